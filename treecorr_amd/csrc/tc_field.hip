// treecorr_amd: device construction of TreeCorr's mean-split ball tree.
//
// Reproduces Field<D,C>::Field + BuildCells (src/Field.cpp:163-277) and
// BuildCell/SplitData/BuildCellData/finishAverages (src/Cell.cpp:89-615) level by
// level on the GPU:
//   * every segment (a node's object range) is owned by one wavefront, which
//     computes the weighted centroid and payload sums in exactly the sequential
//     order of the reference (lanes load 64 objects coalesced, the running sums
//     walk the lanes in order), then size^2 and the bounding box (max/min are
//     order free, so lanes reduce them in parallel);
//   * the split is libstdc++'s std::partition order, computed with one global
//     scan: the k-th misplaced object from the left swaps with the k-th
//     misplaced object from the right;
//   * the rare degenerate split (and split_method='median') runs a serial
//     restatement of std::nth_element (tc_select.h) on one thread;
//   * objects are physically permuted every level so segment loops stay
//     contiguous and coalesced.
#include "tc_engine.h"
#include "tc_select.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

namespace tc {

enum SegMode { SEG_TOP = 0, SEG_BUILD = 1 };

struct Seg {
    long long start, end;
    int node, mode, mintop, maxtop;
    unsigned long long key;
    int depth, root;     // root: which tree of a multi-root build the segment belongs to
};

struct SegOut {
    long long mid;
    double value;
    int split, dim, fallback, childL;
    int child_mode, node;     // node: the node this segment became (-1 if none)
};

// Speculative tree levels.  SPEC_FAST takes a segment's centroid from a parallel
// (unordered) sum and drives the partition with it; SPEC_CHECK recomputes the
// centroid in the reference's order from a snapshot of the same level, writes the
// node records from it and flags any decision or partition side the approximate
// centroid got wrong (the whole build is then redone without speculation).
enum { SPEC_OFF = 0, SPEC_FAST = 1, SPEC_CHECK = 2 };
constexpr int SPEC_NT = 256;            // threads of the slice kernels
constexpr int SPEC_BLOCKS = 1024;       // slices per speculative level, at most
constexpr int SPEC_LEVELS_MAX = 8;      // speculative levels per build, at most
constexpr long long SPEC_MIN_N = 1 << 14;   // objects per segment worth speculating on
constexpr size_t SPEC_SNAP_MAX = size_t(16) << 30;   // bytes of snapshots per build, at most
constexpr int SPEC_P1 = 9;              // per-slice sums: 5 centroid sums, 4 sums of |terms|

struct SpecArgs {
    const double* c1;   // SPEC_FAST: approximate centroids, 4 per segment
    const double* p2;   // SPEC_FAST: per-slice size and bounds, 7 per slice
    int k;              // SPEC_FAST: slices per segment
    int* bad;           // SPEC_CHECK: set when a decision differs
};

struct TopEntry { unsigned long long key; int node, root; };

struct Elems {
    double *x, *y, *z, *wp;
    float *w, *a, *b;
    int* seg;
};

// A speculative level's snapshot for its exact pass: the segments, the decisions
// taken from the approximate centroids and the objects in that level's order.
struct SpecLevel {
    const Seg* segs;
    const SegOut* sout;
    Elems E;
    long long nseg;
    int level0;
};
struct SpecBatch {
    SpecLevel lv[SPEC_LEVELS_MAX];
    int nlev;
};

struct BuildParams {
    double minsizesq, maxsizesq;
    int sm, brute;
};

// ---------------------------------------------------------------------------
// Object setup: the leaf CellData constructors (include/Cell.h:79-81, 131-133,
// 167-169) and CellDataHelper (src/Field.cpp:83-144).
// ---------------------------------------------------------------------------
template <int D, int C>
__global__ void k_init_elems(Elems E, const double* x, const double* y, const double* z,
                             const double* d1, const double* d2, const double* w,
                             const double* wpos, long long n, const long long* bounds,
                             int nroot) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = x[i], py = y[i], pz = (C == Flat) ? 0. : z[i];
    if (C == Sphere) normalize_recip(px, py, pz);
    double wi = w[i];
    E.x[i] = px; E.y[i] = py; E.z[i] = pz;
    E.w[i] = float(wi);
    E.wp[i] = wpos ? wpos[i] : wi;
    float a = 0.f, b = 0.f;
    if (D == KData) a = float(wi * d1[i]);
    if (is_complex(D)) { a = float(wi * d1[i]); b = float(wi * d2[i]); }
    E.a[i] = a; E.b[i] = b;
    // the root segment holding object i: bounds[r] <= i < bounds[r+1]
    int lo = 0, hi = nroot - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bounds[mid] <= i) lo = mid; else hi = mid - 1;
    }
    E.seg[i] = lo;
}

// ---------------------------------------------------------------------------
// Per-segment work: one workgroup of NT threads per segment.
// ---------------------------------------------------------------------------
// Reference-order sums.  Up to NC per-object values are summed over [start, end)
// strictly in object order, one running sum per value, as BuildCellData and
// finishAverages do (src/Cell.cpp:170-354).  The values of a batch of objects are
// computed in parallel and staged in LDS as NC rows; lane c of wave 0 then adds row c
// in order, so all NC sums advance with one dependent FP64 add per object.  With
// NT > 64 the other waves stage batch b+1 while wave 0 adds batch b; the fewer the
// segments (the longer the chains), the more staging waves (NT = 512 near the root).
constexpr long long SEG_WIDE_MAX = 2048;   // levels with at most this many segments use NT=256
constexpr long long SEG_TOP_MAX = 64;      // ... and with at most this many, NT=512

template <int NT> struct SegCfg {
    static constexpr int B = NT > 64 ? 1024 : 128;   // objects per batch
    static constexpr int NC = 5;                      // values per object, at most
    static constexpr int U = NT > 64 ? 32 : 8;        // chain read-ahead (power of 2)
    // LDS batch buffers: with NBUF buffers the stagers fill batch b+NBUF-1 while the
    // adding wave reads batch b (three buffers measured no faster than two)
    static constexpr int NBUF = NT > 64 ? 2 : 1;
    static constexpr int RS = B;   // row stride in LDS (padding measured no faster)
};

template <int NT, int NC, class F>
__device__ __forceinline__ void ordered_sums(long long start, long long end, double init,
                                             double* stg, double* bcast, F value,
                                             double (&out)[NC]) {
    constexpr int B = SegCfg<NT>::B, RS = SegCfg<NT>::RS;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const long long n = end - start;
    const int nbat = (int)((n + B - 1) / B);
    // stagers: waves 1.. when there are several, else the single wave
    const bool stager = NT > 64 ? wave > 0 : true;
    const int st0 = NT > 64 ? t - 64 : t;
    const int nst = NT > 64 ? NT - 64 : NT;
    // Each stager handles PER objects of a batch; all their loads are issued before
    // any value is written to LDS, so one memory latency covers the whole batch.
    constexpr int NST = NT > 64 ? NT - 64 : NT;
    constexpr int PER = (B + NST - 1) / NST;
    auto stage = [&](int bat, double* buf) {
        const long long b0 = start + (long long)bat * B;
        double v[PER][NC];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = st0 + q * nst;
            if (j < B && b0 + j < end) value(b0 + j, v[q]);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = st0 + q * nst;
            if (j < B && b0 + j < end) {
#pragma unroll
                for (int c = 0; c < NC; ++c) buf[c * RS + j] = v[q][c];
            }
        }
    };
    constexpr int NBUF = SegCfg<NT>::NBUF;
    double acc = init;
    if (stager) {
        stage(0, stg);
        for (int b = 1; b < NBUF - 1 && b < nbat; ++b) stage(b, stg + b * NC * RS);
    }
    __syncthreads();
    for (int bat = 0; bat < nbat; ++bat) {
        double* cur = stg + (bat % NBUF) * NC * RS;
        // buffer (bat+NBUF-1) % NBUF was read in the previous batch, before the barrier
        if (NT > 64 && stager && bat + NBUF - 1 < nbat)
            stage(bat + NBUF - 1, stg + ((bat + NBUF - 1) % NBUF) * NC * RS);
        if (wave == 0 && lane < NC) {
            const int cnt = (int)(end - (start + (long long)bat * B) < B
                                      ? end - (start + (long long)bat * B) : B);
            const double* row = cur + lane * RS;
            // Software-pipelined: the next U values are read while the current U are
            // added, so an LDS read has U dependent additions of time to arrive.  A row
            // always has B entries, so reading ahead of cnt stays inside it (those
            // values are never added).
            constexpr int U = SegCfg<NT>::U;
            const int cntu = cnt & ~(U - 1);
            double ra[U], rb[U];
#pragma unroll
            for (int q = 0; q < U; ++q) ra[q] = row[q];
            int j = 0;
            // The scheduling barriers keep each group's reads issued before the
            // additions of the previous group (left alone, the compiler sinks the reads
            // next to their use and every group waits for a full LDS latency).
            while (j < cntu) {
#pragma unroll
                for (int q = 0; q < U; ++q) rb[q] = row[(j + U + q) & (B - 1)];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < U; ++q) acc += ra[q];
                __builtin_amdgcn_sched_barrier(0);
                j += U;
                if (j >= cntu) break;
#pragma unroll
                for (int q = 0; q < U; ++q) ra[q] = row[(j + U + q) & (B - 1)];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < U; ++q) acc += rb[q];
                __builtin_amdgcn_sched_barrier(0);
                j += U;
            }
            for (; j < cnt; ++j) acc += row[j];
        }
        __syncthreads();
        if (NT == 64 && bat + 1 < nbat) {
            stage(bat + 1, stg);
            __syncthreads();
        }
    }
    if (wave == 0 && lane < NC) bcast[lane] = acc;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = bcast[c];
    __syncthreads();
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
    v = wave_max(v);
    if (NT == 64) return v;
    const int t = threadIdx.x;
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int k = 1; k < NT / 64; ++k) m = fmax(m, red[k]);
    __syncthreads();
    return m;
}
// Sum over the workgroup in no particular order (speculative centroids only).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (NT == 64) return v;
    const int t = threadIdx.x;
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int k = 1; k < NT / 64; ++k) m += red[k];
    __syncthreads();
    return m;
}
template <int NT>
__device__ __forceinline__ double block_min(double v, double* red) {
    v = wave_min(v);
    if (NT == 64) return v;
    const int t = threadIdx.x;
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    double m = red[0];
    for (int k = 1; k < NT / 64; ++k) m = fmin(m, red[k]);
    __syncthreads();
    return m;
}

// The centroid of summed (x w, y w, z w, w, w_float) values (BuildCellData,
// src/Cell.cpp:170-200); cx, cy, cz hold the first object's position on entry and
// keep it when the weights sum to zero.
template <int C>
__device__ __forceinline__ void centroid_from_sums(const double* sums, double& cx, double& cy,
                                                   double& cz, float& wf) {
    wf = float(sums[4]);
    if (sums[3] != 0.) {
        cx = sums[0] / sums[3]; cy = sums[1] / sums[3];
        cz = (C == Flat) ? 0. : sums[2] / sums[3];
        if (C == Sphere) normalize_div(cx, cy, cz);
    }
}

// One segment's workgroup: its centroid, size and bounds, the top-cell and split
// decisions, its node record and the split plane.
template <int D, int C, int NT, int SPEC>
__device__ __forceinline__ void segment_body(const long long s, const Seg* segs, SegOut* sout,
                                             const Elems& E, DNode* nodes, int* counters,
                                             TopEntry* tops, const BuildParams& P,
                                             FieldInfo* finfo, int level0, long long skip_n,
                                             const SpecArgs& sp) {
    constexpr int NC = SegCfg<NT>::NC;
    __shared__ double stg[SegCfg<NT>::NBUF * NC * SegCfg<NT>::RS];
    __shared__ double bcast[NC];
    __shared__ double red[NT / 64];
    __shared__ int s_int[2];
    const int tid = threadIdx.x;
    const Seg sg = segs[s];
    const long long start = sg.start, end = sg.end, n = end - start;
    if (n <= skip_n) return;   // handled by k_segment_small

    double cx = E.x[start], cy = E.y[start], cz = E.z[start];
    float wf = E.w[start];
    double sizesq = 0.;
    double mnx = 0, mxx = 0, mny = 0, mxy = 0, mnz = 0, mxz = 0;

    if (n > 1 || level0) {
        double smax = 0.;
        double lnx = INFINITY, lxx = -INFINITY, lny = INFINITY, lxy = -INFINITY;
        double lnz = INFINITY, lxz = -INFINITY;
        if (SPEC == SPEC_FAST) {
            // the approximate centroid and the partial sizes and bounds of the slices
            const double* c1 = sp.c1 + 4 * s;
            cx = c1[0]; cy = c1[1]; cz = c1[2];
            for (int k = tid; k < sp.k; k += NT) {
                const double* q = sp.p2 + ((size_t)s * sp.k + k) * 7;
                smax = fmax(smax, q[0]);
                lnx = fmin(lnx, q[1]); lxx = fmax(lxx, q[2]);
                lny = fmin(lny, q[3]); lxy = fmax(lxy, q[4]);
                lnz = fmin(lnz, q[5]); lxz = fmax(lxz, q[6]);
            }
        } else {
            // BuildCellData (src/Cell.cpp:170-200): sums start from the first object,
            // which -0.0 reproduces exactly (-0.0 + v == v for every v).
            double sums[NC];
            ordered_sums<NT, NC>(start, end, -0.0, stg, bcast,
                [&](long long i, double (&v)[NC]) {
                    const double wp = E.wp[i];
                    v[0] = E.x[i] * wp; v[1] = E.y[i] * wp;
                    v[2] = C == Flat ? 0. : E.z[i] * wp;
                    v[3] = wp; v[4] = double(E.w[i]);
                }, sums);
            centroid_from_sums<C>(sums, cx, cy, cz, wf);
            // CalculateSizeSq (src/Cell.cpp:89-100) and Bounds (include/Bounds.h).  Maxima
            // and minima do not depend on the order, so each thread takes UB objects per
            // round, all their loads in flight together (one workgroup reads the whole
            // segment: near the root, millions of objects).
            auto take = [&](double px, double py, double pz) {
                const double dx = cx - px, dy = cy - py;
                double dsq;
                if (C == Flat) dsq = dx * dx + dy * dy;
                else { const double dz = cz - pz; dsq = dx * dx + dy * dy + dz * dz; }
                smax = fmax(smax, dsq);
                lnx = fmin(lnx, px); lxx = fmax(lxx, px);
                lny = fmin(lny, py); lxy = fmax(lxy, py);
                lnz = fmin(lnz, pz); lxz = fmax(lxz, pz);
            };
            constexpr int UB = 4;
            long long i = start + tid;
            for (; i + (UB - 1) * NT < end; i += UB * NT) {
                double px[UB], py[UB], pz[UB];
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    px[u] = E.x[i + u * NT]; py[u] = E.y[i + u * NT]; pz[u] = E.z[i + u * NT];
                }
#pragma unroll
                for (int u = 0; u < UB; ++u) take(px[u], py[u], pz[u]);
            }
            for (; i < end; i += NT) take(E.x[i], E.y[i], E.z[i]);
        }
        sizesq = block_max<NT>(smax, red);
        mnx = block_min<NT>(lnx, red); mxx = block_max<NT>(lxx, red);
        mny = block_min<NT>(lny, red); mxy = block_max<NT>(lxy, red);
        mnz = block_min<NT>(lnz, red); mxz = block_max<NT>(lxz, red);
        if (level0 && tid == 0 && SPEC != SPEC_FAST) {
            // the roots are the level-0 segments, in order
            finfo[s].cx = cx; finfo[s].cy = cy; finfo[s].cz = cz; finfo[s].sizesq = sizesq;
        }
    }

    // SetupTopLevelCells (src/Field.cpp:27-81) and BuildCell (src/Cell.cpp:556-615).
    int mode = sg.mode;
    int node = sg.node;
    bool split = false;
    const SegOut fast = SPEC == SPEC_CHECK ? sout[s] : SegOut{};
    bool bad = false;     // SPEC_CHECK: a decision differs from the speculative one
    if (mode == SEG_TOP) {
        double ssq = (n == 1) ? 0. : sizesq;
        if (ssq == 0 || (ssq <= P.maxsizesq && sg.mintop <= 0) || sg.maxtop <= 0) {
            if (SPEC == SPEC_CHECK) {
                bad = bad || fast.child_mode != SEG_BUILD;
                node = fast.node;
            } else if (tid == 0) {
                node = atomicAdd(&counters[0], 1);
                int t = atomicAdd(&counters[1], 1);
                tops[t].key = sg.key;
                tops[t].node = node;
                tops[t].root = sg.root;
                s_int[0] = node;
            }
            if (SPEC != SPEC_CHECK) {
                __syncthreads();
                node = s_int[0];
            }
            mode = SEG_BUILD;
        } else {
            split = true;
            if (SPEC == SPEC_CHECK) bad = bad || fast.child_mode != SEG_TOP;
        }
    }
    int childL = -1;
    if (mode == SEG_BUILD && SPEC == SPEC_CHECK && node < 0) bad = true;
    if (mode == SEG_BUILD && !(SPEC == SPEC_CHECK && bad)) {
        DNode nd;
        if (n == 1) {
            long long i = start;
            nd.x = E.x[i]; nd.y = E.y[i]; nd.z = E.z[i];
            nd.w = E.w[i]; nd.a = E.a[i]; nd.b = E.b[i];
            nd.size = 0.f; nd.child = -1; nd.n = 1;
        } else {
            // finishAverages (src/Cell.cpp:208-354); a speculative pass leaves the node
            // record to SPEC_CHECK
            float fa = 0.f, fb = 0.f;
            if (SPEC == SPEC_FAST) {
            } else if (D == KData) {
                double sk[NC];
                ordered_sums<NT, NC>(start, end, 0., stg, bcast,
                    [&](long long i, double (&v)[NC]) {
                        v[0] = double(E.a[i]); v[1] = v[2] = v[3] = v[4] = 0.;
                    }, sk);
                fa = float(sk[0]);
            } else if (is_complex(D)) {
                double ccx = cx, ccy = cy, ccz = cz;
                if (C == ThreeD) normalize_recip(ccx, ccy, ccz);
                double sg2[NC];
                ordered_sums<NT, NC>(start, end, 0., stg, bcast,
                    [&](long long i, double (&v)[NC]) {
                        const double gr = double(E.a[i]), gi = double(E.b[i]);
                        double vr, vi;
                        // spin 0 (Z) is never rotated: SimpleSum / ParallelTransportSum
                        // with exp(-i 0 beta) = 1 (src/Cell.cpp:221-321)
                        if (C == Flat || D == ZData) {
                            vr = gr; vi = gi;
                        } else {
                            // ParallelTransportSum (src/Cell.cpp:276-321), in the
                            // evaluation order of the compiled reference.
                            double px = E.x[i], py = E.y[i], pz = E.z[i];
                            if (C == ThreeD) normalize_recip(px, py, pz);
                            const double z1 = cz, z2 = pz;
                            const double dx = ccx - px, dy = ccy - py, dz = ccz - pz;
                            const double dsq = (C == Sphere) ? (dx * dx + dy * dy) + dz * dz
                                                             : (dy * dy + dz * dz) + dx * dx;
                            if (dsq < 1.e-12) {
                                vr = gr; vi = gi;
                            } else {
                                const double cosA = (z1 - z2) + 0.5 * z2 * dsq;
                                const double sinA = ccy * px - ccx * py;
                                const double cosB = (z2 - z1) + 0.5 * z1 * dsq;
                                const double cosbeta = -cosA * cosB + sinA * sinA;
                                const double sinbeta = (C == Sphere)
                                    ? (cosA + cosB) * sinA
                                    : ((z1 - z2) + (0.5 * z2 * dsq + cosB)) * sinA;
                                double er, ei;
                                expmsialpha<spin_of(D)>(cosbeta, sinbeta, er, ei);
                                vr = gr * er - gi * ei;
                                vi = gr * ei + gi * er;
                            }
                        }
                        v[0] = vr; v[1] = vi; v[2] = v[3] = v[4] = 0.;
                    }, sg2);
                fa = float(sg2[0]); fb = float(sg2[1]);
            }
            nd.x = cx; nd.y = cy; nd.z = cz;
            nd.w = wf; nd.a = fa; nd.b = fb; nd.n = (int)n;
            if (sizesq > P.minsizesq) {
                nd.size = P.brute ? INFINITY : float(sqrt(sizesq));
                split = true;
                if (SPEC == SPEC_CHECK) {
                    bad = bad || !fast.split || fast.childL < 0;
                    childL = fast.childL;
                } else {
                    if (tid == 0) s_int[1] = atomicAdd(&counters[0], 2);
                    __syncthreads();
                    childL = s_int[1];
                }
                nd.child = childL;
            } else {
                nd.size = 0.f;
                nd.child = -1;
                if (SPEC == SPEC_CHECK) bad = bad || fast.split;
            }
        }
        if (tid == 0 && SPEC != SPEC_FAST && !bad) nodes[node] = nd;
    }

    if (SPEC == SPEC_CHECK) {
        // the partition side of every object under the exact split value must be the
        // side the speculative value gave it
        if (!bad && split) {
            int dim;
            if (C == Flat) dim = (mxy - mny) > (mxx - mnx) ? 1 : 0;
            else {
                double xr = mxx - mnx, yr = mxy - mny, zr = mxz - mnz;
                dim = yr > xr ? (zr > yr ? 2 : 1) : (zr > xr ? 2 : 0);
            }
            bad = dim != fast.dim;
            if (!bad && P.sm == Mean) {
                const double v = dim == 2 ? cz : dim == 1 ? cy : cx;
                const double* cc = dim == 2 ? E.z : dim == 1 ? E.y : E.x;
                int diff = 0;
                if (v != fast.value) {
                    constexpr int UB = 8;
                    long long i = start + tid;
                    for (; i + (UB - 1) * NT < end; i += UB * NT) {
                        double c[UB];
#pragma unroll
                        for (int u = 0; u < UB; ++u) c[u] = cc[i + u * NT];
#pragma unroll
                        for (int u = 0; u < UB; ++u) diff |= (c[u] < v) != (c[u] < fast.value);
                    }
                    for (; i < end; i += NT) diff |= (cc[i] < v) != (cc[i] < fast.value);
                }
                bad = block_max<NT>(double(diff), red) != 0.;
            } else if (!bad && P.sm == Middle) {
                const double v = dim == 2 ? (mxz + mnz) / 2. : dim == 1 ? (mxy + mny) / 2.
                                                                       : (mxx + mnx) / 2.;
                bad = v != fast.value;
            }
        }
        if (bad && tid == 0) atomicOr(sp.bad, 1);
        return;
    }

    if (tid == 0) {
        SegOut o;
        o.split = split ? 1 : 0;
        o.childL = childL;
        o.child_mode = mode;
        o.node = mode == SEG_BUILD ? node : -1;
        o.mid = start;
        o.fallback = 0;
        o.dim = 0;
        o.value = 0.;
        if (split) {
            // Bounds<C>::getSplit (include/Bounds.h:65-66, 133-141)
            int dim;
            if (C == Flat) dim = (mxy - mny) > (mxx - mnx) ? 1 : 0;
            else {
                double xr = mxx - mnx, yr = mxy - mny, zr = mxz - mnz;
                dim = yr > xr ? (zr > yr ? 2 : 1) : (zr > xr ? 2 : 0);
            }
            o.dim = dim;
            if (P.sm == Mean) o.value = dim == 2 ? cz : dim == 1 ? cy : cx;
            else if (P.sm == Middle)
                o.value = dim == 2 ? (mxz + mnz) / 2. : dim == 1 ? (mxy + mny) / 2. : (mxx + mnx) / 2.;
            else o.fallback = 1;   // Median
        }
        sout[s] = o;
    }
}

template <int D, int C, int NT, int SPEC = SPEC_OFF>
__global__ __launch_bounds__(NT) void k_segment(const Seg* segs, long long seg0, SegOut* sout,
                                                Elems E, DNode* nodes, int* counters,
                                                TopEntry* tops, BuildParams P, FieldInfo* finfo,
                                                int level0, long long skip_n, SpecArgs sp) {
    segment_body<D, C, NT, SPEC>(seg0 + blockIdx.x, segs, sout, E, nodes, counters, tops, P,
                                 finfo, level0, skip_n, sp);
}

// The exact pass of every speculative level in one launch, the root's segment first
// (the longest chain): workgroup b is segment b of level 0, then level 1's, and so on.
constexpr int CHECK_NT = 512;   // 1024 measured no faster
template <int D, int C>
__global__ __launch_bounds__(CHECK_NT) void k_segment_check(SpecBatch B, DNode* nodes,
                                                            BuildParams P, FieldInfo* finfo,
                                                            int* bad) {
    long long s = blockIdx.x;
    int l = 0;
    while (l + 1 < B.nlev && s >= B.lv[l].nseg) { s -= B.lv[l].nseg; ++l; }
    const SpecLevel& L = B.lv[l];
    SpecArgs sp{};
    sp.bad = bad;
    segment_body<D, C, CHECK_NT, SPEC_CHECK>(s, L.segs, const_cast<SegOut*>(L.sout), L.E, nodes,
                                             nullptr, nullptr, P, finfo, L.level0, 0, sp);
}

// Speculative centroids, sizes and bounds: each segment of the level is cut into K
// slices, one workgroup each (a level near the root holds few segments, too few
// workgroups to read its objects at the memory's rate).
__device__ __forceinline__ void spec_slice(const Seg& sg, int k, int K, long long& a,
                                           long long& b) {
    const long long len = sg.end - sg.start;
    a = sg.start + len * k / K;
    b = sg.start + len * (k + 1) / K;
}

template <int C>
__global__ __launch_bounds__(SPEC_NT) void k_spec_sums(const Seg* segs, Elems E, int K,
                                                       double* p1) {
    __shared__ double red[SPEC_NT / 64];
    const int k = blockIdx.x;
    const long long s = blockIdx.y;
    long long a, b;
    spec_slice(segs[s], k, K, a, b);
    // the centroid sums and, for the bound on their rounding, the sums of |terms|
    double part[SPEC_P1] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
    for (long long i = a + threadIdx.x; i < b; i += SPEC_NT) {
        const double wp = E.wp[i];
        const double tx = E.x[i] * wp, ty = E.y[i] * wp, tz = C == Flat ? 0. : E.z[i] * wp;
        part[0] += tx; part[1] += ty; part[2] += tz;
        part[3] += wp; part[4] += double(E.w[i]);
        part[5] += fabs(tx); part[6] += fabs(ty); part[7] += fabs(tz); part[8] += fabs(wp);
    }
#pragma unroll
    for (int c = 0; c < SPEC_P1; ++c) {
        const double v = block_sum<SPEC_NT>(part[c], red);
        if (threadIdx.x == 0) p1[((size_t)s * K + k) * SPEC_P1 + c] = v;
    }
}

// hb: for each segment and coordinate, a bound on |centroid in the reference's order -
// this centroid|.  Any order of summing n terms is within gamma_n sum |t| of the exact sum
// (gamma_n = n u / (1 - n u), u = 2^-53), so the two orders' sums are within twice that;
// the bound follows the division by the weight sum (and on the sphere the normalisation),
// doubled, plus a few units of rounding.
template <int C>
__global__ __launch_bounds__(SPEC_NT) void k_spec_centroid(const Seg* segs, Elems E, int K,
                                                           const double* p1, double* c1,
                                                           double* hb) {
    __shared__ double red[SPEC_NT / 64];
    const long long s = blockIdx.x;
    double part[SPEC_P1] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
    for (int k = threadIdx.x; k < K; k += SPEC_NT) {
#pragma unroll
        for (int c = 0; c < SPEC_P1; ++c) part[c] += p1[((size_t)s * K + k) * SPEC_P1 + c];
    }
    double sums[SPEC_P1];
#pragma unroll
    for (int c = 0; c < SPEC_P1; ++c) sums[c] = block_sum<SPEC_NT>(part[c], red);
    if (threadIdx.x == 0) {
        const Seg sg = segs[s];
        const long long i = sg.start;
        double cx = E.x[i], cy = E.y[i], cz = E.z[i];
        float wf;
        centroid_from_sums<C>(sums, cx, cy, cz, wf);
        c1[4 * s] = cx; c1[4 * s + 1] = cy; c1[4 * s + 2] = cz; c1[4 * s + 3] = double(wf);
        const double u = 0x1p-53;
        const double nu = double(sg.end - sg.start) * u;
        const double g2 = nu < 0.5 ? 2. * nu / (1. - nu) : INFINITY;
        const double ew = g2 * sums[8];
        const double aw = fabs(sums[3]) - ew;
        double e[3], c[3] = {cx, cy, cz};
        double esum = 0.;
        for (int q = 0; q < 3; ++q) {
            const double v = sums[q] / sums[3];
            e[q] = aw > 0. ? (g2 * sums[5 + q] + fabs(v) * ew) / aw + 4. * u * fabs(v)
                           : INFINITY;
            esum += e[q];
        }
        for (int q = 0; q < 3; ++q) {
            double h = e[q];
            if (C == Sphere) {
                // c = v / |v|: |dc| <= (|dv_q| + |c_q| sum |dv|) / (|v| - sum |dv|)
                const double vn = sqrt((sums[0] * sums[0] + sums[1] * sums[1] +
                                        sums[2] * sums[2])) / fabs(sums[3]);
                h = vn - esum > 0. ? (e[q] + fabs(c[q]) * esum) / (vn - esum) + 8. * u
                                   : INFINITY;
            }
            hb[3 * s + q] = (C == Flat && q == 2) ? 0. : 2. * h + 1e-300;
        }
    }
}

// Validation of a speculative level's decisions without the reference-order sums: a
// segment's partition (a Mean split) is the reference's if no object lies within hb of
// the speculative split value; its top-cell decision (size 0 or not) is if the
// speculative size exceeds the centroid's bound with room to spare.  Any doubt sets
// *flag and the build waits for the exact pass.
__global__ void k_spec_validate_segs(const Seg* segs, long long nseg, int K, const double* p2,
                                     const double* hb, int* flag) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    if (segs[s].end - segs[s].start <= 1) return;
    double ssq = 0.;
    for (int k = 0; k < K; ++k) ssq = fmax(ssq, p2[((size_t)s * K + k) * 7]);
    const double h = hb[3 * s] + hb[3 * s + 1] + hb[3 * s + 2];
    if (!(sqrt(ssq) > 2. * h + 1e-290)) atomicOr(flag, 2);
}

template <int C>
__global__ __launch_bounds__(SPEC_NT) void k_spec_bounds(const Seg* segs, Elems E, int K,
                                                         const double* c1, double* p2) {
    __shared__ double red[SPEC_NT / 64];
    const int k = blockIdx.x;
    const long long s = blockIdx.y;
    long long a, b;
    spec_slice(segs[s], k, K, a, b);
    const double cx = c1[4 * s], cy = c1[4 * s + 1], cz = c1[4 * s + 2];
    double smax = 0.;
    double lnx = INFINITY, lxx = -INFINITY, lny = INFINITY, lxy = -INFINITY;
    double lnz = INFINITY, lxz = -INFINITY;
    for (long long i = a + threadIdx.x; i < b; i += SPEC_NT) {
        const double px = E.x[i], py = E.y[i], pz = E.z[i];
        const double dx = cx - px, dy = cy - py;
        double dsq;
        if (C == Flat) dsq = dx * dx + dy * dy;
        else { const double dz = cz - pz; dsq = dx * dx + dy * dy + dz * dz; }
        smax = fmax(smax, dsq);
        lnx = fmin(lnx, px); lxx = fmax(lxx, px);
        lny = fmin(lny, py); lxy = fmax(lxy, py);
        lnz = fmin(lnz, pz); lxz = fmax(lxz, pz);
    }
    const double v[7] = {block_max<SPEC_NT>(smax, red), block_min<SPEC_NT>(lnx, red),
                         block_max<SPEC_NT>(lxx, red), block_min<SPEC_NT>(lny, red),
                         block_max<SPEC_NT>(lxy, red), block_min<SPEC_NT>(lnz, red),
                         block_max<SPEC_NT>(lxz, red)};
    if (threadIdx.x == 0)
        for (int c = 0; c < 7; ++c) p2[((size_t)s * K + k) * 7 + c] = v[c];
}

// Segments of at most SEG_SMALL_N objects, one thread each (the deep levels hold up to
// n/2 such segments).  The same sums in the same order as k_segment, computed by the
// thread itself; k_segment skips these segments.
constexpr long long SEG_SMALL_N = 32;

template <int D, int C>
__global__ __launch_bounds__(256) void k_segment_small(const Seg* segs, long long nseg,
                                                       SegOut* sout, Elems E, DNode* nodes,
                                                       int* counters, TopEntry* tops,
                                                       BuildParams P, FieldInfo* finfo,
                                                       int level0) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const Seg sg = segs[s];
    const long long start = sg.start, end = sg.end, n = end - start;
    if (n > SEG_SMALL_N) return;
    double cx = E.x[start], cy = E.y[start], cz = E.z[start];
    float wf = E.w[start];
    double sizesq = 0.;
    double mnx = 0, mxx = 0, mny = 0, mxy = 0, mnz = 0, mxz = 0;
    if (n > 1 || level0) {
        // BuildCellData (src/Cell.cpp:170-200)
        double ax = -0.0, ay = -0.0, az = -0.0, asw = -0.0, aww = -0.0;
        for (long long i = start; i < end; ++i) {
            const double wp = E.wp[i];
            ax += E.x[i] * wp; ay += E.y[i] * wp;
            if (C != Flat) az += E.z[i] * wp;
            asw += wp; aww += double(E.w[i]);
        }
        wf = float(aww);
        if (asw != 0.) {
            cx = ax / asw; cy = ay / asw; cz = (C == Flat) ? 0. : az / asw;
            if (C == Sphere) normalize_div(cx, cy, cz);
        }
        // CalculateSizeSq (src/Cell.cpp:89-100) and Bounds (include/Bounds.h)
        mnx = mny = mnz = INFINITY; mxx = mxy = mxz = -INFINITY;
        for (long long i = start; i < end; ++i) {
            const double px = E.x[i], py = E.y[i], pz = E.z[i];
            const double dx = cx - px, dy = cy - py;
            double dsq;
            if (C == Flat) dsq = dx * dx + dy * dy;
            else { const double dz = cz - pz; dsq = dx * dx + dy * dy + dz * dz; }
            sizesq = fmax(sizesq, dsq);
            mnx = fmin(mnx, px); mxx = fmax(mxx, px);
            mny = fmin(mny, py); mxy = fmax(mxy, py);
            mnz = fmin(mnz, pz); mxz = fmax(mxz, pz);
        }
        if (level0) {
            finfo[s].cx = cx; finfo[s].cy = cy; finfo[s].cz = cz; finfo[s].sizesq = sizesq;
        }
    }
    // SetupTopLevelCells (src/Field.cpp:27-81) and BuildCell (src/Cell.cpp:556-615)
    int mode = sg.mode;
    int node = sg.node;
    bool split = false;
    if (mode == SEG_TOP) {
        const double ssq = (n == 1) ? 0. : sizesq;
        if (ssq == 0 || (ssq <= P.maxsizesq && sg.mintop <= 0) || sg.maxtop <= 0) {
            node = atomicAdd(&counters[0], 1);
            const int t = atomicAdd(&counters[1], 1);
            tops[t].key = sg.key;
            tops[t].node = node;
            tops[t].root = sg.root;
            mode = SEG_BUILD;
        } else {
            split = true;
        }
    }
    int childL = -1;
    if (mode == SEG_BUILD) {
        DNode nd;
        if (n == 1) {
            nd.x = E.x[start]; nd.y = E.y[start]; nd.z = E.z[start];
            nd.w = E.w[start]; nd.a = E.a[start]; nd.b = E.b[start];
            nd.size = 0.f; nd.child = -1; nd.n = 1;
        } else {
            // finishAverages (src/Cell.cpp:208-354)
            float fa = 0.f, fb = 0.f;
            if (D == KData) {
                double sa = 0.;
                for (long long i = start; i < end; ++i) sa += double(E.a[i]);
                fa = float(sa);
            } else if (is_complex(D)) {
                double sr = 0., si = 0.;
                double ccx = cx, ccy = cy, ccz = cz;
                if (C == ThreeD) normalize_recip(ccx, ccy, ccz);
                for (long long i = start; i < end; ++i) {
                    const double gr = double(E.a[i]), gi = double(E.b[i]);
                    double vr, vi;
                    if (C == Flat || D == ZData) {
                        vr = gr; vi = gi;
                    } else {
                        // ParallelTransportSum (src/Cell.cpp:276-321), in the evaluation
                        // order of the compiled reference
                        double px = E.x[i], py = E.y[i], pz = E.z[i];
                        if (C == ThreeD) normalize_recip(px, py, pz);
                        const double z1 = cz, z2 = pz;
                        const double dx = ccx - px, dy = ccy - py, dz = ccz - pz;
                        const double dsq = (C == Sphere) ? (dx * dx + dy * dy) + dz * dz
                                                         : (dy * dy + dz * dz) + dx * dx;
                        if (dsq < 1.e-12) {
                            vr = gr; vi = gi;
                        } else {
                            const double cosA = (z1 - z2) + 0.5 * z2 * dsq;
                            const double sinA = ccy * px - ccx * py;
                            const double cosB = (z2 - z1) + 0.5 * z1 * dsq;
                            const double cosbeta = -cosA * cosB + sinA * sinA;
                            const double sinbeta = (C == Sphere)
                                ? (cosA + cosB) * sinA
                                : ((z1 - z2) + (0.5 * z2 * dsq + cosB)) * sinA;
                            double er, ei;
                            expmsialpha<spin_of(D)>(cosbeta, sinbeta, er, ei);
                            vr = gr * er - gi * ei;
                            vi = gr * ei + gi * er;
                        }
                    }
                    sr += vr; si += vi;
                }
                fa = float(sr); fb = float(si);
            }
            nd.x = cx; nd.y = cy; nd.z = cz;
            nd.w = wf; nd.a = fa; nd.b = fb; nd.n = (int)n;
            if (sizesq > P.minsizesq) {
                nd.size = P.brute ? INFINITY : float(sqrt(sizesq));
                split = true;
                childL = atomicAdd(&counters[0], 2);
                nd.child = childL;
            } else {
                nd.size = 0.f;
                nd.child = -1;
            }
        }
        nodes[node] = nd;
    }
    SegOut o;
    o.split = split ? 1 : 0;
    o.childL = childL;
    o.child_mode = mode;
    o.mid = start;
    o.fallback = 0;
    o.dim = 0;
    o.value = 0.;
    if (split) {
        // Bounds<C>::getSplit (include/Bounds.h:65-66, 133-141)
        int dim;
        if (C == Flat) dim = (mxy - mny) > (mxx - mnx) ? 1 : 0;
        else {
            const double xr = mxx - mnx, yr = mxy - mny, zr = mxz - mnz;
            dim = yr > xr ? (zr > yr ? 2 : 1) : (zr > xr ? 2 : 0);
        }
        o.dim = dim;
        if (P.sm == Mean) o.value = dim == 2 ? cz : dim == 1 ? cy : cx;
        else if (P.sm == Middle)
            o.value = dim == 2 ? (mxz + mnz) / 2. : dim == 1 ? (mxy + mny) / 2. : (mxx + mnx) / 2.;
        else o.fallback = 1;   // Median
    }
    sout[s] = o;
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of int flags: out[0..n] with out[n] = total.
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_B = SCAN_T * SCAN_I;

__device__ __forceinline__ int block_exclusive_scan(int v, int* sh, int& total) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int wbase = 0;
    for (int k = 0; k < wv; ++k) wbase += sh[k];
    total = 0;
    for (int k = 0; k < SCAN_T / 64; ++k) total += sh[k];
    __syncthreads();
    return wbase + x - v;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const int* in, long long n, int* bsum) {
    __shared__ int sh[SCAN_T / 64];
    long long base = (long long)blockIdx.x * SCAN_B + threadIdx.x * SCAN_I;
    int s = 0;
    for (int k = 0; k < SCAN_I; ++k) if (base + k < n) s += in[base + k];
    int total;
    block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_bsums(int* bsum, int nb, int* out_total) {
    __shared__ int sh[SCAN_T / 64];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += SCAN_T) {
        int i = c0 + threadIdx.x;
        int v = i < nb ? bsum[i] : 0;
        int total;
        int ex = block_exclusive_scan(v, sh, total);
        if (i < nb) bsum[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const int* in, long long n, const int* bsum,
                                                      int* out) {
    __shared__ int sh[SCAN_T / 64];
    long long base = (long long)blockIdx.x * SCAN_B + threadIdx.x * SCAN_I;
    int v[SCAN_I];
    int s = 0;
    for (int k = 0; k < SCAN_I; ++k) { v[k] = (base + k < n) ? in[base + k] : 0; s += v[k]; }
    int total;
    int ex = block_exclusive_scan(s, sh, total) + bsum[blockIdx.x];
    for (int k = 0; k < SCAN_I; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

static int exclusive_scan(const int* in, long long n, int* out, int* tmp) {
    hipStream_t st = stream();
    int nb = (int)((n + SCAN_B - 1) / SCAN_B);
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_T), 0, st, in, n, tmp);
    hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(SCAN_T), 0, st, tmp, nb, out + n);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_T), 0, st, in, n, tmp, out);
    return 0;
}

// ---------------------------------------------------------------------------
// Partition kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double elem_coord(const Elems& E, long long i, int dim) {
    return dim == 2 ? E.z[i] : dim == 1 ? E.y[i] : E.x[i];
}

__global__ void k_seg_split_flags(const SegOut* sout, int nseg, int* flags) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nseg) flags[s] = sout[s].split;
}

__global__ void k_part_flags(Elems E, const SegOut* sout, long long n, int* flags) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = E.seg[i];
    int f = 0;
    if (s >= 0) {
        const SegOut& o = sout[s];
        if (o.split && !o.fallback) f = elem_coord(E, i, o.dim) < o.value ? 1 : 0;
    }
    flags[i] = f;
}

__global__ void k_spec_validate_objs(Elems E, const SegOut* sout, long long n,
                                     const double* hb, int* flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = E.seg[i];
    if (s < 0) return;
    const SegOut& o = sout[s];
    if (!o.split) return;
    const double v = elem_coord(E, i, o.dim);
    if (fabs(v - o.value) <= hb[3 * s + o.dim]) atomicOr(flag, 1);
}

__global__ void k_part_mid(const Seg* segs, SegOut* sout, int nseg, const int* pre) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    SegOut o = sout[s];
    if (!o.split) return;
    long long st = segs[s].start, en = segs[s].end;
    if (!o.fallback) {
        long long mid = st + (pre[en] - pre[st]);
        if (mid == st || mid == en) o.fallback = 1;   // SplitData, src/Cell.cpp:533-550
        else o.mid = mid;
    }
    if (o.fallback) o.mid = (st + en) / 2;
    sout[s] = o;
}

// Serial std::nth_element for degenerate splits and split_method='median'.
__global__ void k_part_median(const Seg* segs, const SegOut* sout, int nseg, Elems E,
                              double* keybuf, int* ord, long long* dest) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const SegOut o = sout[s];
    if (!o.split || !o.fallback) return;
    long long st = segs[s].start, en = segs[s].end;
    for (long long i = st; i < en; ++i) { keybuf[i] = elem_coord(E, i, o.dim); ord[i] = (int)i; }
    tc_select::KeyLess cmp{keybuf};
    tc_select::nth_element(ord, st, o.mid, en, cmp);
    for (long long p = st; p < en; ++p) dest[ord[p]] = p;
}

// Rank the misplaced objects of each partitioned segment (std::partition swaps the
// k-th !pred object from the left with the k-th pred object from the right).
__global__ void k_part_rank(Elems E, const Seg* segs, const SegOut* sout, long long n,
                            const int* flags, const int* pre, int* lpos, int* rpos) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = E.seg[i];
    if (s < 0) return;
    const SegOut& o = sout[s];
    if (!o.split || o.fallback) return;
    long long st = segs[s].start, en = segs[s].end, mid = o.mid;
    int f = flags[i];
    if (i < mid && !f) {
        long long k = (i - st) - (pre[i] - pre[st]);
        lpos[st + k] = (int)i;
    } else if (i >= mid && f) {
        long long k = pre[en] - pre[i + 1];
        rpos[st + k] = (int)i;
    }
}

__global__ void k_part_dest(Elems E, const Seg* segs, const SegOut* sout, const int* segrank,
                            long long n, const int* flags, const int* pre, const int* lpos,
                            const int* rpos, long long* dest, int* newseg) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = E.seg[i];
    long long d = i;
    int ns = -1;
    if (s >= 0) {
        const SegOut& o = sout[s];
        if (o.split) {
            long long st = segs[s].start, mid = o.mid;
            if (o.fallback) {
                d = dest[i];
            } else {
                int f = flags[i];
                if (i < mid && !f) {
                    long long k = (i - st) - (pre[i] - pre[st]);
                    d = rpos[st + k];
                } else if (i >= mid && f) {
                    long long k = pre[segs[s].end] - pre[i + 1];
                    d = lpos[st + k];
                }
            }
            ns = 2 * segrank[s] + (d < mid ? 0 : 1);
        }
    }
    dest[i] = d;
    newseg[i] = ns;
}

__global__ void k_permute(Elems A, Elems B, long long n, const long long* dest, const int* newseg) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long d = dest[i];
    B.x[d] = A.x[i]; B.y[d] = A.y[i]; B.z[d] = A.z[i]; B.wp[d] = A.wp[i];
    B.w[d] = A.w[i]; B.a[d] = A.a[i]; B.b[d] = A.b[i];
    B.seg[d] = newseg[i];
}

__global__ void k_children(const Seg* segs, const SegOut* sout, const int* segrank, int nseg,
                           Seg* next) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const SegOut o = sout[s];
    if (!o.split) return;
    const Seg sg = segs[s];
    int r = segrank[s];
    Seg L, R;
    L.start = sg.start; L.end = o.mid;
    R.start = o.mid; R.end = sg.end;
    L.mode = R.mode = o.child_mode;
    L.mintop = R.mintop = sg.mintop - 1;
    L.maxtop = R.maxtop = sg.maxtop - 1;
    L.depth = R.depth = sg.depth + 1;
    L.key = sg.key;
    R.key = sg.key | (1ull << (62 - sg.depth));
    L.root = R.root = sg.root;
    if (o.child_mode == SEG_BUILD) { L.node = o.childL; R.node = o.childL + 1; }
    else { L.node = R.node = -1; }
    next[2 * r] = L;
    next[2 * r + 1] = R;
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { dev_free(p, bytes); }
    template <class T> T* as() { return static_cast<T*>(p); }
};

static int dalloc(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    b.p = dev_alloc(bytes);
    if (!b.p) return fail("out of device memory building the tree");
    b.bytes = bytes;
    return 0;
}

// The exact pass of a build whose speculative decisions were validated: it still runs
// on the check streams when the build returns; the buffers it reads and writes (other
// than the node array) are kept here until it has finished.
struct TreeCheck {
    hipEvent_t done[2] = {nullptr, nullptr};
    std::vector<std::pair<void*, size_t>> bufs;
    FieldInfo* finfo = nullptr;   // device, the root's entry (in bufs)
    int* bad = nullptr;           // device flag of the exact pass (in bufs)
    ~TreeCheck() {
        for (hipEvent_t& e : done) {
            if (e) { (void)hipEventSynchronize(e); (void)hipEventDestroy(e); }
        }
        for (auto& b : bufs) dev_free(b.first, b.second);
    }
};

int settle_tree(DeviceTree& T) {
    if (!T.pending) return 0;
    std::shared_ptr<TreeCheck> pc = std::move(T.pending);
    for (hipEvent_t e : pc->done)
        if (e) TC_HIP(hipEventSynchronize(e));
    int bad = 0;
    TC_HIP(hipMemcpy(&bad, pc->bad, sizeof(int), hipMemcpyDeviceToHost));
    TC_HIP(hipMemcpy(&T.info, pc->finfo, sizeof(FieldInfo), hipMemcpyDeviceToHost));
    if (bad) return fail("tree build: the exact pass disagrees with a validated speculative "
                         "decision");
    return 0;
}

static std::atomic<int> g_tree_spec{1};         // tc_set_tree_speculation
static std::atomic<long long> g_tree_spec_redo{0};   // exact rebuilds after a failed check
static std::atomic<long long> g_tree_spec_defer{0};  // builds that did not wait for the check

int set_tree_speculation(int mode) {
    if (mode == -2) return (int)std::min<long long>(g_tree_spec_defer.load(), 1 << 30);
    if (mode < 0) return (int)std::min<long long>(g_tree_spec_redo.load(), 1 << 30);
    if (mode > 3) return fail("tree speculation mode must be 0 to 3");
    g_tree_spec.store(mode);
    return 0;
}

// Builds one tree per root range [bounds[r], bounds[r+1]) of the columns (each range
// non-empty), all in the same level-synchronous pass: the roots are the level-0
// segments.  The trees share one node array.
template <int D, int C>
static int build_trees_t(const double* x, const double* y, const double* z, const double* d1,
                         const double* d2, const double* w, const double* wpos, long long n,
                         const std::vector<long long>& bounds, double minsize, double maxsize,
                         int sm, int brute, int mintop, int maxtop, MultiTree& M) {
    hipStream_t st = stream();
    const int TPB = 256;
    auto grid = [&](long long m) { return dim3((unsigned)((m + TPB - 1) / TPB)); };
    const int nroot = (int)bounds.size() - 1;
    const long long nb_scan = (n + 1 + SCAN_B - 1) / SCAN_B + 1;
    DevBuf ecols[2][8];
    Elems E[2];
    for (int k = 0; k < 2; ++k) {
        for (int c = 0; c < 4; ++c) if (dalloc(ecols[k][c], n * 8)) return 2;
        for (int c = 4; c < 8; ++c) if (dalloc(ecols[k][c], n * 4)) return 2;
        E[k].x = ecols[k][0].as<double>(); E[k].y = ecols[k][1].as<double>();
        E[k].z = ecols[k][2].as<double>(); E[k].wp = ecols[k][3].as<double>();
        E[k].w = ecols[k][4].as<float>(); E[k].a = ecols[k][5].as<float>();
        E[k].b = ecols[k][6].as<float>(); E[k].seg = ecols[k][7].as<int>();
    }
    DevBuf bflags, bpre, blpos, brpos, bdest, bnewseg, bord, bkey, bscan, bbounds;
    DevBuf bsegs[2], bsout, bsegflags, bsegrank, bcnt, btops, bfinfo;
    // top cells: at most one per object, and at most 2^maxtop per root
    const long long maxtops = std::min<long long>(n, (long long)nroot << std::min(maxtop, 40));
    const long long nsegmax = std::max<long long>(n, nroot);
    if (dalloc(bflags, n * 4) || dalloc(bpre, (n + 1) * 4) || dalloc(blpos, n * 4) ||
        dalloc(brpos, n * 4) || dalloc(bdest, n * 8) || dalloc(bnewseg, n * 4) ||
        dalloc(bord, n * 4) || dalloc(bkey, n * 8) || dalloc(bscan, nb_scan * 4) ||
        dalloc(bsegs[0], nsegmax * sizeof(Seg)) || dalloc(bsegs[1], nsegmax * sizeof(Seg)) ||
        dalloc(bsout, nsegmax * sizeof(SegOut)) || dalloc(bsegflags, nsegmax * 4) ||
        dalloc(bsegrank, (nsegmax + 1) * 4) || dalloc(bcnt, 16) ||
        dalloc(btops, sizeof(TopEntry) * (size_t)(maxtops + 1)) ||
        dalloc(bfinfo, sizeof(FieldInfo) * nroot) || dalloc(bbounds, sizeof(long long) * (nroot + 1)))
        return 2;

    // Nodes: at most 2n-1 per root's objects (each split adds two).
    const long long maxnodes = 2 * n;
    M.nodes_bytes = sizeof(DNode) * maxnodes;
    M.nodes = static_cast<DNode*>(dev_alloc(M.nodes_bytes));
    if (!M.nodes) return fail("out of device memory for the tree nodes");

    // Speculation (Mean and Middle splits): the levels near the root partition with
    // centroids summed in parallel; their exact pass runs on the second stream from
    // snapshots, writes the node records and flags any decision that differs.
    const int spec_mode = g_tree_spec.load();
    const bool spec_on = spec_mode != 0 && (sm == Mean || sm == Middle);
    DevBuf bspec, bsnap[SPEC_LEVELS_MAX];
    if (spec_on && dalloc(bspec, sizeof(double) * (7 * SEG_TOP_MAX + (SPEC_P1 + 7) * SPEC_BLOCKS)))
        return 2;
    // A single tree whose speculative decisions all pass the validation does not wait
    // for their exact pass: that pass only writes records above the top cells (the
    // levels speculated on are above min_top) and the field's centre and size.
    const bool defer_ok = spec_on && nroot == 1 && spec_mode != 2;
    bool deferred = false;
    struct SpecGuard {   // declared after the buffers: the exact passes end before they go
        hipStream_t cs[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~SpecGuard() {
            for (int k = 0; k < 2; ++k) {
                if (cs[k]) (void)hipStreamSynchronize(cs[k]);
                if (ev[k]) (void)hipEventDestroy(ev[k]);
            }
        }
    } guard;
    if (spec_on) {
        // the root level's exact pass on one stream, the other levels' on another
        guard.cs[0] = stream2();
        guard.cs[1] = stream3();
        if (!guard.cs[0] || !guard.cs[1]) return fail("could not create the check streams");
        for (int k = 0; k < 2; ++k)
            TC_HIP(hipEventCreateWithFlags(&guard.ev[k], hipEventDisableTiming));
    }
    struct HostBuf {
        int* p = nullptr;
        ~HostBuf() { if (p) (void)hipHostFree(p); }
    } hcnt_buf;
    TC_HIP(hipHostMalloc((void**)&hcnt_buf.p, 16));
    int* hcnt = hcnt_buf.p;
    const BuildParams P{minsize * minsize, maxsize * maxsize, sm, brute};
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };

    int nlevels = 0;
    auto attempt = [&](bool spec) -> int {
        TC_HIP(hipMemcpyAsync(bbounds.p, bounds.data(), sizeof(long long) * (nroot + 1),
                              hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL((k_init_elems<D, C>), grid(n), dim3(TPB), 0, st, E[0], x, y, z, d1,
                           d2, w, wpos, n, bbounds.as<long long>(), nroot);
        std::vector<Seg> roots(nroot);
        for (int r = 0; r < nroot; ++r) {
            Seg& g = roots[r];
            g = Seg{};
            g.start = bounds[r]; g.end = bounds[r + 1]; g.node = -1; g.mode = SEG_TOP;
            g.mintop = mintop; g.maxtop = maxtop; g.key = 0; g.depth = 0; g.root = r;
        }
        TC_HIP(hipMemcpyAsync(bsegs[0].p, roots.data(), sizeof(Seg) * nroot,
                              hipMemcpyHostToDevice, st));
        TC_HIP(hipMemsetAsync(bcnt.p, 0, 16, st));

        SpecBatch SB{};          // speculative levels not yet handed to an exact pass
        long long spec_segs = 0;
        int nspec = 0, nlaunch = 0;
        bool done = false;       // no further speculative levels
        auto launch_check = [&]() -> int {
            hipStream_t cs = guard.cs[nlaunch];
            hipEvent_t ev = guard.ev[nlaunch];
            ++nlaunch;
            TC_HIP(hipEventRecord(ev, st));
            TC_HIP(hipStreamWaitEvent(cs, ev, 0));
            hipLaunchKernelGGL((k_segment_check<D, C>), dim3((unsigned)spec_segs), dim3(CHECK_NT), 0,
                               cs, SB, M.nodes, P, bfinfo.as<FieldInfo>(), bcnt.as<int>() + 2);
            TC_HIP(hipGetLastError());
            SB = SpecBatch{};
            spec_segs = 0;
            return 0;
        };
        int cur = 0;
        long long nseg = nroot;
        int level = 0;
        while (nseg > 0) {
            Seg* segs = bsegs[cur].as<Seg>();
            SegOut* sout = bsout.as<SegOut>();
            const int l0 = level == 0 ? 1 : 0;
            // only levels above the top cells (their segments all have mintop > 0, so the
            // top-cell decision there depends on the centroid only through size == 0)
            bool sp = spec && !done && nspec < SPEC_LEVELS_MAX && nseg <= SEG_TOP_MAX &&
                      level < mintop && (spec_mode >= 2 || n / nseg >= SPEC_MIN_N);
            const size_t segb = al(sizeof(Seg) * nseg), soutb = al(sizeof(SegOut) * nseg);
            const size_t db = al(sizeof(double) * n), fb = al(sizeof(float) * n);
            if (sp && (size_t)(nspec + 1) * (4 * db + 3 * fb) > SPEC_SNAP_MAX) sp = false;
            if (sp) {
                DevBuf& sb = bsnap[nspec];
                if (!sb.p) {   // the first attempt allocates, a second reuses
                    const size_t need = segb + soutb + 4 * db + 3 * fb;
                    sb.p = dev_alloc(need);
                    if (sb.p) sb.bytes = need;
                    else sp = false;   // no room for the snapshot: exact from here on
                }
            }
            if (!sp && nspec > 0 && !done) {
                done = true;
                if (SB.nlev > 0 && launch_check()) return 2;
            }
            if (sp) {
                int K = (int)(SPEC_BLOCKS / nseg);
                const long long per = n / nseg / 1024;   // slices of about 1024 objects or more
                if (K > per) K = (int)per;
                if (K < 1) K = 1;
                double* c1 = bspec.as<double>();
                double* p1 = c1 + 4 * SEG_TOP_MAX;
                double* p2 = p1 + SPEC_P1 * SPEC_BLOCKS;
                double* hb = p2 + 7 * SPEC_BLOCKS;
                hipLaunchKernelGGL((k_spec_sums<C>), dim3(K, (unsigned)nseg), dim3(SPEC_NT), 0,
                                   st, segs, E[cur], K, p1);
                hipLaunchKernelGGL((k_spec_centroid<C>), dim3((unsigned)nseg), dim3(SPEC_NT), 0,
                                   st, segs, E[cur], K, p1, c1, hb);
                hipLaunchKernelGGL((k_spec_bounds<C>), dim3(K, (unsigned)nseg), dim3(SPEC_NT), 0,
                                   st, segs, E[cur], K, c1, p2);
                const SpecArgs sa{c1, p2, K, nullptr};
                hipLaunchKernelGGL((k_segment<D, C, 512, SPEC_FAST>), dim3((unsigned)nseg),
                                   dim3(512), 0, st, segs, 0LL, sout, E[cur], M.nodes,
                                   bcnt.as<int>(), btops.as<TopEntry>(), P,
                                   bfinfo.as<FieldInfo>(), l0, 0LL, sa);
                // the decisions' validation (Middle splits take an exact value)
                hipLaunchKernelGGL(k_spec_validate_segs, grid(nseg), dim3(TPB), 0, st, segs, nseg,
                                   K, p2, hb, bcnt.as<int>() + 3);
                if (sm == Mean)
                    hipLaunchKernelGGL(k_spec_validate_objs, grid(n), dim3(TPB), 0, st, E[cur],
                                       sout, n, hb, bcnt.as<int>() + 3);
                char* q = static_cast<char*>(bsnap[nspec].p);
                SpecLevel& L = SB.lv[SB.nlev++];
                Seg* ssegs = reinterpret_cast<Seg*>(q); q += segb;
                SegOut* ssout = reinterpret_cast<SegOut*>(q); q += soutb;
                Elems S{};
                S.x = reinterpret_cast<double*>(q); q += db;
                S.y = reinterpret_cast<double*>(q); q += db;
                S.z = reinterpret_cast<double*>(q); q += db;
                S.wp = reinterpret_cast<double*>(q); q += db;
                S.w = reinterpret_cast<float*>(q); q += fb;
                S.a = reinterpret_cast<float*>(q); q += fb;
                S.b = reinterpret_cast<float*>(q);
                const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
                TC_HIP(hipMemcpyAsync(ssegs, segs, sizeof(Seg) * nseg, dd, st));
                TC_HIP(hipMemcpyAsync(ssout, sout, sizeof(SegOut) * nseg, dd, st));
                TC_HIP(hipMemcpyAsync(S.x, E[cur].x, sizeof(double) * n, dd, st));
                TC_HIP(hipMemcpyAsync(S.y, E[cur].y, sizeof(double) * n, dd, st));
                TC_HIP(hipMemcpyAsync(S.z, E[cur].z, sizeof(double) * n, dd, st));
                TC_HIP(hipMemcpyAsync(S.wp, E[cur].wp, sizeof(double) * n, dd, st));
                TC_HIP(hipMemcpyAsync(S.w, E[cur].w, sizeof(float) * n, dd, st));
                TC_HIP(hipMemcpyAsync(S.a, E[cur].a, sizeof(float) * n, dd, st));
                TC_HIP(hipMemcpyAsync(S.b, E[cur].b, sizeof(float) * n, dd, st));
                L.segs = ssegs; L.sout = ssout; L.E = S; L.nseg = nseg; L.level0 = l0;
                spec_segs += nseg;
                // the first level's exact pass (the longest chains) starts at once
                if (++nspec == 1 && launch_check()) return 2;
            } else if (nseg <= SEG_TOP_MAX) {
                // Few large segments near the root: 512 threads each (the other waves
                // stage objects while one adds); the longest chains need the most stagers.
                hipLaunchKernelGGL((k_segment<D, C, 512>), dim3((unsigned)nseg), dim3(512),
                                   0, st, segs, 0LL, sout, E[cur], M.nodes, bcnt.as<int>(),
                                   btops.as<TopEntry>(), P, bfinfo.as<FieldInfo>(), l0, 0LL,
                                   SpecArgs{});
            } else if (nseg <= SEG_WIDE_MAX) {
                hipLaunchKernelGGL((k_segment<D, C, 256>), dim3((unsigned)nseg), dim3(256), 0, st,
                                   segs, 0LL, sout, E[cur], M.nodes, bcnt.as<int>(),
                                   btops.as<TopEntry>(), P, bfinfo.as<FieldInfo>(), l0, 0LL,
                                   SpecArgs{});
            } else {
                // many segments: the small ones one thread each, the rest one wave each
                hipLaunchKernelGGL((k_segment_small<D, C>), grid(nseg), dim3(TPB), 0, st, segs,
                                   nseg, sout, E[cur], M.nodes, bcnt.as<int>(),
                                   btops.as<TopEntry>(), P, bfinfo.as<FieldInfo>(), l0);
                hipLaunchKernelGGL((k_segment<D, C, 64>), dim3((unsigned)nseg), dim3(64), 0, st,
                                   segs, 0LL, sout, E[cur], M.nodes, bcnt.as<int>(),
                                   btops.as<TopEntry>(), P, bfinfo.as<FieldInfo>(), l0,
                                   SEG_SMALL_N, SpecArgs{});
            }
            hipLaunchKernelGGL(k_seg_split_flags, grid(nseg), dim3(TPB), 0, st, sout, (int)nseg,
                               bsegflags.as<int>());
            exclusive_scan(bsegflags.as<int>(), nseg, bsegrank.as<int>(), bscan.as<int>());
            TC_HIP(hipMemcpyAsync(hcnt, bsegrank.as<int>() + nseg, sizeof(int),
                                  hipMemcpyDeviceToHost, st));
            hipLaunchKernelGGL(k_part_flags, grid(n), dim3(TPB), 0, st, E[cur], sout, n,
                               bflags.as<int>());
            exclusive_scan(bflags.as<int>(), n, bpre.as<int>(), bscan.as<int>());
            hipLaunchKernelGGL(k_part_mid, grid(nseg), dim3(TPB), 0, st, segs, sout, (int)nseg,
                               bpre.as<int>());
            hipLaunchKernelGGL(k_part_median, grid(nseg), dim3(TPB), 0, st, segs, sout,
                               (int)nseg, E[cur], bkey.as<double>(), bord.as<int>(),
                               bdest.as<long long>());
            hipLaunchKernelGGL(k_part_rank, grid(n), dim3(TPB), 0, st, E[cur], segs, sout, n,
                               bflags.as<int>(), bpre.as<int>(), blpos.as<int>(),
                               brpos.as<int>());
            hipLaunchKernelGGL(k_part_dest, grid(n), dim3(TPB), 0, st, E[cur], segs, sout,
                               bsegrank.as<int>(), n, bflags.as<int>(), bpre.as<int>(),
                               blpos.as<int>(), brpos.as<int>(), bdest.as<long long>(),
                               bnewseg.as<int>());
            hipLaunchKernelGGL(k_permute, grid(n), dim3(TPB), 0, st, E[cur], E[1 - cur], n,
                               bdest.as<long long>(), bnewseg.as<int>());
            hipLaunchKernelGGL(k_children, grid(nseg), dim3(TPB), 0, st, segs, sout,
                               bsegrank.as<int>(), (int)nseg, bsegs[1 - cur].as<Seg>());
            TC_HIP(hipGetLastError());
            TC_HIP(hipStreamSynchronize(st));
            nseg = 2LL * hcnt[0];
            cur = 1 - cur;
            ++level;
            if (level > 4 * 64 + 64) return fail("tree build did not terminate");
        }
        nlevels = level;
        if (SB.nlev > 0 && launch_check()) return 2;
        deferred = false;
        if (spec && defer_ok && nlaunch > 0) {
            int vflag = 1;
            TC_HIP(hipMemcpyAsync(&vflag, bcnt.as<int>() + 3, sizeof(int), hipMemcpyDeviceToHost,
                                  st));
            TC_HIP(hipStreamSynchronize(st));
            deferred = vflag == 0;
        }
        if (!deferred)
            for (int k = 0; k < nlaunch; ++k) TC_HIP(hipStreamSynchronize(guard.cs[k]));
        else {
            // hand the exact pass's streams to the tree: events mark their end
            auto pc = std::make_shared<TreeCheck>();
            for (int k = 0; k < nlaunch; ++k) {
                TC_HIP(hipEventCreateWithFlags(&pc->done[k], hipEventDisableTiming));
                TC_HIP(hipEventRecord(pc->done[k], guard.cs[k]));
            }
            M.pending = pc;
            g_tree_spec_defer.fetch_add(1);
        }
        return 0;
    };

    int rc = attempt(spec_on);
    if (rc == 0 && !deferred && spec_on) {
        int flag = 0;
        TC_HIP(hipMemcpyAsync(&flag, bcnt.as<int>() + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (flag || spec_mode == 2) {
            // a speculative decision differed (or a test forces it): build again exactly
            g_tree_spec_redo.fetch_add(1);
            rc = attempt(false);
        }
    }
    int cnts[2];
    M.info.assign(nroot, FieldInfo{});
    if (rc == 0) {
        TC_HIP(hipMemcpyAsync(cnts, bcnt.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        if (!deferred)   // otherwise the exact pass writes it: settle_tree reads it
            TC_HIP(hipMemcpyAsync(M.info.data(), bfinfo.p, sizeof(FieldInfo) * nroot,
                                  hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
    }
    if (rc) return rc;
    if (rc == 0 && deferred) {
        // the buffers the exact pass still reads or writes go with the tree
        TreeCheck& pc = *M.pending;
        for (int k = 0; k < SPEC_LEVELS_MAX; ++k)
            if (bsnap[k].p) { pc.bufs.push_back({bsnap[k].p, bsnap[k].bytes}); bsnap[k].p = nullptr; }
        pc.finfo = bfinfo.as<FieldInfo>();
        pc.bufs.push_back({bfinfo.p, bfinfo.bytes});
        bfinfo.p = nullptr;
        pc.bad = bcnt.as<int>() + 2;
        pc.bufs.push_back({bcnt.p, bcnt.bytes});
        bcnt.p = nullptr;
        guard.cs[0] = guard.cs[1] = nullptr;     // the tree waits for them, not the build
    }
    M.nnodes = cnts[0];
    M.depth = nlevels;
    std::vector<TopEntry> tops(cnts[1]);
    TC_HIP(hipMemcpy(tops.data(), btops.p, sizeof(TopEntry) * cnts[1], hipMemcpyDeviceToHost));
    // each root's top cells in the reference's order (the path keys' order)
    std::sort(tops.begin(), tops.end(), [](const TopEntry& a, const TopEntry& b) {
        return a.root != b.root ? a.root < b.root : a.key < b.key;
    });
    M.top.assign(nroot, std::vector<int>());
    for (const TopEntry& t : tops) M.top[t.root].push_back(t.node);
    return 0;
}

int build_trees(int D, int C, const double* x, const double* y, const double* z, const double* d1,
                const double* d2, const double* w, const double* wpos, long long n,
                const std::vector<long long>& bounds, double minsize, double maxsize, int sm,
                int brute, int mintop, int maxtop, MultiTree& M) {
    if (bounds.size() < 2 || bounds.front() != 0 || bounds.back() != n)
        return fail("invalid root ranges");
    for (size_t r = 0; r + 1 < bounds.size(); ++r)
        if (bounds[r + 1] <= bounds[r]) return fail("every root range must hold objects");
#define TC_BT(DD, CC) return build_trees_t<DD, CC>(x, y, z, d1, d2, w, wpos, n, bounds, minsize, maxsize, sm, brute, mintop, maxtop, M)
    if (D == NData) { if (C == Flat) TC_BT(NData, Flat); if (C == ThreeD) TC_BT(NData, ThreeD); TC_BT(NData, Sphere); }
    if (D == KData) { if (C == Flat) TC_BT(KData, Flat); if (C == ThreeD) TC_BT(KData, ThreeD); TC_BT(KData, Sphere); }
    if (D == ZData) { if (C == Flat) TC_BT(ZData, Flat); if (C == ThreeD) TC_BT(ZData, ThreeD); TC_BT(ZData, Sphere); }
    if (D == VData) { if (C == Flat) TC_BT(VData, Flat); if (C == ThreeD) TC_BT(VData, ThreeD); TC_BT(VData, Sphere); }
    if (D == TData) { if (C == Flat) TC_BT(TData, Flat); if (C == ThreeD) TC_BT(TData, ThreeD); TC_BT(TData, Sphere); }
    if (D == QData) { if (C == Flat) TC_BT(QData, Flat); if (C == ThreeD) TC_BT(QData, ThreeD); TC_BT(QData, Sphere); }
    if (C == Flat) TC_BT(GData, Flat); if (C == ThreeD) TC_BT(GData, ThreeD); TC_BT(GData, Sphere);
#undef TC_BT
}

int build_tree(int D, int C, const double* x, const double* y, const double* z, const double* d1,
               const double* d2, const double* w, const double* wpos, long long n,
               double minsize, double maxsize, int sm, int brute, int mintop, int maxtop,
               DeviceTree& T) {
    MultiTree M;
    const int rc = build_trees(D, C, x, y, z, d1, d2, w, wpos, n, {0, n}, minsize, maxsize, sm,
                               brute, mintop, maxtop, M);
    T.nodes = M.nodes;      // the tree owns the node array from here on
    T.nodes_bytes = M.nodes_bytes;
    M.nodes = nullptr;
    T.pending = std::move(M.pending);
    if (rc) return rc;
    T.nnodes = M.nnodes;
    T.depth = M.depth;
    T.top = std::move(M.top[0]);
    T.info = M.info[0];
    T.dtop_bytes = sizeof(int) * std::max<size_t>(1, T.top.size());
    T.dtop = static_cast<int*>(dev_alloc(T.dtop_bytes));
    if (!T.dtop) return fail("out of device memory for the top cells");
    TC_HIP(hipMemcpy(T.dtop, T.top.data(), sizeof(int) * T.top.size(), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace tc
