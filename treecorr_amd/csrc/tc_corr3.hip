// treecorr_amd: the 3-point multipole algorithm (Porth et al.) on the GPU,
// auto-correlation of one field for NNN and GGG (LogMultipole binning).
//
// TreeCorr (src/Corr3.cpp:1113-2849) recurses per top cell on each CPU thread:
// multipoleSplit1 splits c1 until it is small enough for the largest bin, then
// multipoleFinish walks c1 depth first, accumulating the Fourier moments Gn of the
// accepted c2 cells into a scratch (MultipoleScratch) that is copied whenever c1 is
// split, and contracts finished rows into zeta.  Here every level of both recursions
// is one breadth-first launch over all cells of the level, driven from the host:
//   1. k_mp_split_level: one workgroup per (c1, c2 list) of multipoleSplit1
//      (:1374-1408, splitC2Cells :1323-1372).  Children and the "units" (c1 small
//      enough, with its c2 list) go to fixed per-task positions; a scan compacts them
//      in task order, so every later order is independent of scheduling.
//   2. k_mp_stage: one workgroup per multipoleFinish stage (:1548-1613,
//      splitC2CellsOrCalculateGn :1452-1546).  The stage's scratch lives in a slot in
//      HBM, is loaded into LDS, receives the accepted pairs' moments (calculateGn
//      :1681-1828) and is written back.  A split of c1 hands the left child the slot
//      and the right child a copy in a spare slot (mp.duplicate()).  A stage that can
//      contract rows of zeta records a job (the slot and the row range).
//   3. k_mp_zeta: the jobs of a level, contracted (calculateZeta :1861-1974,
//      MultipoleHelper<5> :2582-2849) by one workgroup per (job chunk, row k2) into an
//      LDS tile of the compressed outputs, written per chunk and summed in chunk order.
//   4. Free slots of finished stages and unused spares return to a free list (scan).
// The outputs are kept compressed (k3 >= k2, n >= 0: the other three quarters of the
// reference's arrays are copies or conjugates) and expanded into TreeCorr's layout at
// the end (k_mp_expand).
#include "tc_mp.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

namespace tc {

// Output arrays per slot: zeta0..7, meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri
enum { O_Z0 = 0, O_MEAND2 = 8, O_MEANLOGD2, O_MEAND3, O_MEANLOGD3, O_WEIGHT, O_WEIGHT_IM, O_NTRI,
       NOUT3 = 15 };
// counters: Gn calls, zeta calls, zeta rows, finish stages, zeta cells (k2 row, k3 >= k2)
enum { C3_GN = 0, C3_ZETA, C3_ROWS, C3_STAGES, C3_CELLS, NC3 = 5 };

constexpr int MP_THREADS = 256;

template <int C>
__device__ __forceinline__ double dsq3(const DNode& a, const DNode& b) {
    double dx = a.x - b.x, dy = a.y - b.y;
    if (C == Flat) return dx * dx + dy * dy;
    double dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}

// BinTypeHelper<Log>::singleBin, include/BinType.h:93-151
__device__ __forceinline__ bool single_bin3(const C3Params& P, double rsq, double s1ps2, int& ik,
                                            double& r, double& logr) {
    if (s1ps2 == 0.) return true;
    const double s1ps2sq = s1ps2 * s1ps2;
    if (s1ps2sq > P.asq * rsq) return false;
    if (s1ps2sq <= P.bsq * rsq) return true;
    if (s1ps2sq > 0.25 * ((P.binsize + P.b) * (P.binsize + P.b)) * rsq) return false;
    // a decision within a few ulp of a bin edge or of the refined test's threshold is
    // taken again with the correctly rounded log, as the host libm rounds it
    double L = log(rsq);
    logr = 0.5 * L;
    double kk = (logr - P.logminsep) / P.binsize;
    const double sl = kk_slack(logr, P.logminsep, 1. / P.binsize, kk);
    bool near = near_int(kk, sl);
    bool fits = refined_fits(kk, s1ps2sq, rsq, P.binsize, P.b, sl, near);
    if (near) {
        L = log_cr(rsq, L);
        logr = 0.5 * L;
        kk = (logr - P.logminsep) / P.binsize;
        fits = refined_fits(kk, s1ps2sq, rsq, P.binsize, P.b, 0., near);
    }
    ik = int(kk);
    if (!fits) return false;
    r = sqrt(rsq);
    return true;
}

__device__ __forceinline__ bool too_small3(const C3Params& P, double rsq, double s1ps2) {
    return rsq < P.minsepsq && s1ps2 < P.minsep && rsq < (P.minsep - s1ps2) * (P.minsep - s1ps2);
}
__device__ __forceinline__ bool too_large3(const C3Params& P, double rsq, double s1ps2) {
    return rsq >= P.maxsepsq && rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
}

// ProjectHelper<Sphere>::calculate_direction (include/ProjectHelper.h:272-345)
__device__ __forceinline__ void direction3(double x1, double y1, double z1, double x2, double y2,
                                           double z2, double& re, double& im) {
    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
    double dsq = dx * dx + dy * dy + dz * dz;
    re = y1 * x2 - x1 * y2;
    im = -((z1 - z2) + 0.5 * z2 * dsq);
}

// ---------------------------------------------------------------------------
// Node sums, bottom up (iterated until every node is done)
// ---------------------------------------------------------------------------
// A node is ready once both children were finished by an earlier launch; flags are
// read from done_in and written to done_out so no launch depends on its own writes.
__global__ void k_node_sums(const DNode* nodes, int nn, NodeSum* sums, const int* done_in,
                            int* done_out, int* changed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    if (done_in[i]) { done_out[i] = 1; return; }
    const DNode c = nodes[i];
    NodeSum s;
    if (c.size == 0.f || c.n == 1 || c.child < 0) {
        double w = c.w, a = c.a, b = c.b;
        s.wsq = w * w;
        s.wzr = a * a - b * b;
        s.wzi = a * b + b * a;
        s.abswz = a * a + b * b;
    } else {
        const int L = c.child;
        if (L >= nn - 1) { done_out[i] = 1; *changed = 1; sums[i] = NodeSum{0., 0., 0., 0.}; return; }
        if (!done_in[L] || !done_in[L + 1]) { done_out[i] = 0; return; }
        const NodeSum l = sums[L];
        const NodeSum r = sums[L + 1];
        s.wsq = l.wsq + r.wsq;
        s.wzr = l.wzr + r.wzr;
        s.wzi = l.wzi + r.wzi;
        s.abswz = l.abswz + r.abswz;
    }
    sums[i] = s;
    done_out[i] = 1;
    *changed = 1;
}

// Corr3::calculateGn + MultipoleScratch::calculateGn (src/Corr3.cpp:1681-1828) for one
// accepted (c1,c2), added into the LDS scratch with FP64 atomics.
//
// The lanes of a wavefront usually hold pairs of the same bin k, so a plain loop over
// m would make every lane of an atomic add hit the same address.  Each lane therefore
// walks the multipoles in its own cyclic order, starting at m0 = lane mod (M+2): one
// power sequence p = z^m (m = 0..M+1) gives W(k,m) += w p, G(k,m) += g p and
// G(k,-m) += g conj(p) (|z| = 1, so conj(z)^m = conj(z^m)).  The reference multiplies
// the powers up one at a time from m = 0; a lane here starts from z^m0 by repeated
// squaring, so the terms differ from the reference's in the last bits only.
__device__ __forceinline__ void cpow_int(double zr, double zi, int e, double& pr, double& pi) {
    double br = zr, bi = zi;
    pr = 1.;
    pi = 0.;
    for (int bit = 0; bit < 7; ++bit) {
        if (e & 1) {
            double tr, ti;
            cmul3(pr, pi, br, bi, tr, ti);
            pr = tr; pi = ti;
        }
        e >>= 1;
        if (e == 0) break;
        double tr, ti;
        cmul3(br, bi, br, bi, tr, ti);
        br = tr; bi = ti;
    }
}

template <int D, int C>
__device__ __forceinline__ void add_gn(const ScratchLayout& Lo, double* sc, const DNode& c1,
                                       const DNode& c2, const NodeSum& s2, double r,
                                       double logr, int k, int maxn, int plain) {
    const double n = double(c2.n), w = c2.w;
    atomicAdd(&sc[Lo.S(S_NPAIRS, k)], n);
    atomicAdd(&sc[Lo.S(S_SUMW, k)], w);
    atomicAdd(&sc[Lo.S(S_SUMWR, k)], w * r);
    atomicAdd(&sc[Lo.S(S_SUMWLOGR, k)], w * logr);
    atomicAdd(&sc[Lo.S(S_SUMWW, k)], s2.wsq);
    atomicAdd(&sc[Lo.S(S_SUMWWR, k)], s2.wsq * r);
    atomicAdd(&sc[Lo.S(S_SUMWWLOGR, k)], s2.wsq * logr);
    // ProjectHelper<C>::ExpIPhi (ProjectHelper.h:264-266, 526-535, 681-688)
    double zr, zi;
    double x1 = c1.x, y1 = c1.y, z1 = c1.z, x2 = c2.x, y2 = c2.y, z2 = c2.z;
    if (C == Flat) {
        zr = (x2 - x1) / r;
        zi = (y2 - y1) / r;
    } else {
        if (C == ThreeD) { normalize_recip(x1, y1, z1); normalize_recip(x2, y2, z2); }
        double dr, di;
        direction3(x2, y2, z2, x1, y1, z1, dr, di);
        double nn = dr * dr + di * di;
        nn = nn > 0. ? nn : 1.;
        double sn = sqrt(nn);
        zr = dr / sn;
        zi = di / sn;
    }
    double gr = 0., gi = 0.;
    if (D == GData) {
        // ProjectWithSq<GData> (ProjectHelper.h:115-124, 357-368)
        double er, ei;
        if (C == Flat) {
            expm2ialpha(x2 - x1, y2 - y1, er, ei);
        } else {
            double dr, di;
            direction3(x1, y1, z1, x2, y2, z2, dr, di);
            expm2ialpha(dr, di, er, ei);
        }
        cmul3(c2.a, c2.b, er, ei, gr, gi);
        double e2r, e2i;
        cmul3(er, ei, er, ei, e2r, e2i);
        double qr, qi;
        cmul3(s2.wzr, s2.wzi, e2r, e2i, qr, qi);          // wgsq
        double zsr, zsi;
        cmul3(zr, zi, zr, zi, zsr, zsi);                   // z^2
        double tr, ti;
        cmul3(qr, qi, zsr, -zsi, tr, ti);
        atomicAdd(&sc[Lo.GG(0, k)], tr); atomicAdd(&sc[Lo.GG(0, k) + 1], ti);
        cmul3(qr, qi, zsr, zsi, tr, ti);
        atomicAdd(&sc[Lo.GG(1, k)], tr); atomicAdd(&sc[Lo.GG(1, k) + 1], ti);
        cmul3(s2.abswz, 0., zsr, -zsi, tr, ti);
        atomicAdd(&sc[Lo.GG(2, k)], tr); atomicAdd(&sc[Lo.GG(2, k) + 1], ti);
    }
    // W(k, 0..M); G(k, -(M+1)..M+1) for GGG
    const int nm = D == GData ? maxn + 2 : maxn + 1;
    int m = plain ? 0 : (int)(threadIdx.x & 63u) % nm;
    double pr, pi;
    cpow_int(zr, zi, m, pr, pi);
    for (int t = 0; t < nm; ++t) {
        if (m <= maxn) {
            atomicAdd(&sc[Lo.W(k, m)], w * pr);
            atomicAdd(&sc[Lo.W(k, m) + 1], w * pi);
        }
        if (D == GData) {
            double tr, ti;
            cmul3(gr, gi, pr, pi, tr, ti);
            atomicAdd(&sc[Lo.G(k, m)], tr);
            atomicAdd(&sc[Lo.G(k, m) + 1], ti);
            if (m > 0) {
                cmul3(gr, gi, pr, -pi, tr, ti);
                atomicAdd(&sc[Lo.G(k, -m)], tr);
                atomicAdd(&sc[Lo.G(k, -m) + 1], ti);
            }
        }
        // next multipole, cyclically
        double tr, ti;
        cmul3(pr, pi, zr, zi, tr, ti);
        pr = tr; pi = ti;
        if (++m == nm) { m = 0; pr = 1.; pi = 0.; }
    }
}
// ---------------------------------------------------------------------------
// Records of the two breadth-first recursions
// ---------------------------------------------------------------------------
// multipoleSplit1(c1, list): the list is pool[off, off + len)
struct SplitTask { int c1, off, len, pad; };
// multipoleFinish(c1, list, mp, mink_zeta, maxr): mp lives in slot `slot`; `spare` is a
// free slot for the right child of a split; fresh: the slot's scratch starts at zero.
// With two lists, maxr is the first list's largest remaining separation and maxr3 the
// second's (multipoleFinish's maxr2, maxr3, src/Corr3.cpp:1615-1679).
struct MPTask {
    int c1, off, len, mink;
    double maxr;
    int slot, spare, fresh, pad;
    double maxr3;
};
__device__ __forceinline__ bool has_children(const DNode& c, int nnodes) {
    return c.child >= 0 && c.child < nnodes - 1;
}

// ---------------------------------------------------------------------------
// Workgroup helpers (every thread of the workgroup calls them)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int block_excl_scan(int v, int* sh, int& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int base = 0;
    total = 0;
    for (int k = 0; k < nw; ++k) {
        const int s = sh[k];
        if (k < wv) base += s;
        total += s;
    }
    __syncthreads();
    return base + x - v;
}

__device__ __forceinline__ double block_max(double v, double* shd) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_max(v);
    if (lane == 0) shd[wv] = v;
    __syncthreads();
    double m = shd[0];
    for (int k = 1; k < nw; ++k) m = fmax(m, shd[k]);
    __syncthreads();
    return m;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v,
                                                            unsigned long long* shl) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) shl[wv] = v;
    __syncthreads();
    unsigned long long s = 0;
    for (int k = 0; k < nw; ++k) s += shl[k];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// The per-c2 decisions
// ---------------------------------------------------------------------------
// splitC2Cells (src/Corr3.cpp:1323-1372): 0 dropped, 1 kept, 2 replaced by its children
template <int C>
__device__ __forceinline__ int split_class(const C3Params& P, const DNode& c1, double s1,
                                           const DNode& c2) {
    const double s2 = c2.size;
    const double rsq = dsq3<C>(c1, c2);
    const double s1ps2 = s1 + s2;
    if (too_small3(P, rsq, s1ps2) || too_large3(P, rsq, s1ps2)) return 0;
    int k = -1;
    double r = 0., logr = 0.;
    const bool split = !single_bin3(P, rsq, s1ps2, k, r, logr);
    return (split && s2 > s1) ? 2 : 1;
}

// splitC2CellsOrCalculateGn (src/Corr3.cpp:1486-1545) for one c2: 0 dropped, 1 accepted
// into bin k (with r, logr), 2 kept for the next stage (children if split2), with its
// largest remaining separation rtot and whether c1 needs a split for it (CalcSplitSq,
// include/Split.h:82-101).
template <int C>
__device__ __forceinline__ int finish_class(const C3Params& P, const DNode& c1, double s1,
                                            const DNode& c2, int& k, double& r, double& logr,
                                            double& rtot, bool& split1, bool& split2) {
    const double s2 = c2.size;
    const double rsq = dsq3<C>(c1, c2);
    const double s1ps2 = s1 + s2;
    if (too_small3(P, rsq, s1ps2) || too_large3(P, rsq, s1ps2)) return 0;
    k = -1;
    r = sqrt(rsq);
    logr = 0.;
    if (single_bin3(P, rsq, s1ps2, k, r, logr)) {
        if (!(rsq >= P.minsepsq && rsq < P.maxsepsq)) return 0;
        if (k < 0) k = guarded_bin(r, log(r), P.logminsep, P.binsize, logr);
        return 1;
    }
    rtot = r + s1ps2;
    const double bsq_eff = rsq * (P.asq < P.bsq ? P.asq : P.bsq);
    if (s2 > s1) {
        split2 = true;
        split1 = (s2 > 2. * s1) ? false : (s1 * s1 > 0.3422 * bsq_eff);
    } else if (s1 > 2. * s2) {
        split1 = true;
        split2 = false;
    } else {
        split1 = true;
        split2 = s2 * s2 > 0.3422 * bsq_eff;
    }
    return 2;
}

// ---------------------------------------------------------------------------
// Phase 1: one level of multipoleSplit1
// ---------------------------------------------------------------------------
// c1 runs over c1top, every task's list is ltop (BaseCorr3::multipole, src/Corr3.cpp:
// 1124-1305: the list of all top cells of the other field, or fields)
__global__ void k_split_init(const int* __restrict__ c1top, int nc1, const int* __restrict__ ltop,
                             int nl, SplitTask* __restrict__ tasks, int* __restrict__ pool) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nl) pool[i] = ltop[i];
    if (i < nc1) tasks[i] = SplitTask{c1top[i], 0, nl, 0};
}

template <int C>
__global__ __launch_bounds__(MP_THREADS) void k_mp_split_level(
    const DNode* __restrict__ nodes, int nnodes, const SplitTask* __restrict__ tasks,
    int* __restrict__ pool, unsigned long long* pool_top, unsigned long long pool_cap, C3Params P,
    SplitTask* __restrict__ child_out, SplitTask* __restrict__ unit_out, int* err) {
    __shared__ int sh[MP_THREADS / 64];
    __shared__ unsigned long long s_base;
    __shared__ int s_bad;
    const int t = blockIdx.x;
    const SplitTask tk = tasks[t];
    const SplitTask none{-1, 0, 0, 0};
    const bool ok = (unsigned)tk.c1 < (unsigned)nnodes && tk.off >= 0 && tk.len >= 0 &&
                    (unsigned long long)tk.off + (unsigned long long)tk.len <= pool_cap;
    if (!ok) {
        if (threadIdx.x == 0) {
            atomicOr(err, E_INDEX);
            child_out[2 * t] = none; child_out[2 * t + 1] = none; unit_out[t] = none;
        }
        return;
    }
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    const DNode c1 = nodes[tk.c1];
    const double s1 = c1.size;
    const int* list = pool + tk.off;
    int total = 0;
    for (int base = 0; base < tk.len; base += blockDim.x) {
        const int i = base + threadIdx.x;
        int no = 0;
        if (i < tk.len) {
            const int i2 = list[i];
            if ((unsigned)i2 >= (unsigned)nnodes) {
                s_bad = 1;
            } else {
                const DNode c2 = nodes[i2];
                const int cl = split_class<C>(P, c1, s1, c2);
                if (cl == 2 && !has_children(c2, nnodes)) s_bad = 1;
                else no = cl;
            }
        }
        int tot;
        block_excl_scan(no, sh, tot);
        total += tot;
    }
    if (threadIdx.x == 0) s_base = total > 0 ? atomicAdd(pool_top, (unsigned long long)total) : 0ull;
    __syncthreads();
    const unsigned long long obase = s_base;
    const bool split_c1 = s1 * s1 > P.maxbsq_eff;
    const int bad = s_bad ? E_INDEX : (obase + (unsigned long long)total > pool_cap ||
                                       obase + (unsigned long long)total > 0x7fffffffull) ? E_POOL
                    : (split_c1 && !has_children(c1, nnodes)) ? E_INDEX : 0;
    if (bad) {
        if (threadIdx.x == 0) {
            atomicOr(err, bad);
            child_out[2 * t] = none; child_out[2 * t + 1] = none; unit_out[t] = none;
        }
        return;
    }
    int wpos = 0;
    for (int base = 0; base < tk.len; base += blockDim.x) {
        const int i = base + threadIdx.x;
        int no = 0, o0 = 0;
        if (i < tk.len) {
            const int i2 = list[i];
            const DNode c2 = nodes[i2];
            no = split_class<C>(P, c1, s1, c2);
            o0 = no == 2 ? c2.child : i2;
        }
        int tot;
        const int ex = block_excl_scan(no, sh, tot);
        if (no >= 1) pool[obase + wpos + ex] = o0;
        if (no == 2) pool[obase + wpos + ex + 1] = o0 + 1;
        wpos += tot;
    }
    if (threadIdx.x == 0) {
        const int noff = (int)obase;
        if (split_c1) {
            child_out[2 * t] = SplitTask{c1.child, noff, total, 0};
            child_out[2 * t + 1] = SplitTask{c1.child + 1, noff, total, 0};
            unit_out[t] = none;
        } else {
            child_out[2 * t] = none; child_out[2 * t + 1] = none;
            unit_out[t] = SplitTask{tk.c1, noff, total, 0};
        }
    }
}

// ---------------------------------------------------------------------------
// Order-preserving compaction of fixed-position records (c1 < 0 or id < 0: empty)
// ---------------------------------------------------------------------------
constexpr int CP_TPB = 256, CP_ITEMS = 4, CP_BLOCK = CP_TPB * CP_ITEMS;

__device__ __forceinline__ bool rec_valid(const SplitTask& r) { return r.c1 >= 0; }
__device__ __forceinline__ bool rec_valid(const MPTask& r) { return r.c1 >= 0; }
__device__ __forceinline__ bool rec_valid(const int& r) { return r >= 0; }
__device__ __forceinline__ unsigned long long rec_len(const SplitTask& r) { return (unsigned)r.len; }
__device__ __forceinline__ unsigned long long rec_len(const MPTask& r) { return (unsigned)r.len; }
__device__ __forceinline__ unsigned long long rec_len(const int&) { return 0ull; }

template <class R>
__global__ __launch_bounds__(CP_TPB) void k_cp_count(const R* __restrict__ in, long long n,
                                                      int* __restrict__ bcount,
                                                      unsigned long long* __restrict__ blen) {
    __shared__ int sh[CP_TPB / 64];
    __shared__ unsigned long long shl[CP_TPB / 64];
    const long long base = (long long)blockIdx.x * CP_BLOCK;
    int c = 0;
    unsigned long long l = 0;
    for (int q = 0; q < CP_ITEMS; ++q) {
        const long long i = base + q * CP_TPB + threadIdx.x;
        if (i < n) {
            const R r = in[i];
            if (rec_valid(r)) { ++c; l += rec_len(r); }
        }
    }
    int tot;
    block_excl_scan(c, sh, tot);
    const unsigned long long ltot = block_sum_u64(l, shl);
    if (threadIdx.x == 0) { bcount[blockIdx.x] = tot; blen[blockIdx.x] = ltot; }
}

// one workgroup: exclusive offsets of the block counts; totals = {count, sum of len}
__global__ __launch_bounds__(1024) void k_cp_scan(const int* __restrict__ bcount,
                                                  const unsigned long long* __restrict__ blen,
                                                  int nblk, int* __restrict__ boff,
                                                  unsigned long long* __restrict__ totals) {
    __shared__ int sh[1024 / 64];
    __shared__ unsigned long long shl[1024 / 64];
    const int per = (nblk + (int)blockDim.x - 1) / (int)blockDim.x;
    const int b0 = threadIdx.x * per;
    int c = 0;
    unsigned long long l = 0;
    for (int b = b0; b < b0 + per && b < nblk; ++b) { c += bcount[b]; l += blen[b]; }
    int tot;
    const int ex = block_excl_scan(c, sh, tot);
    const unsigned long long ltot = block_sum_u64(l, shl);
    int run = ex;
    for (int b = b0; b < b0 + per && b < nblk; ++b) { boff[b] = run; run += bcount[b]; }
    if (threadIdx.x == 0) { totals[0] = (unsigned long long)tot; totals[1] = ltot; }
}

template <class R>
__global__ __launch_bounds__(CP_TPB) void k_cp_scatter(const R* __restrict__ in, long long n,
                                                        const int* __restrict__ boff,
                                                        R* __restrict__ out, long long out_base) {
    __shared__ int sh[CP_TPB / 64];
    const long long base = (long long)blockIdx.x * CP_BLOCK;
    long long run = out_base + boff[blockIdx.x];
    for (int q = 0; q < CP_ITEMS; ++q) {
        const long long i = base + q * CP_TPB + threadIdx.x;
        R r{};
        int v = 0;
        if (i < n) {
            r = in[i];
            v = rec_valid(r) ? 1 : 0;
        }
        int tot;
        const int ex = block_excl_scan(v, sh, tot);
        if (v) out[run + ex] = r;
        run += tot;
    }
}

// ---------------------------------------------------------------------------
// Phase 2: one level of multipoleFinish stages
// ---------------------------------------------------------------------------
__global__ void k_units_to_tasks(const SplitTask* __restrict__ units, int n, int nbins,
                                 MPTask* __restrict__ tasks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SplitTask u = units[i];
    MPTask t;
    t.c1 = u.c1; t.off = u.off; t.len = u.len; t.mink = nbins; t.maxr = 0.; t.maxr3 = 0.;
    t.slot = -1; t.spare = -1; t.fresh = 1; t.pad = 0;
    tasks[i] = t;
}

// Spare slot for every task (the right child of a split), and a slot for tasks that have
// none yet, from the top of the free list in task order.
__global__ void k_mp_prep(MPTask* __restrict__ tasks, int n, const int* __restrict__ free_ids,
                          int free_top) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    MPTask t = tasks[i];
    t.spare = free_ids[free_top - 1 - i];
    if (t.slot < 0) { t.slot = free_ids[free_top - 1 - n - i]; t.fresh = 1; }
    tasks[i] = t;
}

// X: the two-list form (src/Corr3.cpp:1615-1679).  A slot then holds two scratches, the
// first list's at offset 0 and the second's at Lo.total; a list entry with node index
// >= P.base3 is the second list's and adds into the second scratch, and each list keeps
// its own largest remaining separation (maxr, maxr3) for its split threshold.
template <int D, int C, bool X>
__global__ __launch_bounds__(MP_THREADS) void k_mp_stage(
    const DNode* __restrict__ nodes, int nnodes, const NodeSum* __restrict__ nsum,
    const MPTask* __restrict__ tasks, const int* __restrict__ pool_in,
    unsigned long long pool_in_cap, int* __restrict__ pool_out, unsigned long long* pool_out_top,
    unsigned long long pool_out_cap, double* const* __restrict__ slots, int nslots, C3Params P,
    MPTask* __restrict__ child_out, MPJob* __restrict__ job_out, int* __restrict__ ret_out,
    unsigned long long* __restrict__ counters, int* err) {
    HIP_DYNAMIC_SHARED(double, sc)
    __shared__ int sh[MP_THREADS / 64];
    __shared__ double shd[MP_THREADS / 64];
    __shared__ unsigned long long shl[MP_THREADS / 64];
    __shared__ unsigned long long s_base;
    __shared__ int s_bad;
    ScratchLayout Lo;
    Lo.init(P.nbins, P.maxn, D);
    const int stot = X ? 2 * Lo.total : Lo.total;      // doubles per slot
    const int t = blockIdx.x;
    const MPTask tk = tasks[t];
    MPTask none;
    none.c1 = -1; none.off = 0; none.len = 0; none.mink = 0; none.maxr = 0.; none.maxr3 = 0.;
    none.slot = -1; none.spare = -1; none.fresh = 0; none.pad = 0;
    const MPJob nojob{-1, -1, 0, 0};
    const bool ok = (unsigned)tk.c1 < (unsigned)nnodes && tk.off >= 0 && tk.len >= 0 &&
                    (unsigned long long)tk.off + (unsigned long long)tk.len <= pool_in_cap &&
                    (unsigned)tk.slot < (unsigned)nslots && (unsigned)tk.spare < (unsigned)nslots &&
                    tk.slot != tk.spare && tk.mink >= 0 && tk.mink <= P.nbins;
    if (!ok) {
        if (threadIdx.x == 0) {
            atomicOr(err, E_SLOT);
            child_out[2 * t] = none; child_out[2 * t + 1] = none; job_out[t] = nojob;
            ret_out[2 * t] = -1; ret_out[2 * t + 1] = -1;
        }
        return;
    }
    double* gs = slot_at(slots, tk.slot, stot);
    if (tk.fresh) {
        for (int i = threadIdx.x; i < stot; i += blockDim.x) sc[i] = 0.;
    } else {
        for (int i = threadIdx.x; i < stot; i += blockDim.x) sc[i] = gs[i];
    }
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    const DNode c1 = nodes[tk.c1];
    const double s1 = c1.size;
    // BinTypeHelper<Log>::oneBinLessThan of each list's previous largest separation
    const double thresh1 = (1. - P.binsize) * tk.maxr;
    const double thresh13 = (1. - P.binsize) * tk.maxr3;
    const int* list = pool_in + tk.off;
    // pass 1: accepted pairs into the scratch, the size of the next list, the largest
    // remaining separation and whether any c2 needs c1 split
    int nlen = 0, lany = 0;
    double lmax = 0., lmax3 = 0.;
    unsigned long long c_gn = 0;
    for (int base = 0; base < tk.len; base += blockDim.x) {
        const int i = base + threadIdx.x;
        int no = 0;
        if (i < tk.len) {
            const int i2 = list[i];
            if ((unsigned)i2 >= (unsigned)nnodes) {
                s_bad = 1;
            } else {
                const bool second = X && i2 >= P.base3;
                const DNode c2 = nodes[i2];
                int k = -1;
                double r = 0., logr = 0., rtot = 0.;
                bool sp1 = false, sp2 = false;
                const int cl = finish_class<C>(P, c1, s1, c2, k, r, logr, rtot, sp1, sp2);
                if (cl == 1) {
                    if (k >= 0 && k < P.nbins) {
                        ++c_gn;
                        add_gn<D, C>(Lo, second ? sc + Lo.total : sc, c1, c2, nsum[i2], r, logr,
                                     k, P.maxn, P.plain);
                    } else {
                        s_bad = 1;     // r at the very edge of the range: no bin
                    }
                } else if (cl == 2) {
                    if (second) lmax3 = fmax(lmax3, rtot);
                    else lmax = fmax(lmax, rtot);
                    if (sp1 && rtot > (second ? thresh13 : thresh1)) lany = 1;
                    if (sp2 && !has_children(c2, nnodes)) s_bad = 1;
                    no = sp2 ? 2 : 1;
                }
            }
        }
        int tot;
        block_excl_scan(no, sh, tot);
        nlen += tot;
    }
    const double nmaxr2 = block_max(lmax, shd);
    const double nmaxr3 = X ? block_max(lmax3, shd) : 0.;
    const double nmaxr = fmax(nmaxr2, nmaxr3);
    int nany;
    block_excl_scan(lany, sh, nany);
    const unsigned long long gn_tot = block_sum_u64(c_gn, shl);
    if (threadIdx.x == 0) s_base = nlen > 0 ? atomicAdd(pool_out_top, (unsigned long long)nlen) : 0ull;
    __syncthreads();
    const unsigned long long obase = s_base;
    const bool split = nlen > 0 && nany > 0;
    const int bad = s_bad ? E_INDEX
                    : (obase + (unsigned long long)nlen > pool_out_cap ||
                       obase + (unsigned long long)nlen > 0x7fffffffull) ? E_POOL
                    : (split && !has_children(c1, nnodes)) ? E_INDEX : 0;
    if (bad) {
        if (threadIdx.x == 0) {
            atomicOr(err, bad);
            child_out[2 * t] = none; child_out[2 * t + 1] = none; job_out[t] = nojob;
            ret_out[2 * t] = -1; ret_out[2 * t + 1] = -1;
        }
        return;
    }
    // pass 2: the next list, in list order (the same decisions as pass 1)
    if (nlen > 0) {
        int wpos = 0;
        for (int base = 0; base < tk.len; base += blockDim.x) {
            const int i = base + threadIdx.x;
            int no = 0, o0 = 0;
            if (i < tk.len) {
                const int i2 = list[i];
                const DNode c2 = nodes[i2];
                int k = -1;
                double r = 0., logr = 0., rtot = 0.;
                bool sp1 = false, sp2 = false;
                const int cl = finish_class<C>(P, c1, s1, c2, k, r, logr, rtot, sp1, sp2);
                if (cl == 2) { no = sp2 ? 2 : 1; o0 = sp2 ? c2.child : i2; }
            }
            int tot;
            const int ex = block_excl_scan(no, sh, tot);
            if (no >= 1) pool_out[obase + wpos + ex] = o0;
            if (no == 2) pool_out[obase + wpos + ex + 1] = o0 + 1;
            wpos += tot;
        }
    }
    // multipoleFinish's decisions after splitC2CellsOrCalculateGn (src/Corr3.cpp:1574-1611,
    // 1637-1677)
    MPJob job = nojob;
    int nmink = tk.mink;
    if (nlen > 0) {
        if (nmaxr < P.maxsep) {
            double lm = 0.;
            const int k = nmaxr < P.minsep ? 0
                        : guarded_bin(nmaxr, log(nmaxr), P.logminsep, P.binsize, lm) + 1;
            if (k >= 0 && k < tk.mink) { job = MPJob{tk.c1, tk.slot, k, tk.mink}; nmink = k; }
        }
    } else {
        job = MPJob{tk.c1, tk.slot, 0, tk.mink};
    }
    // the scratch after this stage: this level's zeta reads it from the slot; the right
    // child of a split continues from a copy (mp.duplicate())
    double* gsp = split ? slot_at(slots, tk.spare, stot) : nullptr;
    for (int i = threadIdx.x; i < stot; i += blockDim.x) {
        const double v = sc[i];
        gs[i] = v;
        if (split) gsp[i] = v;
    }
    if (threadIdx.x == 0) {
        job_out[t] = job;
        MPTask a = none, b = none;
        int r0 = -1, r1 = -1;
        if (nlen > 0) {
            a.off = (int)obase; a.len = nlen; a.mink = nmink;
            a.maxr = X ? nmaxr2 : nmaxr; a.maxr3 = nmaxr3;
            if (split) {
                b = a;
                a.c1 = c1.child; a.slot = tk.slot;
                b.c1 = c1.child + 1; b.slot = tk.spare;
            } else {
                a.c1 = tk.c1; a.slot = tk.slot;
                r0 = tk.spare;
            }
        } else {
            r0 = tk.spare;
            r1 = tk.slot;       // free after this level's zeta
        }
        child_out[2 * t] = a; child_out[2 * t + 1] = b;
        ret_out[2 * t] = r0; ret_out[2 * t + 1] = r1;
        atomicAdd(&counters[C3_GN], gn_tot);
        atomicAdd(&counters[C3_STAGES], 1ull);
        if (job.c1 >= 0) {
            const int nb = P.nbins;
            atomicAdd(&counters[C3_ZETA], 1ull);
            if (X) {
                // rows [ks, nb); cells: {k2, k3 >= ks} without {k2, k3 >= mk}
                const long long a2 = nb - job.ks, b2 = nb - job.mk;
                atomicAdd(&counters[C3_ROWS], (unsigned long long)a2);
                atomicAdd(&counters[C3_CELLS], (unsigned long long)(a2 * a2 - b2 * b2));
            } else {
                atomicAdd(&counters[C3_ROWS], (unsigned long long)(job.mk - job.ks));
                // cells (k2, k3 >= k2) of the rows [ks, mk)
                const long long a2 = nb - job.ks, b2 = nb - job.mk;
                atomicAdd(&counters[C3_CELLS], (unsigned long long)((a2 * (a2 + 1) - b2 * (b2 + 1)) / 2));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Phase 3: zeta of the jobs of one level
// ---------------------------------------------------------------------------
// Compressed outputs per (k2, k3 >= k2, m >= 0) cell, NV doubles:
//   [0,1]      w1 W(k2,m) W(k3,m)*  (k3 == k2: w1 (|W(k2,m)|^2 - sumww), real)
//   GGG [2,9]  P0..P3: zeta0..3 at (k2, k3, +m); [10,17] Q0..Q3: zeta0..3 at (k3, k2, +m)
//   last 5     ntri, meand2, meanlogd2, meand3, meanlogd3 at (k2, k3, n = 0) (m = 0 only)
// The other entries of the reference's arrays follow from these (k_mp_expand).
template <int D>
__device__ __forceinline__ void zeta_cell(const ScratchLayout& Lo, const double* __restrict__ sc,
                                          const DNode& c1, int k2, int k3, int m,
                                          const C3Params& P, double* __restrict__ v) {
    const double w1 = c1.w, n1 = double(c1.n);
    constexpr int NV = zeta_nv(D);
    double* two = v + NV - 5;
    // Corr3::calculateZeta, src/Corr3.cpp:1872-1940
    if (k3 == k2) {
        if (m == 0 && !P.quick) {
            const double np2 = sc[Lo.S(S_NPAIRS, k2)], sw2 = sc[Lo.S(S_SUMW, k2)];
            const double swr = sc[Lo.S(S_SUMWR, k2)], swlr = sc[Lo.S(S_SUMWLOGR, k2)];
            const double swwr = sc[Lo.S(S_SUMWWR, k2)], swwlr = sc[Lo.S(S_SUMWWLOGR, k2)];
            two[0] += n1 * np2 * (np2 - 1);
            two[1] += w1 * (sw2 * swr - swwr);
            two[2] += w1 * (sw2 * swlr - swwlr);
            two[3] += w1 * (sw2 * swr - swwr);
            two[4] += w1 * (sw2 * swlr - swwlr);
        }
        const double wr = sc[Lo.W(k2, m)], wi = sc[Lo.W(k2, m) + 1];
        v[0] += w1 * ((wr * wr + wi * wi) - sc[Lo.S(S_SUMWW, k2)]);
    } else {
        if (m == 0 && !P.quick) {
            const double np2 = sc[Lo.S(S_NPAIRS, k2)], sw2 = sc[Lo.S(S_SUMW, k2)];
            const double np3 = sc[Lo.S(S_NPAIRS, k3)], sw3 = sc[Lo.S(S_SUMW, k3)];
            const double ww12 = w1 * sw2, ww13 = w1 * sw3;
            two[0] += n1 * np2 * np3;
            two[1] += ww13 * sc[Lo.S(S_SUMWR, k2)];
            two[2] += ww13 * sc[Lo.S(S_SUMWLOGR, k2)];
            two[3] += ww12 * sc[Lo.S(S_SUMWR, k3)];
            two[4] += ww12 * sc[Lo.S(S_SUMWLOGR, k3)];
        }
        const double ar = w1 * sc[Lo.W(k2, m)], ai = w1 * sc[Lo.W(k2, m) + 1];
        double xr, xi;
        cmul3(ar, ai, sc[Lo.W(k3, m)], -sc[Lo.W(k3, m) + 1], xr, xi);
        v[0] += xr;
        v[1] += xi;
    }
    if (D != GData) return;
    // MultipoleHelper<5>::CalculateZeta, src/Corr3.cpp:2582-2849
    auto Gr = [&](int k, int n) { return sc[Lo.G(k, n)]; };
    auto Gi = [&](int k, int n) { return sc[Lo.G(k, n) + 1]; };
    double* Pv = v + 2;
    double* Qv = v + 10;
    const double g1r = c1.a, g1i = c1.b;
    const double g1cr = g1r, g1ci = -g1i;
    if (k3 == k2) {
        double c0r, c0i, c1r, c1i, c2r, c2i;
        cmul3(g1r, g1i, sc[Lo.GG(0, k2)], sc[Lo.GG(0, k2) + 1], c0r, c0i);
        cmul3(g1cr, g1ci, sc[Lo.GG(1, k2)], sc[Lo.GG(1, k2) + 1], c1r, c1i);
        cmul3(g1r, g1i, sc[Lo.GG(2, k2)], sc[Lo.GG(2, k2) + 1], c2r, c2i);
        if (m == 0) {
            double ar, ai, br, bi, tr, ti;
            cmul3(g1r, g1i, Gr(k2, -1), Gi(k2, -1), ar, ai);
            cmul3(g1cr, g1ci, Gr(k2, 1), Gi(k2, 1), br, bi);
            cmul3(ar, ai, Gr(k2, -1), Gi(k2, -1), tr, ti); Pv[0] += tr - c0r; Pv[1] += ti - c0i;
            cmul3(br, bi, Gr(k2, 1), Gi(k2, 1), tr, ti); Pv[2] += tr - c1r; Pv[3] += ti - c1i;
            cmul3(ar, ai, Gr(k2, 1), -Gi(k2, 1), tr, ti);
            Pv[4] += tr - c2r; Pv[5] += ti - c2i;
            Pv[6] += tr - c2r; Pv[7] += ti - c2i;
        } else {
            const double ppr = Gr(k2, m + 1), ppi = Gi(k2, m + 1);
            const double pmr = Gr(k2, m - 1), pmi = Gi(k2, m - 1);
            const double mpr = Gr(k2, -m + 1), mpi = Gi(k2, -m + 1);
            const double mmr = Gr(k2, -m - 1), mmi = Gi(k2, -m - 1);
            double ar, ai, br, bi, tr, ti;
            cmul3(g1r, g1i, mmr, mmi, ar, ai);
            cmul3(g1cr, g1ci, ppr, ppi, br, bi);
            cmul3(ar, ai, pmr, pmi, tr, ti); Pv[0] += tr - c0r; Pv[1] += ti - c0i;
            cmul3(br, bi, mpr, mpi, tr, ti); Pv[2] += tr - c1r; Pv[3] += ti - c1i;
            cmul3(br, -bi, pmr, pmi, tr, ti); Pv[4] += tr - c2r; Pv[5] += ti - c2i;
            cmul3(ar, ai, mpr, -mpi, tr, ti); Pv[6] += tr - c2r; Pv[7] += ti - c2i;
        }
    } else if (m == 0) {
        double ar, ai, br, bi, tr, ti;
        cmul3(g1r, g1i, Gr(k2, -1), Gi(k2, -1), ar, ai);
        cmul3(g1cr, g1ci, Gr(k2, 1), Gi(k2, 1), br, bi);
        cmul3(ar, ai, Gr(k3, -1), Gi(k3, -1), tr, ti); Pv[0] += tr; Pv[1] += ti;
        cmul3(br, bi, Gr(k3, 1), Gi(k3, 1), tr, ti); Pv[2] += tr; Pv[3] += ti;
        cmul3(ar, ai, Gr(k3, 1), -Gi(k3, 1), tr, ti); Pv[4] += tr; Pv[5] += ti;
        cmul3(br, -bi, Gr(k3, -1), Gi(k3, -1), tr, ti); Pv[6] += tr; Pv[7] += ti;
    } else {
        const double G2ppr = Gr(k2, m + 1), G2ppi = Gi(k2, m + 1);
        const double G2pmr = Gr(k2, m - 1), G2pmi = Gi(k2, m - 1);
        const double G2mpr = Gr(k2, -m + 1), G2mpi = Gi(k2, -m + 1);
        const double G2mmr = Gr(k2, -m - 1), G2mmi = Gi(k2, -m - 1);
        const double G3ppr = Gr(k3, m + 1), G3ppi = Gi(k3, m + 1);
        const double G3pmr = Gr(k3, m - 1), G3pmi = Gi(k3, m - 1);
        const double G3mpr = Gr(k3, -m + 1), G3mpi = Gi(k3, -m + 1);
        const double G3mmr = Gr(k3, -m - 1), G3mmi = Gi(k3, -m - 1);
        double pmR, pmI, cpR, cpI, mcR, mcI, mmR, mmI, tr, ti;
        cmul3(g1r, g1i, G2pmr, G2pmi, pmR, pmI);
        cmul3(g1cr, g1ci, G2ppr, G2ppi, cpR, cpI);
        cmul3(g1r, g1i, G2mpr, -G2mpi, mcR, mcI);
        cmul3(g1r, g1i, G2mmr, G2mmi, mmR, mmI);
        cmul3(pmR, pmI, G3mmr, G3mmi, tr, ti); Pv[0] += tr; Pv[1] += ti;
        cmul3(cpR, cpI, G3mpr, G3mpi, tr, ti); Pv[2] += tr; Pv[3] += ti;
        cmul3(pmR, pmI, G3ppr, -G3ppi, tr, ti); Pv[4] += tr; Pv[5] += ti;
        cmul3(mcR, mcI, G3mmr, G3mmi, tr, ti); Pv[6] += tr; Pv[7] += ti;
        cmul3(mmR, mmI, G3pmr, G3pmi, tr, ti); Qv[0] += tr; Qv[1] += ti;
        cmul3(mcR, -mcI, G3ppr, G3ppi, tr, ti); Qv[2] += tr; Qv[3] += ti;
        cmul3(cpR, -cpI, G3pmr, G3pmi, tr, ti); Qv[4] += tr; Qv[5] += ti;
        cmul3(mmR, mmI, G3mpr, -G3mpi, tr, ti); Qv[6] += tr; Qv[7] += ti;
    }
}

constexpr int ZT_THREADS = 512;
constexpr int ZT_MR = 4;            // multipoles m per workgroup, at most
constexpr int ZT_IPT = 2;           // cells per thread: mr * nbins (nbins + 1) / 2 <= ZT_IPT * ZT_THREADS

// One workgroup per (job chunk, range of mr multipoles m).  The chunk's jobs are
// neighbours in task order, so their row ranges [ks, mk) are alike: the workgroup takes
// the union [kmin, kmax) of its jobs' rows and every thread owns up to two cells
// (k2, k3 >= k2, m) of those rows, their sums kept in registers.  For each job, the part
// of its scratch these multipoles read (W(k, m), G(k, m-1..m+1), G(k, -m-1..-m+1) for
// every bin, and the per-bin sums) is copied from the slot into LDS, at the slot's own
// offsets; the cells whose row the job contracts (ks <= k2 < mk) add their terms.  The
// sums are written to the chunk's partial array and the rows to chunk_rows.
template <int D>
__global__ __launch_bounds__(ZT_THREADS) void k_mp_zeta(
    const MPJob* __restrict__ jobs, int njobs, const DNode* __restrict__ nodes,
    double* const* __restrict__ slots, C3Params P, int mr, int jpc, double* __restrict__ partial,
    RowRange* __restrict__ chunk_rows, int nq, int nsub, int diag_only) {
    HIP_DYNAMIC_SHARED(double, sc)
    __shared__ int s_kmin[ZT_THREADS / 64], s_kmax[ZT_THREADS / 64];
    constexpr int NV = zeta_nv(D);
    int q, sub;
    if (!xcd_chunk(nq, nsub, q, sub)) return;
    const int m0 = sub * mr;
    const int nb = P.nbins, M = P.maxn;
    ScratchLayout Lo;
    Lo.init(nb, M, D);
    const int j0 = q * jpc;
    const int j1 = njobs < j0 + jpc ? njobs : j0 + jpc;
    // the union of the chunk's row ranges
    int kmin = nb, kmax = 0;
    for (int j = j0 + (int)threadIdx.x; j < j1; j += blockDim.x) {
        const MPJob jb = jobs[j];
        if (jb.c1 >= 0 && jb.ks < jb.mk) {
            kmin = jb.ks < kmin ? jb.ks : kmin;
            kmax = jb.mk > kmax ? jb.mk : kmax;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(kmin, o, 64), b = __shfl_xor(kmax, o, 64);
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
    }
    if ((threadIdx.x & 63) == 0) { s_kmin[threadIdx.x >> 6] = kmin; s_kmax[threadIdx.x >> 6] = kmax; }
    __syncthreads();
    kmin = nb; kmax = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        kmin = s_kmin[w] < kmin ? s_kmin[w] : kmin;
        kmax = s_kmax[w] > kmax ? s_kmax[w] : kmax;
    }
    if (threadIdx.x == 0 && sub == 0) chunk_rows[q] = RowRange{kmin, kmax};
    if (kmin >= kmax) return;            // no rows in this chunk (uniform)
    // this thread's cells: c = threadIdx.x + u * ZT_THREADS, m fastest, then k3, then k2
    int ck2[ZT_IPT], ck3[ZT_IPT], cm[ZT_IPT];
#pragma unroll
    for (int u = 0; u < ZT_IPT; ++u) {
        const int c = (int)threadIdx.x + u * ZT_THREADS;
        const int ml = c % mr;
        int pr = c / mr;
        int k2 = kmin;
        while (k2 < kmax && pr >= nb - k2) { pr -= nb - k2; ++k2; }
        ck2[u] = k2;
        ck3[u] = k2 + pr;
        cm[u] = (k2 < kmax && m0 + ml <= M) ? m0 + ml : -1;
        // with the contraction on the matrix cores, this kernel keeps only the n = 0 sums
        // (ntri, meand2, ...), which stay FP64 (mode 2: launched for m = 0 alone)
        if (diag_only && cm[u] > 0) cm[u] = -1;
    }
    double a0[NV], a1[NV];
    static_assert(ZT_IPT == 2, "two cells per thread");
#pragma unroll
    for (int v = 0; v < NV; ++v) { a0[v] = 0.; a1[v] = 0.; }
    // the scratch entries these multipoles read: per bin W(k, m0..mhi), and for GGG
    // G(k, gl..mhi+1) and G(k, -(mhi+1)..-gl); then all per-bin sums
    const int mhi = (m0 + mr - 1 < M ? m0 + mr - 1 : M);
    const int nw = 2 * (mhi - m0 + 1);
    const int gl = m0 - 1 < 0 ? 0 : m0 - 1;
    const int ng = D == GData ? 2 * (mhi + 1 - gl + 1) : 0;
    const int per_k = nw + 2 * ng;
    const int nsum = Lo.total - Lo.sums;
    const int ncopy = nb * per_k + nsum;
    for (int j = j0; j < j1; ++j) {
        const MPJob jb = jobs[j];
        if (jb.c1 < 0 || jb.ks >= jb.mk) continue;
        const double* gsl = slot_at(slots, jb.slot, Lo.total);
        __syncthreads();                 // the previous job's entries are no longer read
        for (int e = threadIdx.x; e < ncopy; e += blockDim.x) {
            int off;
            if (e >= nb * per_k) {
                off = Lo.sums + (e - nb * per_k);
            } else {
                const int k = e / per_k;
                const int r = e - k * per_k;
                if (r < nw) off = Lo.W(k, m0) + r;
                else if (r < nw + ng) off = Lo.G(k, gl) + (r - nw);
                else off = Lo.G(k, -(mhi + 1)) + (r - nw - ng);
            }
            sc[off] = gsl[off];
        }
        __syncthreads();
        const DNode c1 = nodes[jb.c1];
        if (cm[0] >= 0 && ck2[0] >= jb.ks && ck2[0] < jb.mk)
            zeta_cell<D>(Lo, sc, c1, ck2[0], ck3[0], cm[0], P, a0);
        if (cm[1] >= 0 && ck2[1] >= jb.ks && ck2[1] < jb.mk)
            zeta_cell<D>(Lo, sc, c1, ck2[1], ck3[1], cm[1], P, a1);
    }
    const size_t ncomp_chunk = (size_t)nb * nb * (M + 1) * NV;
    if (cm[0] >= 0) {
        double* dst = partial + (size_t)q * ncomp_chunk +
                      (((size_t)ck2[0] * nb + ck3[0]) * (M + 1) + cm[0]) * NV;
        const int v0 = diag_only ? NV - 5 : 0;
#pragma unroll
        for (int v = 0; v < NV; ++v) if (v >= v0) dst[v] = a0[v];
    }
    if (cm[1] >= 0) {
        double* dst = partial + (size_t)q * ncomp_chunk +
                      (((size_t)ck2[1] * nb + ck3[1]) * (M + 1) + cm[1]) * NV;
        const int v0 = diag_only ? NV - 5 : 0;
#pragma unroll
        for (int v = 0; v < NV; ++v) if (v >= v0) dst[v] = a1[v];
    }
}

// ---------------------------------------------------------------------------
// Phase 3, two lists: zeta of the jobs of one level (calculateZeta with mp2, mp3,
// src/Corr3.cpp:1976-2085; MultipoleHelper<5>, :2851-3003)
// ---------------------------------------------------------------------------
// There is no symmetry between the lists, so the outputs are kept per (k2, k3, m >= 0)
// for every k3, NV doubles: the weight at n = +m (n = -m is its conjugate), for GGG the
// four gam at n = +m [2,9] and at n = -m [10,17], then ntri, meand2, meanlogd2, meand3,
// meanlogd3 at n = 0 (m = 0 only).  s2, s3: the first and second lists' scratches, which
// hold the sums of the cells across from vertex 2 and vertex 3 respectively.
template <int D>
__device__ __forceinline__ void zeta_cell_x(const ScratchLayout& Lo, const double* __restrict__ s2,
                                            const double* __restrict__ s3, const DNode& c1, int k2,
                                            int k3, int m, const C3Params& P,
                                            double* __restrict__ v) {
    constexpr int NV = zeta_nv(D);
    const double w1 = c1.w, n1 = double(c1.n);
    const bool swap23 = P.order3 == 1;
    double* two = v + NV - 5;
    if (m == 0 && !P.quick) {
        auto S2 = [&](int q, int k) { return s2[Lo.S(q, k)]; };
        auto S3 = [&](int q, int k) { return s3[Lo.S(q, k)]; };
        if (!swap23) {
            two[0] += n1 * S3(S_NPAIRS, k2) * S2(S_NPAIRS, k3);
            const double ww12 = w1 * S3(S_SUMW, k2), ww13 = w1 * S2(S_SUMW, k3);
            two[1] += ww13 * S3(S_SUMWR, k2);
            two[2] += ww13 * S3(S_SUMWLOGR, k2);
            two[3] += ww12 * S2(S_SUMWR, k3);
            two[4] += ww12 * S2(S_SUMWLOGR, k3);
        } else {
            two[0] += n1 * (S2(S_NPAIRS, k2) * S3(S_NPAIRS, k3) + S3(S_NPAIRS, k2) * S2(S_NPAIRS, k3));
            two[1] += w1 * (S2(S_SUMW, k3) * S3(S_SUMWR, k2) + S3(S_SUMW, k3) * S2(S_SUMWR, k2));
            two[2] += w1 * (S2(S_SUMW, k3) * S3(S_SUMWLOGR, k2) + S3(S_SUMW, k3) * S2(S_SUMWLOGR, k2));
            two[3] += w1 * (S2(S_SUMW, k2) * S3(S_SUMWR, k3) + S3(S_SUMW, k2) * S2(S_SUMWR, k3));
            // the second term pairs sumw and sumwlogr of bin k3, as the reference does
            two[4] += w1 * (S2(S_SUMW, k2) * S3(S_SUMWLOGR, k3) + S3(S_SUMW, k3) * S2(S_SUMWLOGR, k3));
        }
    }
    // weight: w1 W3(k2, m) conj(W2(k3, m)) (+ w1 W2(k2, m) conj(W3(k3, m)))
    {
        double xr, xi;
        cmul3(w1 * s3[Lo.W(k2, m)], w1 * s3[Lo.W(k2, m) + 1], s2[Lo.W(k3, m)], -s2[Lo.W(k3, m) + 1],
              xr, xi);
        v[0] += xr;
        v[1] += xi;
        if (swap23) {
            cmul3(w1 * s2[Lo.W(k2, m)], w1 * s2[Lo.W(k2, m) + 1], s3[Lo.W(k3, m)],
                  -s3[Lo.W(k3, m) + 1], xr, xi);
            v[0] += xr;
            v[1] += xi;
        }
    }
    if (D != GData) return;
    const double g1r = c1.a, g1i = c1.b;
    // gam_mu(n) from A's G at bin k2 and B's G at bin k3 (the general form of :2883-2897)
    auto gam = [&](const double* A, const double* B, int n, double* g) {
        double ar, ai, tr, ti;
        const int a_nm = Lo.G(k2, n - 1), a_np = Lo.G(k2, n + 1), a_mnp = Lo.G(k2, -n + 1);
        const int b_mnm = Lo.G(k3, -n - 1), b_mnp = Lo.G(k3, -n + 1), b_np = Lo.G(k3, n + 1);
        cmul3(g1r, g1i, A[a_nm], A[a_nm + 1], ar, ai);
        cmul3(ar, ai, B[b_mnm], B[b_mnm + 1], tr, ti); g[0] += tr; g[1] += ti;
        cmul3(ar, ai, B[b_np], -B[b_np + 1], tr, ti); g[4] += tr; g[5] += ti;
        cmul3(g1r, -g1i, A[a_np], A[a_np + 1], ar, ai);
        cmul3(ar, ai, B[b_mnp], B[b_mnp + 1], tr, ti); g[2] += tr; g[3] += ti;
        cmul3(g1r, g1i, A[a_mnp], -A[a_mnp + 1], ar, ai);
        cmul3(ar, ai, B[b_mnm], B[b_mnm + 1], tr, ti); g[6] += tr; g[7] += ti;
    };
    gam(s3, s2, m, v + 2);
    if (swap23) gam(s2, s3, m, v + 2);
    if (m > 0) {
        gam(s3, s2, -m, v + 10);
        if (swap23) gam(s2, s3, -m, v + 10);
    }
}

// One workgroup per (job chunk, range of mr multipoles, group of kg rows k2).  The
// chunk's rows are k2 >= kmin, the smallest ks of its jobs; each thread owns up to
// ZT_IPT cells (k2, k3, m) of its group's rows (every k3), sums in registers.  For each
// job, the entries of both scratches that these multipoles read are copied into LDS
// (the first list's at the slot's offsets, the second's after them); a cell takes the
// job's terms when k2, k3 >= ks and not both >= mk.  Every cell of the group's rows
// k2 >= kmin is written to the chunk's partial array (zero where no job contributes).
template <int D>
__global__ __launch_bounds__(ZT_THREADS) void k_mp_zeta_x(
    const MPJob* __restrict__ jobs, int njobs, const DNode* __restrict__ nodes,
    double* const* __restrict__ slots, C3Params P, int mr, int kg, int jpc,
    double* __restrict__ partial, RowRange* __restrict__ chunk_rows, int nq, int nsub) {
    HIP_DYNAMIC_SHARED(double, sc)
    __shared__ int s_kmin[ZT_THREADS / 64];
    constexpr int NV = zeta_nv(D);
    int q, sub;
    if (!xcd_chunk(nq, nsub, q, sub)) return;
    const int nb = P.nbins, M = P.maxn;
    const int nmr = (M + 1 + mr - 1) / mr;
    const int m0 = (sub % nmr) * mr;
    const int g0 = (sub / nmr) * kg;
    ScratchLayout Lo;
    Lo.init(nb, M, D);
    const int j0 = q * jpc;
    const int j1 = njobs < j0 + jpc ? njobs : j0 + jpc;
    int kmin = nb;
    for (int j = j0 + (int)threadIdx.x; j < j1; j += blockDim.x) {
        const MPJob jb = jobs[j];
        if (jb.c1 >= 0 && jb.ks < jb.mk) kmin = jb.ks < kmin ? jb.ks : kmin;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int a = __shfl_xor(kmin, o, 64);
        kmin = a < kmin ? a : kmin;
    }
    if ((threadIdx.x & 63) == 0) s_kmin[threadIdx.x >> 6] = kmin;
    __syncthreads();
    kmin = nb;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) kmin = s_kmin[w] < kmin ? s_kmin[w] : kmin;
    if (threadIdx.x == 0 && sub == 0) chunk_rows[q] = RowRange{kmin, nb};
    if (kmin >= nb) return;              // no jobs in this chunk (uniform)
    int ck2[ZT_IPT], ck3[ZT_IPT], cm[ZT_IPT];
#pragma unroll
    for (int u = 0; u < ZT_IPT; ++u) {
        const int c = (int)threadIdx.x + u * ZT_THREADS;
        const int ml = c % mr;
        const int pr = c / mr;
        const int k2 = g0 + pr / nb;
        ck2[u] = k2;
        ck3[u] = pr % nb;
        cm[u] = (k2 < nb && k2 < g0 + kg && k2 >= kmin && m0 + ml <= M) ? m0 + ml : -1;
    }
    double a0[NV], a1[NV];
    static_assert(ZT_IPT == 2, "two cells per thread");
#pragma unroll
    for (int v = 0; v < NV; ++v) { a0[v] = 0.; a1[v] = 0.; }
    const int mhi = (m0 + mr - 1 < M ? m0 + mr - 1 : M);
    const int nw = 2 * (mhi - m0 + 1);
    const int gl = m0 - 1 < 0 ? 0 : m0 - 1;
    const int ng = D == GData ? 2 * (mhi + 1 - gl + 1) : 0;
    const int per_k = nw + 2 * ng;
    const int nsum = Lo.total - Lo.sums;
    const int ncopy = nb * per_k + nsum;
    for (int j = j0; j < j1; ++j) {
        const MPJob jb = jobs[j];
        if (jb.c1 < 0 || jb.ks >= jb.mk) continue;
        const double* gsl = slot_at(slots, jb.slot, 2 * Lo.total);
        __syncthreads();                 // the previous job's entries are no longer read
        for (int e = threadIdx.x; e < 2 * ncopy; e += blockDim.x) {
            const int half = e >= ncopy ? 1 : 0;
            const int ee = e - half * ncopy;
            int off;
            if (ee >= nb * per_k) {
                off = Lo.sums + (ee - nb * per_k);
            } else {
                const int k = ee / per_k;
                const int r = ee - k * per_k;
                if (r < nw) off = Lo.W(k, m0) + r;
                else if (r < nw + ng) off = Lo.G(k, gl) + (r - nw);
                else off = Lo.G(k, -(mhi + 1)) + (r - nw - ng);
            }
            off += half * Lo.total;
            sc[off] = gsl[off];
        }
        __syncthreads();
        const DNode c1 = nodes[jb.c1];
        auto takes = [&](int k2, int k3) {
            return k2 >= jb.ks && k3 >= jb.ks && (k2 < jb.mk || k3 < jb.mk);
        };
        if (cm[0] >= 0 && takes(ck2[0], ck3[0]))
            zeta_cell_x<D>(Lo, sc, sc + Lo.total, c1, ck2[0], ck3[0], cm[0], P, a0);
        if (cm[1] >= 0 && takes(ck2[1], ck3[1]))
            zeta_cell_x<D>(Lo, sc, sc + Lo.total, c1, ck2[1], ck3[1], cm[1], P, a1);
    }
    const size_t ncomp_chunk = (size_t)nb * nb * (M + 1) * NV;
    if (cm[0] >= 0) {
        double* dst = partial + (size_t)q * ncomp_chunk +
                      (((size_t)ck2[0] * nb + ck3[0]) * (M + 1) + cm[0]) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) dst[v] = a0[v];
    }
    if (cm[1] >= 0) {
        double* dst = partial + (size_t)q * ncomp_chunk +
                      (((size_t)ck2[1] * nb + ck3[1]) * (M + 1) + cm[1]) * NV;
#pragma unroll
        for (int v = 0; v < NV; ++v) dst[v] = a1[v];
    }
}

// total += sum over the chunks (in chunk order) of the partial sums of the chunks whose
// rows include this entry's row k2, compensated
// (full: the two-list layout, every k3)
__global__ void k_mp_zeta_reduce(const double* __restrict__ partial,
                                 const RowRange* __restrict__ chunk_rows, int nchunks, long long n,
                                 long long row_stride, int nbins, double* __restrict__ hi,
                                 double* __restrict__ lo, int full) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int k2 = (int)(e / row_stride);
    const int k3 = (int)((e % row_stride) / (row_stride / nbins));
    if (!full && k3 < k2) return;        // not a cell of the compressed layout
    double s = hi[e], c = lo[e];
    for (int q = 0; q < nchunks; ++q) {
        const RowRange rows = chunk_rows[q];
        if (k2 < rows.lo || k2 >= rows.hi) continue;
        const double v = partial[(size_t)q * n + e];
        const double tt = s + v;
        c += (fabs(s) >= fabs(v)) ? (s - tt) + v : (v - tt) + s;
        s = tt;
    }
    hi[e] = s;
    lo[e] = c;
}

// The compressed sums -> TreeCorr's arrays zeta0..7, meand2, meanlogd2, meand3,
// meanlogd3, weight, weight_im, ntri at index (a*nb + b)*(2M+1) + M + n.
template <int D>
__global__ void k_mp_expand(const double* __restrict__ hi, const double* __restrict__ lo,
                            C3Params P, double* __restrict__ out) {
    constexpr int NV = zeta_nv(D);
    const int nb = P.nbins, M = P.maxn, nn = 2 * M + 1;
    const long long ntot = (long long)nb * nb * nn;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntot) return;
    const int a = (int)(idx / ((long long)nb * nn));
    const int b = (int)((idx / nn) % nb);
    const int n = (int)(idx % nn) - M;
    const int m = n >= 0 ? n : -n;
    const int k2 = a <= b ? a : b, k3 = a <= b ? b : a;
    const long long cb = (((long long)k2 * nb + k3) * (M + 1) + m) * NV;
    auto V = [&](int i) { return hi[cb + i] + lo[cb + i]; };
    const long long cb0 = (((long long)k2 * nb + k3) * (M + 1)) * NV;
    auto V0 = [&](int i) { return hi[cb0 + i] + lo[cb0 + i]; };
    // weight, weight_im (src/Corr3.cpp:1917-1938)
    double wre = V(0), wim = V(1);
    if (a == b) wim = 0.;
    else if (a < b) wim = n >= 0 ? wim : -wim;
    else wim = n >= 0 ? -wim : wim;
    out[O_WEIGHT * ntot + idx] = wre;
    out[O_WEIGHT_IM * ntot + idx] = wim;
    // ntri, meand2, meanlogd2, meand3, meanlogd3 at n = 0 (:1877-1907)
    double t5[5] = {0., 0., 0., 0., 0.};
    if (n == 0) {
        for (int i = 0; i < 5; ++i) t5[i] = V0(NV - 5 + i);
        if (a > b) {   // the (k3, k2) entries: d2 and d3 exchanged
            double x = t5[1]; t5[1] = t5[3]; t5[3] = x;
            x = t5[2]; t5[2] = t5[4]; t5[4] = x;
        }
    }
    out[O_NTRI * ntot + idx] = t5[0];
    out[O_MEAND2 * ntot + idx] = t5[1];
    out[O_MEANLOGD2 * ntot + idx] = t5[2];
    out[O_MEAND3 * ntot + idx] = t5[3];
    out[O_MEANLOGD3 * ntot + idx] = t5[4];
    if (D != GData) return;
    // zeta0..3 as (re, im) pairs; perm exchanges zeta2 and zeta3
    double z[8];
    bool perm = false;
    int base = 2;            // P
    if (a == b) {
        perm = n < 0;
    } else if (a < b) {
        if (n < 0) { base = 10; perm = true; }
    } else {
        if (n > 0) base = 10;
        else perm = true;    // n = 0: P with zeta2, zeta3 exchanged; n < 0 as well
    }
    for (int i = 0; i < 8; ++i) z[i] = V(base + i);
    if (perm) {
        double x = z[4]; z[4] = z[6]; z[6] = x;
        x = z[5]; z[5] = z[7]; z[7] = x;
    }
    for (int i = 0; i < 8; ++i) out[(O_Z0 + i) * ntot + idx] = z[i];
}

// The two-list sums -> TreeCorr's arrays (index (k2*nb + k3)*(2M+1) + M + n)
template <int D>
__global__ void k_mp_expand_x(const double* __restrict__ hi, const double* __restrict__ lo,
                              C3Params P, double* __restrict__ out) {
    constexpr int NV = zeta_nv(D);
    const int nb = P.nbins, M = P.maxn, nn = 2 * M + 1;
    const long long ntot = (long long)nb * nb * nn;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ntot) return;
    const int k2 = (int)(idx / ((long long)nb * nn));
    const int k3 = (int)((idx / nn) % nb);
    const int n = (int)(idx % nn) - M;
    const int m = n >= 0 ? n : -n;
    const long long cb = (((long long)k2 * nb + k3) * (M + 1) + m) * NV;
    auto V = [&](int i) { return hi[cb + i] + lo[cb + i]; };
    out[O_WEIGHT * ntot + idx] = V(0);
    out[O_WEIGHT_IM * ntot + idx] = n >= 0 ? V(1) : -V(1);
    const int oi[5] = {O_NTRI, O_MEAND2, O_MEANLOGD2, O_MEAND3, O_MEANLOGD3};
    for (int i = 0; i < 5; ++i) out[oi[i] * ntot + idx] = n == 0 ? V(NV - 5 + i) : 0.;
    if (D != GData) return;
    const int base = n >= 0 ? 2 : 10;
    for (int i = 0; i < 8; ++i) out[(O_Z0 + i) * ntot + idx] = V(base + i);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;       // elements
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    // A block goes back to the caching allocator only once no queued kernel can still
    // use it (the allocator hands it out again at once).
    ~DBuf() { release(); }
    void release() {
        if (p) {
            (void)hipStreamSynchronize(stream());
            dev_free(p, bytes);
        }
        p = nullptr; n = 0; bytes = 0;
    }
    int alloc(size_t count) {
        release();
        bytes = sizeof(T) * std::max<size_t>(count, 1);
        p = static_cast<T*>(dev_alloc(bytes));
        n = p ? std::max<size_t>(count, 1) : 0;
        if (!p) { bytes = 0; return fail("out of device memory in the multipole engine"); }
        return 0;
    }
    // at least `count` elements; the first `keep` are preserved
    int reserve(size_t count, size_t keep, hipStream_t st) {
        if (count <= n) return 0;
        const size_t want = std::max(count, n + n / 2);
        T* q = static_cast<T*>(dev_alloc(sizeof(T) * want));
        if (!q) return fail("out of device memory in the multipole engine");
        if (p && keep) TC_HIP(hipMemcpyAsync(q, p, sizeof(T) * std::min(keep, n), hipMemcpyDeviceToDevice, st));
        if (p) {
            TC_HIP(hipStreamSynchronize(st));   // the copy has read the old block
            dev_free(p, bytes);
        }
        p = q; n = want; bytes = sizeof(T) * want;
        (void)keep;
        return 0;
    }
    void swap(DBuf& o) { std::swap(p, o.p); std::swap(n, o.n); std::swap(bytes, o.bytes); }
};

using SplitFn = void (*)(const DNode*, int, const SplitTask*, int*, unsigned long long*,
                         unsigned long long, C3Params, SplitTask*, SplitTask*, int*);
using StageFn = void (*)(const DNode*, int, const NodeSum*, const MPTask*, const int*,
                         unsigned long long, int*, unsigned long long*, unsigned long long,
                         double* const*, int, C3Params, MPTask*, MPJob*, int*,
                         unsigned long long*, int*);
using ZetaFn = void (*)(const MPJob*, int, const DNode*, double* const*, C3Params, int, int, double*,
                        RowRange*, int, int, int);
using ExpandFn = void (*)(const double*, const double*, C3Params, double*);

static SplitFn pick_split(int C) {
    if (C == Flat) return &k_mp_split_level<Flat>;
    if (C == ThreeD) return &k_mp_split_level<ThreeD>;
    return &k_mp_split_level<Sphere>;
}
template <bool X>
static StageFn pick_stage_t(int D, int C) {
#define TC_SF(DD, CC) return &k_mp_stage<DD, CC, X>
    if (D == NData) { if (C == Flat) TC_SF(NData, Flat); if (C == ThreeD) TC_SF(NData, ThreeD); TC_SF(NData, Sphere); }
    if (C == Flat) TC_SF(GData, Flat); if (C == ThreeD) TC_SF(GData, ThreeD); TC_SF(GData, Sphere);
#undef TC_SF
}
static StageFn pick_stage(int D, int C, bool X) {
    return X ? pick_stage_t<true>(D, C) : pick_stage_t<false>(D, C);
}
using ZetaXFn = void (*)(const MPJob*, int, const DNode*, double* const*, C3Params, int, int, int,
                         double*, RowRange*, int, int);

static int set_err_bits(const char* phase, int bits) {
    std::string m = std::string(phase) + " stopped:";
    if (bits & E_POOL) m += " list pool overflow;";
    if (bits & E_INDEX) m += " inconsistent tree or list data;";
    if (bits & E_SLOT) m += " inconsistent stage record;";
    return fail(m);
}

// Scan compaction of n fixed-position records into out[out_base ...]; the count and the
// sum of their list lengths land in totals[0], totals[1] (device, read later).
template <class R>
struct Compactor {
    DBuf<int> bcount, boff;
    DBuf<unsigned long long> blen;
    int run(const R* in, long long n, R* out, long long out_base, unsigned long long* totals,
            hipStream_t st) {
        const int nblk = (int)std::max<long long>(1, (n + CP_BLOCK - 1) / CP_BLOCK);
        if (bcount.reserve(nblk, 0, st) || boff.reserve(nblk, 0, st) || blen.reserve(nblk, 0, st))
            return 2;
        if (n > 0)
            hipLaunchKernelGGL(k_cp_count<R>, dim3(nblk), dim3(CP_TPB), 0, st, in, n, bcount.p,
                               blen.p);
        else {
            TC_HIP(hipMemsetAsync(bcount.p, 0, sizeof(int), st));
            TC_HIP(hipMemsetAsync(blen.p, 0, sizeof(unsigned long long), st));
        }
        TC_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_cp_scan, dim3(1), dim3(1024), 0, st, bcount.p, blen.p, nblk, boff.p,
                           totals);
        TC_HIP(hipGetLastError());
        if (n > 0) {
            hipLaunchKernelGGL(k_cp_scatter<R>, dim3(nblk), dim3(CP_TPB), 0, st, in, n, boff.p,
                               out, out_base);
            TC_HIP(hipGetLastError());
        }
        return 0;
    }
};

static std::atomic<int> g_zeta_mode{0};
int set_zeta_mode(int mode) {
    if (mode < 0) return g_zeta_mode.load();
    if (mode > 2) return fail("zeta mode must be 0 (FP64), 1 (FP32 MFMA) or 2 (split-BF16 MFMA)");
    g_zeta_mode.store(mode);
    return 0;
}

static std::atomic<int> g_mp_plain{0};
int set_multipole_plain(int on) {
    if (on < 0) return g_mp_plain.load();
    g_mp_plain.store(on ? 1 : 0);
    return 0;
}

// Runs the multipole auto-correlation of one device tree; writes NOUT3*ntot reduced
// sums (zeta0..7, meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri) into
// out (host) and the work counters into cnt (NC3 = 5 values: Gn calls, zeta calls, zeta
// rows, stages, zeta cells).
int corr3_run(int D, int C, const DNode* nodes, int nnodes, const int* dtop, int ntop,
              const C3Params& P_in, double* out, unsigned long long* cnt) {
    C3Params P = P_in;
    P.order3 = 0;
    return corr3_run_x(D, C, nodes, nnodes, dtop, ntop, dtop, ntop, P, out, cnt);
}

int corr3_run_x(int D, int C, const DNode* nodes, int nnodes, const int* c1top, int nc1,
                const int* ltop, int nl, const C3Params& P_in, double* out,
                unsigned long long* cnt) {
    hipStream_t st = stream();
    C3Params P = P_in;
    P.plain = g_mp_plain.load();
    const bool X = P.order3 != 0;
    if (X && P.order3 != 1 && P.order3 != 4) return fail("two-list multipole: ordered must be 1 or 4");
    const int nb = P.nbins, M = P.maxn;
    const long long ntot = (long long)nb * nb * (2 * M + 1);
    ScratchLayout Lo;
    Lo.init(nb, M, D);
    const int stot = X ? 2 * Lo.total : Lo.total;        // doubles per scratch slot
    const size_t shbytes = sizeof(double) * stot;
    if (shbytes > 150 * 1024) return fail("nbins and max_n too large for the LDS scratch");
    if (nc1 <= 0 || nl <= 0 || nnodes <= 0) return fail("empty field");
    const int ntop = nc1;
    StageFn sfn = pick_stage(D, C, X);
    TC_HIP(hipFuncSetAttribute((const void*)sfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shbytes));
    const int NV = zeta_nv(D);
    const long long ncell = (long long)nb * (M + 1);
    const long long npair = (long long)nb * (nb + 1) / 2;
    const int zmr = (int)std::min<long long>(ZT_MR, (long long)ZT_IPT * ZT_THREADS / npair);
    if (zmr < 1) return fail("nbins too large for the zeta kernel's cells (at most 45 bins)");
    const int nmr = (M + 1 + zmr - 1) / zmr;
    ZetaFn zfn = D == GData ? &k_mp_zeta<GData> : &k_mp_zeta<NData>;
    const size_t zt_bytes = shbytes;    // the slot's layout, partly filled per job
    TC_HIP(hipFuncSetAttribute((const void*)zfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)zt_bytes));
    // two lists: cells (k2, k3, m) of every k3, groups of kg rows and xmr multipoles
    ZetaXFn zxfn = D == GData ? &k_mp_zeta_x<GData> : &k_mp_zeta_x<NData>;
    const int xcells = ZT_IPT * ZT_THREADS;
    const int xmr = std::max(1, std::min(ZT_MR, xcells / (nb * nb)));
    const int xkg = std::min(nb, xcells / (nb * xmr));
    if (X && xkg < 1) return fail("nbins too large for the two-list zeta kernel (at most 1024 bins)");
    const int xsub = X ? ((M + 1 + xmr - 1) / xmr) * ((nb + xkg - 1) / xkg) : 0;
    if (X) TC_HIP(hipFuncSetAttribute((const void*)zxfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)shbytes));
    // the off-diagonal zeta cells on the matrix cores (at most 32 bins; one list only)
    const int zmode = (nb <= 32 && !X) ? g_zeta_mode.load() : 0;
    int mf_threads = 0;
    size_t mf_bytes = 0;
    ZetaMfmaFn mfn = pick_zeta_mfma(D, zmode, nb, mf_threads, mf_bytes);
    if (mfn) TC_HIP(hipFuncSetAttribute((const void*)mfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)mf_bytes));
    KernelStat& ks_stage = kernel_stat(KS_MULTIPOLE);
    KernelStat& ks_zeta = kernel_stat(KS_ZETA);
    const bool timing = timing_enabled();

    // node sums, bottom up: each launch finishes the nodes whose children are done
    DBuf<NodeSum> nsum;
    DBuf<int> doneA, doneB, changed;
    if (nsum.alloc(nnodes) || doneA.alloc(nnodes) || doneB.alloc(nnodes) || changed.alloc(1))
        return 2;
    TC_HIP(hipMemsetAsync(doneA.p, 0, sizeof(int) * nnodes, st));
    {
        int hchanged = 1;
        int* din = doneA.p;
        int* dout = doneB.p;
        for (int it = 0; it < 4096 && hchanged; ++it) {
            TC_HIP(hipMemsetAsync(changed.p, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_node_sums, dim3((nnodes + 255) / 256), dim3(256), 0, st, nodes,
                               nnodes, nsum.p, din, dout, changed.p);
            TC_HIP(hipGetLastError());
            TC_HIP(hipMemcpyAsync(&hchanged, changed.p, sizeof(int), hipMemcpyDeviceToHost, st));
            TC_HIP(hipStreamSynchronize(st));
            std::swap(din, dout);
        }
        if (hchanged) return fail("node sums did not converge: the tree has a cycle");
    }

    DBuf<int> err;
    DBuf<unsigned long long> counters, tot, ptop;     // tot: 3 compaction totals x 2
    if (err.alloc(1) || counters.alloc(NC3) || tot.alloc(6) || ptop.alloc(1)) return 2;
    TC_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
    TC_HIP(hipMemsetAsync(counters.p, 0, sizeof(unsigned long long) * NC3, st));

    // ---- phase 1: multipoleSplit1, one launch per level ----
    DBuf<int> spool;
    DBuf<SplitTask> stA, stB, schild, sunit_out, units;
    if (spool.alloc((size_t)nl + 2ull * nc1 * nl) || stA.alloc(nc1) || units.alloc(1024))
        return 2;
    hipLaunchKernelGGL(k_split_init, dim3((std::max(nc1, nl) + 255) / 256), dim3(256), 0, st, c1top,
                       nc1, ltop, nl, stA.p, spool.p);
    TC_HIP(hipGetLastError());
    unsigned long long spool_top = (unsigned long long)nl;
    unsigned long long sumlen = (unsigned long long)nc1 * nl;
    long long ntask = ntop, nunits = 0;
    unsigned long long usum = 0;      // the units' list lengths (bounds the first next lists)
    Compactor<SplitTask> cps;
    Compactor<MPTask> cpt;
    Compactor<int> cpi;
    SplitFn splfn = pick_split(C);
    for (int level = 0; ntask > 0; ++level) {
        if (level > 200) return fail("multipole split phase: too many levels");
        const unsigned long long need = spool_top + 2 * sumlen;
        if (need > 0x7fffffffull) return fail("multipole split phase: list pool beyond 2^31 entries");
        if (spool.reserve(need, spool_top, st) || schild.reserve(2 * ntask, 0, st) ||
            sunit_out.reserve(ntask, 0, st) || stB.reserve(2 * ntask, 0, st) ||
            units.reserve(nunits + ntask, nunits, st))
            return 2;
        TC_HIP(hipMemcpyAsync(ptop.p, &spool_top, sizeof(spool_top), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(splfn, dim3((unsigned)ntask), dim3(MP_THREADS), 0, st, nodes, nnodes,
                           stA.p, spool.p, ptop.p, (unsigned long long)spool.n, P, schild.p,
                           sunit_out.p, err.p);
        TC_HIP(hipGetLastError());
        if (cps.run(schild.p, 2 * ntask, stB.p, 0, tot.p, st) ||
            cps.run(sunit_out.p, ntask, units.p, nunits, tot.p + 2, st))
            return 2;
        unsigned long long h[4];
        int herr = 0;
        TC_HIP(hipMemcpyAsync(h, tot.p, sizeof(h), hipMemcpyDeviceToHost, st));
        TC_HIP(hipMemcpyAsync(&spool_top, ptop.p, sizeof(spool_top), hipMemcpyDeviceToHost, st));
        TC_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (herr) return set_err_bits("multipole split phase", herr);
        ntask = (long long)h[0];
        sumlen = h[1];
        nunits += (long long)h[2];
        usum += h[3];
        stA.swap(stB);
    }

    // ---- phase 2+3: multipoleFinish stages and zeta, one launch of each per level ----
    DBuf<MPTask> tq, tq_next, tchild;
    DBuf<MPJob> jobs;
    DBuf<int> ret, free_ids;
    DBuf<int> poolA, poolB;
    DBuf<double> partial, zhi, zlo, red;
    DBuf<RowRange> chunk_rows;
    DBuf<double*> chunk_tab;
    std::vector<double*> chunks;
    struct ChunkGuard {
        std::vector<double*>* c; size_t bytes;
        ~ChunkGuard() {
            (void)hipStreamSynchronize(stream());
            for (double* p : *c) dev_free(p, bytes);
        }
    } cguard{&chunks, sizeof(double) * (size_t)stot << SLOT_SHIFT};
    const long long ncomp = (long long)nb * ncell * NV;
    if (tq.alloc(std::max<long long>(nunits, 1)) || zhi.alloc(ncomp) || zlo.alloc(ncomp) ||
        red.alloc((size_t)NOUT3 * ntot) || chunk_tab.alloc(4096) || free_ids.alloc(1 << SLOT_SHIFT))
        return 2;
    TC_HIP(hipMemsetAsync(zhi.p, 0, sizeof(double) * ncomp, st));
    TC_HIP(hipMemsetAsync(zlo.p, 0, sizeof(double) * ncomp, st));
    if (nunits > 0)
        hipLaunchKernelGGL(k_units_to_tasks, dim3((unsigned)((nunits + 255) / 256)), dim3(256), 0,
                           st, units.p, (int)nunits, nb, tq.p);
    TC_HIP(hipGetLastError());
    // the units' lists stay in the split pool, which is the first level's input pool
    int* pin = spool.p;
    unsigned long long pin_cap = spool_top;
    bool pin_is_split = true;
    int free_top = 0;
    long long nslots = 0;
    std::vector<int> new_ids;
    ntask = nunits;
    sumlen = usum;
    for (int level = 0; ntask > 0; ++level) {
        if (level > 100000) return fail("multipole finish phase: too many levels");
        const long long pops = ntask * (level == 0 ? 2 : 1);
        // enough free slots for the pops of this level
        while (free_top < pops) {
            if ((long long)chunks.size() >= 4096 || nslots + (1 << SLOT_SHIFT) > 0x7fffffffll)
                return fail("multipole finish phase: too many scratch slots");
            double* c = static_cast<double*>(dev_alloc(cguard.bytes));
            if (!c) return fail("out of device memory for the multipole scratch slots");
            chunks.push_back(c);
            TC_HIP(hipMemcpyAsync(chunk_tab.p + chunks.size() - 1, &chunks.back(), sizeof(double*),
                                  hipMemcpyHostToDevice, st));
            if (free_ids.reserve(nslots + (1 << SLOT_SHIFT), free_top, st)) return 2;
            new_ids.resize(1 << SLOT_SHIFT);
            for (int i = 0; i < (1 << SLOT_SHIFT); ++i) new_ids[i] = (int)(nslots + (1 << SLOT_SHIFT) - 1 - i);
            TC_HIP(hipMemcpyAsync(free_ids.p + free_top, new_ids.data(), sizeof(int) * new_ids.size(),
                                  hipMemcpyHostToDevice, st));
            TC_HIP(hipStreamSynchronize(st));     // new_ids is reused
            free_top += 1 << SLOT_SHIFT;
            nslots += 1 << SLOT_SHIFT;
        }
        DBuf<int>& pout = (!pin_is_split && pin == poolA.p) ? poolB : poolA;
        const unsigned long long need = std::max<unsigned long long>(2 * sumlen, 1);
        if (need > 0x7fffffffull) return fail("multipole finish phase: list pool beyond 2^31 entries");
        if (pout.reserve(need, 0, st) || tchild.reserve(2 * ntask, 0, st) ||
            tq_next.reserve(2 * ntask, 0, st) || jobs.reserve(ntask, 0, st) ||
            ret.reserve(2 * ntask, 0, st))
            return 2;
        hipLaunchKernelGGL(k_mp_prep, dim3((unsigned)((ntask + 255) / 256)), dim3(256), 0, st,
                           tq.p, (int)ntask, free_ids.p, free_top);
        TC_HIP(hipGetLastError());
        const int top_after = free_top - (int)pops;
        const unsigned long long zero = 0;
        TC_HIP(hipMemcpyAsync(ptop.p, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
        if (timing) {
            TC_HIP(hipEventCreate(&e0)); TC_HIP(hipEventCreate(&e1)); TC_HIP(hipEventCreate(&e2));
            TC_HIP(hipEventRecord(e0, st));
        }
        hipLaunchKernelGGL(sfn, dim3((unsigned)ntask), dim3(MP_THREADS), shbytes, st, nodes, nnodes,
                           nsum.p, tq.p, pin, pin_cap, pout.p, ptop.p, (unsigned long long)pout.n,
                           chunk_tab.p, (int)nslots, P, tchild.p, jobs.p, ret.p, counters.p, err.p);
        TC_HIP(hipGetLastError());
        if (timing) TC_HIP(hipEventRecord(e1, st));
        // this level's zeta jobs, in task order
        // enough chunks for about 4096 workgroups, at least 8 jobs each
        const long long jpc = std::max<long long>(8, std::min<long long>(2048, (ntask * (X ? xsub : nmr) + 4095) / 4096));
        const int nchunks = (int)((ntask + jpc - 1) / jpc);
        if (chunk_rows.reserve(nchunks, 0, st)) return 2;
        if ((size_t)nchunks * ncomp > partial.n) {
            // entries with k3 < k2 are never written: they stay zero
            if (partial.reserve((size_t)nchunks * ncomp, 0, st)) return 2;
            TC_HIP(hipMemsetAsync(partial.p, 0, sizeof(double) * partial.n, st));
        }
        if (X) {
            const int qpad = (nchunks + 7) / 8 * 8;
            hipLaunchKernelGGL(zxfn, dim3((unsigned)(qpad * xsub)), dim3(ZT_THREADS), shbytes, st,
                               jobs.p, (int)ntask, nodes, chunk_tab.p, P, xmr, xkg, (int)jpc,
                               partial.p, chunk_rows.p, nchunks, xsub);
            TC_HIP(hipGetLastError());
        } else {
            // FP64: every cell; with the matrix cores, only the n = 0 sums, for m = 0
            const int zr_mr = mfn ? 1 : zmr, zr_sub = mfn ? 1 : nmr;
            const int qpad = (nchunks + 7) / 8 * 8;
            hipLaunchKernelGGL(zfn, dim3((unsigned)(qpad * zr_sub)), dim3(ZT_THREADS), zt_bytes, st,
                               jobs.p, (int)ntask, nodes, chunk_tab.p, P, zr_mr, (int)jpc, partial.p,
                               chunk_rows.p, nchunks, zr_sub, mfn ? 1 : 0);
            TC_HIP(hipGetLastError());
            if (mfn)
                hipLaunchKernelGGL(mfn, dim3((unsigned)(qpad * (M + 1))), dim3(mf_threads), mf_bytes,
                                   st, jobs.p, (int)ntask, nodes, chunk_tab.p, P, (int)jpc,
                                   partial.p, nchunks);
            TC_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_mp_zeta_reduce, dim3((unsigned)((ncomp + 255) / 256)), dim3(256), 0, st,
                           partial.p, chunk_rows.p, nchunks, ncomp, (long long)nb * (M + 1) * NV,
                           nb, zhi.p, zlo.p, X ? 1 : 0);
        TC_HIP(hipGetLastError());
        if (timing) TC_HIP(hipEventRecord(e2, st));
        // next tasks, and the slots that became free, in task order
        if (cpt.run(tchild.p, 2 * ntask, tq_next.p, 0, tot.p, st) ||
            cpi.run(ret.p, 2 * ntask, free_ids.p, top_after, tot.p + 2, st))
            return 2;
        unsigned long long h[3] = {0, 0, 0};
        int herr = 0;
        unsigned long long used = 0;
        TC_HIP(hipMemcpyAsync(h, tot.p, sizeof(h), hipMemcpyDeviceToHost, st));
        TC_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        TC_HIP(hipMemcpyAsync(&used, ptop.p, sizeof(used), hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (timing) {
            float ms1 = 0.f, ms2 = 0.f;
            TC_HIP(hipEventElapsedTime(&ms1, e0, e1));
            TC_HIP(hipEventElapsedTime(&ms2, e1, e2));
            ks_stage.ms += ms1; ks_stage.launches += 1;
            ks_zeta.ms += ms2; ks_zeta.launches += 1;
            (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2);
        }
        if (herr) return set_err_bits("multipole finish phase", herr);
        free_top = top_after + (int)h[2];
        ntask = (long long)h[0];
        sumlen = h[1];
        tq.swap(tq_next);
        pin = pout.p;
        pin_cap = used;
        pin_is_split = false;
        if (level == 0) { spool.release(); units.release(); }   // the units' lists are consumed
    }
    ExpandFn efn = X ? (D == GData ? &k_mp_expand_x<GData> : &k_mp_expand_x<NData>)
                     : (D == GData ? &k_mp_expand<GData> : &k_mp_expand<NData>);
    hipLaunchKernelGGL(efn, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, st, zhi.p, zlo.p, P,
                       red.p);
    TC_HIP(hipGetLastError());
    TC_HIP(hipMemcpyAsync(out, red.p, sizeof(double) * NOUT3 * ntot, hipMemcpyDeviceToHost, st));
    TC_HIP(hipMemcpyAsync(cnt, counters.p, sizeof(unsigned long long) * NC3, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    return 0;
}


// ---------------------------------------------------------------------------
// Multipole passes over two or three fields (the cross-correlation drivers)
// ---------------------------------------------------------------------------
// dst = src with the child links moved by `off` (the trees laid end to end in one array)
__global__ void k_nodes_shifted(const DNode* __restrict__ src, int n, int off,
                                DNode* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DNode d = src[i];
    if (d.child >= 0) d.child += off;
    dst[i] = d;
}
__global__ void k_ints_shifted(const int* __restrict__ src, int n, int off, int* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + off;
}

int corr3_run_trees(int D, int C, const DeviceTree& t1, const DeviceTree& t2,
                    const DeviceTree* t3, int ordered, const C3Params& P_in, double* out,
                    unsigned long long* cnt) {
    C3Params P = P_in;
    if (!t3 && &t1 == &t2) {
        P.order3 = 0;
        return corr3_run_x(D, C, t1.nodes, t1.nnodes, t1.dtop, (int)t1.top.size(), t1.dtop,
                           (int)t1.top.size(), P, out, cnt);
    }
    const DeviceTree* ts[3] = {&t1, &t2, t3};
    const int nt = t3 ? 3 : 2;
    long long total = 0;
    int offs[3] = {0, 0, 0};
    for (int i = 0; i < nt; ++i) {
        offs[i] = (int)total;
        total += ts[i]->nnodes;
    }
    if (total > 0x7ffffff0ll) return fail("multipole: the fields have more than 2^31 nodes in all");
    const int nc1 = (int)t1.top.size();
    int nl = 0;
    for (int i = 1; i < nt; ++i) nl += (int)ts[i]->top.size();
    DBuf<DNode> nodes;
    DBuf<int> c1top, ltop;
    if (nodes.alloc((size_t)total) || c1top.alloc(nc1) || ltop.alloc(nl)) return 2;
    hipStream_t st = stream();
    int lpos = 0;
    for (int i = 0; i < nt; ++i) {
        const int n = ts[i]->nnodes, nto = (int)ts[i]->top.size();
        hipLaunchKernelGGL(k_nodes_shifted, dim3((n + 255) / 256), dim3(256), 0, st, ts[i]->nodes,
                           n, offs[i], nodes.p + offs[i]);
        TC_HIP(hipGetLastError());
        if (i == 0) {
            hipLaunchKernelGGL(k_ints_shifted, dim3((nto + 255) / 256), dim3(256), 0, st,
                               ts[i]->dtop, nto, 0, c1top.p);
        } else {
            hipLaunchKernelGGL(k_ints_shifted, dim3((nto + 255) / 256), dim3(256), 0, st,
                               ts[i]->dtop, nto, offs[i], ltop.p + lpos);
            lpos += nto;
        }
        TC_HIP(hipGetLastError());
    }
    if (t3) {
        P.order3 = ordered;
        P.base3 = offs[2];
    } else {
        P.order3 = 0;
    }
    return corr3_run_x(D, C, nodes.p, (int)total, c1top.p, nc1, ltop.p, nl, P, out, cnt);
}

}  // namespace tc
