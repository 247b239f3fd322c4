// treecorr_amd: the 3-point multipole algorithm (Porth et al.) on the GPU,
// auto-correlation of one field for NNN and GGG (LogMultipole binning).
//
// TreeCorr (src/Corr3.cpp:1113-2849) recurses per top cell on each CPU thread.
// Here:
//   1. k_mp_split1: one workgroup per top cell walks multipoleSplit1
//      (:1374-1408, splitC2Cells :1323-1372) and emits independent "units":
//      a c1 cell small enough for the largest bin plus its c2 list.
//   2. k_mp_finish: persistent workgroups pull units and run multipoleFinish
//      (:1548-1613) depth first.  The multipole scratch (Wn, Gn, the pair sums;
//      include/MultipoleScratch.h) lives in LDS; a c1 split pushes a copy of it
//      onto a per-workgroup stack in HBM (TreeCorr's mp.duplicate()).  Every stage
//      processes its c2 list with one lane per c2 (splitC2CellsOrCalculateGn
//      :1452-1546): accepted pairs add their Fourier moments into LDS with FP64
//      atomics, split c2 cells are compacted in list order into the next list.
//      Zeta rows [kstart, mink_zeta) are contracted (calculateZeta :1861-1974,
//      MultipoleHelper<5> :2582-2849) by the whole workgroup into the
//      workgroup's private output slot.
//   3. k_mp_reduce: fixed-order compensated sum of the slots.
#include "tc_engine.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace tc {

struct NodeSum {          // BaseCell::calculateSumWSq / SumWZSq / SumAbsWZSq (src/Cell.cpp:143-168)
    double wsq, wzr, wzi, abswz;
};

// Scratch layout (doubles) in LDS, for nb bins and maxn M:
//   Wn  : nb*(M+1) complex           TreeCorr's Windex(k)+n (MultipoleScratch.h:80-81)
//   Gn  : nb*(2M+3) complex (GGG)    Gindex(k)+n, n in [-(M+1), M+1] (:334-335)
//   sums: npairs sumw sumwr sumwlogr sumww sumwwr sumwwlogr (7*nb)
//   gg  : sumwwgg0,1,2 (GGG) 3*nb complex
struct ScratchLayout {
    int nb, M, D;
    int wn, gn, sums, gg, total;
    __host__ __device__ void init(int nb_, int M_, int D_) {
        nb = nb_; M = M_; D = D_;
        wn = 0;
        gn = wn + 2 * nb * (M + 1);
        sums = gn + (D == GData ? 2 * nb * (2 * M + 3) : 0);
        gg = sums + 7 * nb;
        total = gg + (D == GData ? 6 * nb : 0);
    }
    __device__ int W(int k, int m) const { return wn + 2 * (k * (M + 1) + m); }
    __device__ int G(int k, int m) const { return gn + 2 * (k * (2 * M + 3) + (M + 1) + m); }
    __device__ int S(int q, int k) const { return sums + q * nb + k; }
    __device__ int GG(int q, int k) const { return gg + 2 * (q * nb + k); }
};

enum { S_NPAIRS = 0, S_SUMW, S_SUMWR, S_SUMWLOGR, S_SUMWW, S_SUMWWR, S_SUMWWLOGR };
// Output arrays per slot: zeta0..7, meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri
enum { O_Z0 = 0, O_MEAND2 = 8, O_MEANLOGD2, O_MEAND3, O_MEANLOGD3, O_WEIGHT, O_WEIGHT_IM, O_NTRI,
       NOUT3 = 15 };
enum { C3_GN = 0, C3_ZETA, C3_ROWS, C3_STAGES, NC3 = 4 };

constexpr int MP_THREADS = 256;
constexpr int MP_STACK = 96;               // DFS entries per workgroup
constexpr long long MP_ARENA0 = 1 << 20;  // initial list entries per workgroup (grown on demand)

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
    logr = 0.5 * log(rsq);
    const double kk = (logr - P.logminsep) / P.binsize;
    ik = int(kk);
    const double frackk = kk - ik;
    const double omf = 1. - frackk;
    const double f = (omf < frackk) ? omf : frackk;
    const double t1 = f * P.binsize + P.b;
    if (s1ps2sq > (t1 * t1) * rsq) return false;
    const double t2 = frackk * P.binsize + P.b - s1ps2sq / rsq;
    if (s1ps2sq > (t2 * t2) * rsq) return false;
    r = sqrt(rsq);
    return true;
}

__device__ __forceinline__ bool too_small3(const C3Params& P, double rsq, double s1ps2) {
    return rsq < P.minsepsq && s1ps2 < P.minsep && rsq < (P.minsep - s1ps2) * (P.minsep - s1ps2);
}
__device__ __forceinline__ bool too_large3(const C3Params& P, double rsq, double s1ps2) {
    return rsq >= P.maxsepsq && rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
}

__device__ __forceinline__ void cmul3(double ar, double ai, double br, double bi, double& cr,
                                      double& ci) {
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
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

// ---------------------------------------------------------------------------
// Block helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int block_scan_excl(int v, int* sh, int& total) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int base = 0;
    total = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        if (k < wv) base += sh[k];
        total += sh[k];
    }
    __syncthreads();
    return base + x - v;
}

struct Unit { int c1, off, len, pad; };

// Error bits reported through `err` (all of them end the kernel cleanly).
enum { E_POOL = 1, E_SPLITSTACK = 2, E_UNITS = 4, E_ARENA = 8, E_STACK = 16, E_INDEX = 32,
       E_BUDGET = 64 };

// Global data written by some waves of a workgroup and read by others after the
// barrier (c2 lists in the pool and the arena, zeta output slots).  The workgroup
// barrier alone waits only for LDS operations; the device-scope fences make this
// thread's global stores complete before the barrier and drop stale L1 lines after it.
__device__ __forceinline__ bool has_children(const DNode& c, int nnodes) {
    return c.child >= 0 && c.child < nnodes - 1;
}

__device__ __forceinline__ void block_sync_global() {
    __threadfence();
    __syncthreads();
    __threadfence();
}

// ---------------------------------------------------------------------------
// Phase 1: multipoleSplit1 per top cell -> units
// ---------------------------------------------------------------------------
// Every node index read from memory is checked against nnodes before it is
// dereferenced and every loop is bounded, so inconsistent data raises an error bit
// instead of an out-of-range access.
template <int C>
__global__ __launch_bounds__(MP_THREADS) void k_mp_split1(
    const DNode* __restrict__ nodes, int nnodes, const int* __restrict__ top, int ntop, C3Params P,
    int* __restrict__ pool, unsigned long long* pool_top, unsigned long long pool_cap,
    Unit* __restrict__ units, unsigned int* nunits, unsigned int unit_cap, int* err) {
    __shared__ int st_c1[MP_STACK], st_off[MP_STACK], st_len[MP_STACK];
    __shared__ int sh[MP_THREADS / 64];
    __shared__ unsigned long long s_base;
    __shared__ int s_bad;
    int sp = 0;
    if (threadIdx.x == 0) { st_c1[0] = top[blockIdx.x]; st_off[0] = -1; st_len[0] = ntop; }
    sp = 1;
    __syncthreads();
    for (int iter = 0; sp > 0; ++iter) {
        --sp;
        const int i1 = st_c1[sp], off = st_off[sp], len = st_len[sp];
        if (threadIdx.x == 0) s_bad = 0;
        __syncthreads();
        const bool ok = iter <= nnodes && (unsigned)i1 < (unsigned)nnodes && len >= 0 &&
                        (off < 0 ? len == ntop : (unsigned long long)off + len <= pool_cap);
        if (!ok) {
            if (threadIdx.x == 0) atomicOr(err, iter > nnodes ? E_BUDGET : E_INDEX);
            return;
        }
        const DNode c1 = nodes[i1];
        const double s1 = c1.size;
        const int* list = off < 0 ? top : pool + off;
        // splitC2Cells, src/Corr3.cpp:1323-1372
        // pass 1: count outputs
        int total_out = 0;
        for (int base = 0; base < len; base += blockDim.x) {
            int t = base + threadIdx.x;
            int no = 0;
            if (t < len) {
                const int i2 = list[t];
                if ((unsigned)i2 >= (unsigned)nnodes) {
                    s_bad = 1;
                } else {
                    const DNode c2 = nodes[i2];
                    const double s2 = c2.size;
                    const double rsq = dsq3<C>(c1, c2);
                    const double s1ps2 = s1 + s2;
                    if (!too_small3(P, rsq, s1ps2) && !too_large3(P, rsq, s1ps2)) {
                        int k = -1; double r = 0, logr = 0;
                        bool split = !single_bin3(P, rsq, s1ps2, k, r, logr);
                        if (split && s2 > s1) {
                            if (!has_children(c2, nnodes)) s_bad = 1;
                            else no = 2;
                        } else {
                            no = 1;
                        }
                    }
                }
            }
            int tot;
            block_scan_excl(no, sh, tot);
            total_out += tot;
        }
        if (s_bad) {
            if (threadIdx.x == 0) atomicOr(err, E_INDEX);
            return;
        }
        if (threadIdx.x == 0) {
            s_base = atomicAdd(pool_top, (unsigned long long)total_out);
            if (s_base + total_out > pool_cap) atomicOr(err, E_POOL);
        }
        __syncthreads();
        const unsigned long long obase = s_base;
        if (obase + total_out > pool_cap) return;
        // pass 2: write outputs in list order (the same decisions as pass 1)
        int wpos = 0;
        for (int base = 0; base < len; base += blockDim.x) {
            int t = base + threadIdx.x;
            int no = 0, o0 = 0, o1 = 0;
            if (t < len) {
                const int i2 = list[t];
                const DNode c2 = nodes[i2];
                const double s2 = c2.size;
                const double rsq = dsq3<C>(c1, c2);
                const double s1ps2 = s1 + s2;
                if (!too_small3(P, rsq, s1ps2) && !too_large3(P, rsq, s1ps2)) {
                    int k = -1; double r = 0, logr = 0;
                    bool split = !single_bin3(P, rsq, s1ps2, k, r, logr);
                    if (split && s2 > s1) { no = 2; o0 = c2.child; o1 = c2.child + 1; }
                    else { no = 1; o0 = i2; }
                }
            }
            int tot;
            int ex = block_scan_excl(no, sh, tot);
            if (wpos + ex + no <= total_out) {
                if (no > 0) pool[obase + wpos + ex] = o0;
                if (no > 1) pool[obase + wpos + ex + 1] = o1;
            }
            wpos += tot;
        }
        // the children's passes read this list from other waves
        block_sync_global();
        const int noff = (int)obase;
        if (s1 * s1 > P.maxbsq_eff) {
            if (sp + 2 > MP_STACK || !has_children(c1, nnodes)) {
                if (threadIdx.x == 0) atomicOr(err, sp + 2 > MP_STACK ? E_SPLITSTACK : E_INDEX);
                return;
            }
            if (threadIdx.x == 0) {
                st_c1[sp] = c1.child + 1; st_off[sp] = noff; st_len[sp] = total_out;
                st_c1[sp + 1] = c1.child; st_off[sp + 1] = noff; st_len[sp + 1] = total_out;
            }
            sp += 2;
        } else if (threadIdx.x == 0) {
            unsigned int u = atomicAdd(nunits, 1u);
            if (u < unit_cap) units[u] = Unit{i1, noff, total_out, 0};
            else atomicOr(err, E_UNITS);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Phase 2: multipoleFinish per unit
// ---------------------------------------------------------------------------
struct MPWork {
    int* arena;            // per workgroup: arena_cap ints
    double* scratch;       // per workgroup: MP_STACK * layout.total doubles
    int* st;               // per workgroup: MP_STACK * 8 ints (c1, off, len, mink, maxr_lo, maxr_hi)
};

// Corr3::calculateGn + MultipoleScratch::calculateGn for one accepted (c1,c2)
template <int D, int C>
__device__ __forceinline__ void add_gn(const ScratchLayout& Lo, double* sc, const DNode& c1,
                                       const DNode& c2, const NodeSum& s2, double r,
                                       double logr, int k, int maxn) {
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
    atomicAdd(&sc[Lo.W(k, 0)], w);
    double wr = w, wi = 0.;
    for (int m = 1; m <= maxn; ++m) {
        double tr, ti;
        cmul3(wr, wi, zr, zi, tr, ti);
        wr = tr; wi = ti;
        atomicAdd(&sc[Lo.W(k, m)], wr);
        atomicAdd(&sc[Lo.W(k, m) + 1], wi);
    }
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
        double gr, gi;
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
        atomicAdd(&sc[Lo.G(k, 0)], gr);
        atomicAdd(&sc[Lo.G(k, 0) + 1], gi);
        double ar = gr, ai = gi;
        for (int m = 1; m <= maxn + 1; ++m) {
            cmul3(ar, ai, zr, zi, tr, ti);
            ar = tr; ai = ti;
            atomicAdd(&sc[Lo.G(k, m)], ar);
            atomicAdd(&sc[Lo.G(k, m) + 1], ai);
        }
        ar = gr; ai = gi;
        for (int m = 1; m <= maxn + 1; ++m) {
            cmul3(ar, ai, zr, -zi, tr, ti);
            ar = tr; ai = ti;
            atomicAdd(&sc[Lo.G(k, -m)], ar);
            atomicAdd(&sc[Lo.G(k, -m) + 1], ai);
        }
    }
}

// Corr3::calculateZeta + MultipoleHelper<0>/<5>, written by the whole workgroup into
// this workgroup's output slot.  Each (k2, k3>=k2, m>=0) triple is one work item; the
// outputs it touches are disjoint from every other item's.
template <int D>
__device__ void zeta_block(const ScratchLayout& Lo, const double* sc, const DNode& c1,
                           int kstart, int mink, const C3Params& P, double* slot, int ntot) {
    const int nb = P.nbins, M = P.maxn, nn = 2 * M + 1;
    const double w1 = c1.w, n1 = double(c1.n);
    const double g1r = c1.a, g1i = c1.b;
    // items: for k2 in [kstart, mink): (nb - k2) * (M + 1)
    int nitems = 0;
    for (int k2 = kstart; k2 < mink; ++k2) nitems += (nb - k2) * (M + 1);
    for (int it = threadIdx.x; it < nitems; it += blockDim.x) {
        int rem = it, k2 = kstart;
        while (rem >= (nb - k2) * (M + 1)) { rem -= (nb - k2) * (M + 1); ++k2; }
        const int k3 = k2 + rem / (M + 1);
        const int m = rem % (M + 1);
        const double np2 = sc[Lo.S(S_NPAIRS, k2)], sw2 = sc[Lo.S(S_SUMW, k2)];
        if (k3 == k2) {
            const int iz22 = (k2 * nb + k2) * nn + M;
            if (m == 0 && !P.quick) {
                const double swr = sc[Lo.S(S_SUMWR, k2)], swlr = sc[Lo.S(S_SUMWLOGR, k2)];
                const double swwr = sc[Lo.S(S_SUMWWR, k2)], swwlr = sc[Lo.S(S_SUMWWLOGR, k2)];
                slot[O_NTRI * ntot + iz22] += n1 * np2 * (np2 - 1);
                slot[O_MEAND2 * ntot + iz22] += w1 * (sw2 * swr - swwr);
                slot[O_MEANLOGD2 * ntot + iz22] += w1 * (sw2 * swlr - swwlr);
                slot[O_MEAND3 * ntot + iz22] += w1 * (sw2 * swr - swwr);
                slot[O_MEANLOGD3 * ntot + iz22] += w1 * (sw2 * swlr - swwlr);
            }
            const double wr = sc[Lo.W(k2, m)], wi = sc[Lo.W(k2, m) + 1];
            const double www = w1 * ((wr * wr + wi * wi) - sc[Lo.S(S_SUMWW, k2)]);
            slot[O_WEIGHT * ntot + iz22 + m] += www;
            if (m > 0) slot[O_WEIGHT * ntot + iz22 - m] += www;
        } else {
            const int i23 = (k2 * nb + k3) * nn + M;
            const int i32 = (k3 * nb + k2) * nn + M;
            if (m == 0 && !P.quick) {
                const double np3 = sc[Lo.S(S_NPAIRS, k3)], sw3 = sc[Lo.S(S_SUMW, k3)];
                const double nnn = n1 * np2 * np3;
                slot[O_NTRI * ntot + i23] += nnn;
                slot[O_NTRI * ntot + i32] += nnn;
                const double ww12 = w1 * sw2, ww13 = w1 * sw3;
                const double wwwd2 = ww13 * sc[Lo.S(S_SUMWR, k2)];
                slot[O_MEAND2 * ntot + i23] += wwwd2;
                slot[O_MEAND3 * ntot + i32] += wwwd2;
                const double wwwlogd2 = ww13 * sc[Lo.S(S_SUMWLOGR, k2)];
                slot[O_MEANLOGD2 * ntot + i23] += wwwlogd2;
                slot[O_MEANLOGD3 * ntot + i32] += wwwlogd2;
                const double wwwd3 = ww12 * sc[Lo.S(S_SUMWR, k3)];
                slot[O_MEAND3 * ntot + i23] += wwwd3;
                slot[O_MEAND2 * ntot + i32] += wwwd3;
                const double wwwlogd3 = ww12 * sc[Lo.S(S_SUMWLOGR, k3)];
                slot[O_MEANLOGD3 * ntot + i23] += wwwlogd3;
                slot[O_MEANLOGD2 * ntot + i32] += wwwlogd3;
            }
            // weight: w1 * Wn[k2,m] * conj(Wn[k3,m])
            const double ar = w1 * sc[Lo.W(k2, m)], ai = w1 * sc[Lo.W(k2, m) + 1];
            double xr, xi;
            cmul3(ar, ai, sc[Lo.W(k3, m)], -sc[Lo.W(k3, m) + 1], xr, xi);
            slot[O_WEIGHT * ntot + i23 + m] += xr;
            slot[O_WEIGHT_IM * ntot + i23 + m] += xi;
            slot[O_WEIGHT * ntot + i32 + m] += xr;
            slot[O_WEIGHT_IM * ntot + i32 + m] -= xi;
            if (m > 0) {
                slot[O_WEIGHT * ntot + i23 - m] += xr;
                slot[O_WEIGHT_IM * ntot + i23 - m] -= xi;
                slot[O_WEIGHT * ntot + i32 - m] += xr;
                slot[O_WEIGHT_IM * ntot + i32 - m] += xi;
            }
        }
        if (D != GData) continue;
        // MultipoleHelper<5>::CalculateZeta, src/Corr3.cpp:2582-2849
        auto Gr = [&](int k, int n) { return sc[Lo.G(k, n)]; };
        auto Gi = [&](int k, int n) { return sc[Lo.G(k, n) + 1]; };
        auto add8 = [&](int idx, double a0r, double a0i, double a1r, double a1i, double a2r,
                        double a2i, double a3r, double a3i) {
            slot[(O_Z0 + 0) * ntot + idx] += a0r; slot[(O_Z0 + 1) * ntot + idx] += a0i;
            slot[(O_Z0 + 2) * ntot + idx] += a1r; slot[(O_Z0 + 3) * ntot + idx] += a1i;
            slot[(O_Z0 + 4) * ntot + idx] += a2r; slot[(O_Z0 + 5) * ntot + idx] += a2i;
            slot[(O_Z0 + 6) * ntot + idx] += a3r; slot[(O_Z0 + 7) * ntot + idx] += a3i;
        };
        const double g1cr = g1r, g1ci = -g1i;
        if (k3 == k2) {
            const int iz22 = (k2 * nb + k2) * nn + M;
            double c0r, c0i, c1r, c1i, c2r, c2i;
            cmul3(g1r, g1i, sc[Lo.GG(0, k2)], sc[Lo.GG(0, k2) + 1], c0r, c0i);
            cmul3(g1cr, g1ci, sc[Lo.GG(1, k2)], sc[Lo.GG(1, k2) + 1], c1r, c1i);
            cmul3(g1r, g1i, sc[Lo.GG(2, k2)], sc[Lo.GG(2, k2) + 1], c2r, c2i);
            if (m == 0) {
                double ar, ai, br, bi, tr, ti;
                cmul3(g1r, g1i, Gr(k2, -1), Gi(k2, -1), ar, ai);     // g1G2m
                cmul3(g1cr, g1ci, Gr(k2, 1), Gi(k2, 1), br, bi);     // g1cG2p
                double x0r, x0i, x1r, x1i, x2r, x2i;
                cmul3(ar, ai, Gr(k2, -1), Gi(k2, -1), tr, ti); x0r = tr - c0r; x0i = ti - c0i;
                cmul3(br, bi, Gr(k2, 1), Gi(k2, 1), tr, ti); x1r = tr - c1r; x1i = ti - c1i;
                cmul3(ar, ai, Gr(k2, 1), -Gi(k2, 1), tr, ti); x2r = tr - c2r; x2i = ti - c2i;
                add8(iz22, x0r, x0i, x1r, x1i, x2r, x2i, x2r, x2i);
            } else {
                const double ppr = Gr(k2, m + 1), ppi = Gi(k2, m + 1);
                const double pmr = Gr(k2, m - 1), pmi = Gi(k2, m - 1);
                const double mpr = Gr(k2, -m + 1), mpi = Gi(k2, -m + 1);
                const double mmr = Gr(k2, -m - 1), mmi = Gi(k2, -m - 1);
                double ar, ai, br, bi, tr, ti;
                cmul3(g1r, g1i, mmr, mmi, ar, ai);       // g1G2mm
                cmul3(g1cr, g1ci, ppr, ppi, br, bi);     // g1cG2pp
                double x0r, x0i, x1r, x1i, x2r, x2i, x3r, x3i;
                cmul3(ar, ai, pmr, pmi, tr, ti); x0r = tr - c0r; x0i = ti - c0i;
                cmul3(br, bi, mpr, mpi, tr, ti); x1r = tr - c1r; x1i = ti - c1i;
                cmul3(br, -bi, pmr, pmi, tr, ti); x2r = tr - c2r; x2i = ti - c2i;
                cmul3(ar, ai, mpr, -mpi, tr, ti); x3r = tr - c2r; x3i = ti - c2i;
                add8(iz22 + m, x0r, x0i, x1r, x1i, x2r, x2i, x3r, x3i);
                add8(iz22 - m, x0r, x0i, x1r, x1i, x3r, x3i, x2r, x2i);
            }
        } else {
            const int iz23 = (k2 * nb + k3) * nn + M;
            const int iz32 = (k3 * nb + k2) * nn + M;
            if (m == 0) {
                double ar, ai, br, bi;
                cmul3(g1r, g1i, Gr(k2, -1), Gi(k2, -1), ar, ai);     // g1G2m
                cmul3(g1cr, g1ci, Gr(k2, 1), Gi(k2, 1), br, bi);     // g1cG2p
                double a0r, a0i, a1r, a1i, a2r, a2i, a3r, a3i;
                cmul3(ar, ai, Gr(k3, -1), Gi(k3, -1), a0r, a0i);
                cmul3(br, bi, Gr(k3, 1), Gi(k3, 1), a1r, a1i);
                cmul3(ar, ai, Gr(k3, 1), -Gi(k3, 1), a2r, a2i);
                cmul3(br, -bi, Gr(k3, -1), Gi(k3, -1), a3r, a3i);
                add8(iz23, a0r, a0i, a1r, a1i, a2r, a2i, a3r, a3i);
                add8(iz32, a0r, a0i, a1r, a1i, a3r, a3i, a2r, a2i);
            } else {
                const double G2ppr = Gr(k2, m + 1), G2ppi = Gi(k2, m + 1);
                const double G2pmr = Gr(k2, m - 1), G2pmi = Gi(k2, m - 1);
                const double G2mpr = Gr(k2, -m + 1), G2mpi = Gi(k2, -m + 1);
                const double G2mmr = Gr(k2, -m - 1), G2mmi = Gi(k2, -m - 1);
                const double G3ppr = Gr(k3, m + 1), G3ppi = Gi(k3, m + 1);
                const double G3pmr = Gr(k3, m - 1), G3pmi = Gi(k3, m - 1);
                const double G3mpr = Gr(k3, -m + 1), G3mpi = Gi(k3, -m + 1);
                const double G3mmr = Gr(k3, -m - 1), G3mmi = Gi(k3, -m - 1);
                double pmR, pmI, cpR, cpI, mcR, mcI, mmR, mmI;
                cmul3(g1r, g1i, G2pmr, G2pmi, pmR, pmI);        // g1G2pm
                cmul3(g1cr, g1ci, G2ppr, G2ppi, cpR, cpI);      // g1cG2pp
                cmul3(g1r, g1i, G2mpr, -G2mpi, mcR, mcI);       // g1G2mpc
                cmul3(g1r, g1i, G2mmr, G2mmi, mmR, mmI);        // g1G2mm
                double p0r, p0i, p1r, p1i, p2r, p2i, p3r, p3i;
                cmul3(pmR, pmI, G3mmr, G3mmi, p0r, p0i);
                cmul3(cpR, cpI, G3mpr, G3mpi, p1r, p1i);
                cmul3(pmR, pmI, G3ppr, -G3ppi, p2r, p2i);
                cmul3(mcR, mcI, G3mmr, G3mmi, p3r, p3i);
                add8(iz23 + m, p0r, p0i, p1r, p1i, p2r, p2i, p3r, p3i);
                add8(iz32 - m, p0r, p0i, p1r, p1i, p3r, p3i, p2r, p2i);
                double q0r, q0i, q1r, q1i, q2r, q2i, q3r, q3i;
                cmul3(mmR, mmI, G3pmr, G3pmi, q0r, q0i);
                cmul3(mcR, -mcI, G3ppr, G3ppi, q1r, q1i);
                cmul3(cpR, -cpI, G3pmr, G3pmi, q2r, q2i);
                cmul3(mmR, mmI, G3mpr, -G3mpi, q3r, q3i);
                add8(iz32 + m, q0r, q0i, q1r, q1i, q2r, q2i, q3r, q3i);
                add8(iz23 - m, q0r, q0i, q1r, q1i, q3r, q3i, q2r, q2i);
            }
        }
    }
}

// Task record in LDS: every control value of the depth-first walk is read from
// memory by thread 0 only and broadcast through LDS, so all waves of the workgroup
// always agree on it.
enum { T_C1 = 0, T_OFF, T_LEN, T_MINK, T_MAXR_LO, T_MAXR_HI, T_KIND, T_N };
enum { TASK_NONE = 0, TASK_POP = 1, TASK_UNIT = 2, TASK_BAD = 3, TASK_STOP = 4 };

template <int D, int C>
__global__ __launch_bounds__(MP_THREADS) void k_mp_finish(
    const DNode* __restrict__ nodes, int nnodes, const NodeSum* __restrict__ nsum,
    const int* __restrict__ pool, unsigned long long pool_cap, const Unit* __restrict__ units,
    unsigned int nunits, unsigned int* next_unit, unsigned long long* stage_count,
    unsigned long long stage_budget, C3Params P, int* arena_all, long long arena_cap,
    double* scratch_all,
    int* stack_all, double* slots, int ntot, unsigned long long* counters, int* err) {
    HIP_DYNAMIC_SHARED(double, sc)
    __shared__ int sh[MP_THREADS / 64];
    __shared__ int s_task[T_N];
    __shared__ int s_anysplit, s_bad, s_stop;
    __shared__ double s_maxr;
    ScratchLayout Lo;
    Lo.init(P.nbins, P.maxn, D);
    int* arena = arena_all + (size_t)blockIdx.x * arena_cap;
    double* sstack = scratch_all + (size_t)blockIdx.x * MP_STACK * Lo.total;
    int* st = stack_all + (size_t)blockIdx.x * MP_STACK * 8;
    double* slot = slots + (size_t)blockIdx.x * NOUT3 * ntot;
    unsigned long long c_gn = 0, c_zeta = 0, c_rows = 0, c_stage = 0;

    int sp = 0;           // DFS stack depth (uniform)
    int have = 0;         // current task valid (uniform)
    int c1i = 0, loff = 0, llen = 0, mink = 0;
    bool in_pool = true;  // current list lives in the unit pool (else in the arena)
    double maxr = 0.;
    long long atop = 0;   // arena top (uniform)
    while (true) {
        if (!have) {
            // next task: the DFS stack of this workgroup, else a new unit
            if (threadIdx.x == 0) {
                int kind = TASK_NONE;
                if (sp > 0) {
                    const int* e = st + (sp - 1) * 8;     // written by this same thread
                    for (int q = 0; q < 6; ++q) s_task[q] = e[q];
                    kind = TASK_POP;
                } else {
                    const unsigned int u = atomicAdd(next_unit, 1u);
                    if (u < nunits) {
                        const Unit un = units[u];
                        s_task[T_C1] = un.c1; s_task[T_OFF] = un.off; s_task[T_LEN] = un.len;
                        s_task[T_MINK] = P.nbins; s_task[T_MAXR_LO] = 0; s_task[T_MAXR_HI] = 0;
                        kind = TASK_UNIT;
                    }
                }
                if (kind != TASK_NONE) {
                    const int c = s_task[T_C1], o = s_task[T_OFF], l = s_task[T_LEN];
                    const int mk = s_task[T_MINK];
                    const unsigned long long cap = kind == TASK_UNIT ? pool_cap : (unsigned long long)arena_cap;
                    if ((unsigned)c >= (unsigned)nnodes || o < 0 || l < 0 ||
                        (unsigned long long)o + l > cap || mk < 0 || mk > P.nbins)
                        kind = TASK_BAD;
                }
                s_task[T_KIND] = kind;
            }
            __syncthreads();
            const int kind = s_task[T_KIND];
            if (kind == TASK_NONE) break;
            if (kind == TASK_BAD) {
                if (threadIdx.x == 0) atomicOr(err, E_INDEX);
                break;
            }
            c1i = s_task[T_C1]; loff = s_task[T_OFF]; llen = s_task[T_LEN]; mink = s_task[T_MINK];
            maxr = __hiloint2double(s_task[T_MAXR_HI], s_task[T_MAXR_LO]);
            if (kind == TASK_POP) {
                --sp;
                in_pool = false;
                atop = (long long)loff + llen;
                for (int t = threadIdx.x; t < Lo.total; t += blockDim.x)
                    sc[t] = sstack[(size_t)sp * Lo.total + t];     // same thread wrote it
            } else {
                in_pool = true;
                atop = 0;
                for (int t = threadIdx.x; t < Lo.total; t += blockDim.x) sc[t] = 0.;
            }
            __syncthreads();
            have = 1;
        }
        // ---- one multipoleFinish stage for (c1, list) ----
        ++c_stage;
        if (threadIdx.x == 0) {
            s_anysplit = 0; s_maxr = 0.; s_bad = 0;
            s_stop = atomicAdd(stage_count, 1ull) >= stage_budget;
        }
        __syncthreads();
        if (s_stop) {
            if (threadIdx.x == 0) atomicOr(err, E_BUDGET);
            break;
        }
        const DNode c1 = nodes[c1i];
        const double s1 = c1.size;
        const int* list = in_pool ? pool + loff : arena + loff;
        const double thresh1 = (1. - P.binsize) * maxr;
        const long long nlo = atop;      // new list starts at the arena top
        int nlen = 0;
        double lmax = 0.;
        int lany = 0;
        for (int base = 0; base < llen; base += blockDim.x) {
            const int t = base + threadIdx.x;
            int no = 0, o0 = 0, o1 = 0;
            const int i2 = t < llen ? list[t] : 0;
            if (t < llen && (unsigned)i2 >= (unsigned)nnodes) {
                s_bad = 1;
            } else if (t < llen) {
                const DNode c2 = nodes[i2];
                const double s2 = c2.size;
                const double rsq = dsq3<C>(c1, c2);
                const double s1ps2 = s1 + s2;
                if (!too_small3(P, rsq, s1ps2) && !too_large3(P, rsq, s1ps2)) {
                    int k = -1;
                    double r = sqrt(rsq), logr = 0.;
                    if (single_bin3(P, rsq, s1ps2, k, r, logr)) {
                        if (rsq >= P.minsepsq && rsq < P.maxsepsq) {
                            if (k < 0) {
                                logr = log(r);
                                k = int((logr - P.logminsep) / P.binsize);
                            }
                            if (k >= 0 && k < P.nbins) {
                                ++c_gn;
                                add_gn<D, C>(Lo, sc, c1, c2, nsum[i2], r, logr, k, P.maxn);
                            } else {
                                s_bad = 1;     // r at the very edge of the range: no bin
                            }
                        }
                    } else {
                        const double rtot = r + s1ps2;
                        lmax = fmax(lmax, rtot);
                        const double bsq_eff = rsq * (P.asq < P.bsq ? P.asq : P.bsq);
                        bool sp1, sp2;
                        if (s2 > s1) {
                            sp2 = true;
                            sp1 = (s2 > 2. * s1) ? false : (s1 * s1 > 0.3422 * bsq_eff);
                        } else if (s1 > 2. * s2) {
                            sp1 = true; sp2 = false;
                        } else {
                            sp1 = true;
                            sp2 = s2 * s2 > 0.3422 * bsq_eff;
                        }
                        if (sp1 && rtot > thresh1) lany = 1;
                        if (sp2) {
                            if (has_children(c2, nnodes)) { no = 2; o0 = c2.child; o1 = c2.child + 1; }
                            else s_bad = 1;
                        } else {
                            no = 1; o0 = i2;
                        }
                    }
                }
            }
            int tot;
            const int ex = block_scan_excl(no, sh, tot);
            if (nlo + nlen + tot > arena_cap) {
                if (threadIdx.x == 0) atomicOr(err, E_ARENA);
            } else {
                if (no > 0) arena[nlo + nlen + ex] = o0;
                if (no > 1) arena[nlo + nlen + ex + 1] = o1;
            }
            nlen += tot;
        }
        if (lany) atomicOr(&s_anysplit, 1);
        // block max of the remaining separation (all values >= 0)
        {
            double v = lmax;
            for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
            if ((threadIdx.x & 63) == 0)
                atomicMax((unsigned long long*)&s_maxr, (unsigned long long)__double_as_longlong(v));
        }
        // the next stage reads the new list from other waves
        block_sync_global();
        if (s_bad) {
            if (threadIdx.x == 0) atomicOr(err, E_INDEX);
            break;
        }
        if (nlo + nlen > arena_cap) break;   // overflow (flagged above): stop cleanly
        const int anysplit1 = s_anysplit;
        const double nmaxr = s_maxr;
        __syncthreads();
        if (nlen > 0) {
            int nmink = mink;
            if (nmaxr < P.maxsep) {
                const int k = nmaxr < P.minsep ? 0
                              : int((log(nmaxr) - P.logminsep) / P.binsize) + 1;
                if (k >= 0 && k < mink) {
                    ++c_zeta; c_rows += (unsigned long long)(mink - k);
                    zeta_block<D>(Lo, sc, c1, k, mink, P, slot, ntot);
                    block_sync_global();
                    nmink = k;
                }
            }
            if (anysplit1) {
                if (sp >= MP_STACK || !has_children(c1, nnodes)) {
                    if (threadIdx.x == 0) atomicOr(err, sp >= MP_STACK ? E_STACK : E_INDEX);
                    break;
                }
                if (threadIdx.x == 0) {
                    int* e = st + sp * 8;
                    e[0] = c1.child + 1; e[1] = (int)nlo; e[2] = nlen; e[3] = nmink;
                    e[4] = __double2loint(nmaxr); e[5] = __double2hiint(nmaxr);
                }
                for (int t = threadIdx.x; t < Lo.total; t += blockDim.x)
                    sstack[(size_t)sp * Lo.total + t] = sc[t];
                ++sp;
                c1i = c1.child;
            }
            loff = (int)nlo; llen = nlen; in_pool = false; mink = nmink; maxr = nmaxr;
            atop = nlo + nlen;
        } else {
            ++c_zeta; c_rows += (unsigned long long)mink;
            zeta_block<D>(Lo, sc, c1, 0, mink, P, slot, ntot);
            block_sync_global();
            have = 0;
        }
    }
    atomicAdd(&counters[C3_GN], c_gn);
    if (threadIdx.x == 0) {
        atomicAdd(&counters[C3_ZETA], c_zeta);
        atomicAdd(&counters[C3_ROWS], c_rows);
        atomicAdd(&counters[C3_STAGES], c_stage);
    }
}

__global__ void k_mp_reduce(const double* slots, int nslots, long long n, double* out) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = 0., c = 0.;
    for (int k = 0; k < nslots; ++k) {
        const double v = slots[(size_t)k * n + t];
        const double tt = s + v;
        c += (fabs(s) >= fabs(v)) ? (s - tt) + v : (v - tt) + s;
        s = tt;
    }
    out[t] = s + c;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t bytes = 0;
    ~DBuf() { dev_free(p, bytes); }
    int alloc(size_t n) {
        dev_free(p, bytes);
        bytes = sizeof(T) * std::max<size_t>(n, 1);
        p = static_cast<T*>(dev_alloc(bytes));
        if (!p) { bytes = 0; return fail("out of device memory in the multipole engine"); }
        return 0;
    }
};

using SplitFn = void (*)(const DNode*, int, const int*, int, C3Params, int*, unsigned long long*,
                         unsigned long long, Unit*, unsigned int*, unsigned int, int*);
using FinishFn = void (*)(const DNode*, int, const NodeSum*, const int*, unsigned long long,
                          const Unit*, unsigned int, unsigned int*, unsigned long long*,
                          unsigned long long, C3Params, int*, long long, double*, int*,
                          double*, int, unsigned long long*, int*);

static SplitFn pick_split(int C) {
    if (C == Flat) return &k_mp_split1<Flat>;
    if (C == ThreeD) return &k_mp_split1<ThreeD>;
    return &k_mp_split1<Sphere>;
}
static FinishFn pick_finish(int D, int C) {
#define TC_FF(DD, CC) return &k_mp_finish<DD, CC>
    if (D == NData) { if (C == Flat) TC_FF(NData, Flat); if (C == ThreeD) TC_FF(NData, ThreeD); TC_FF(NData, Sphere); }
    if (C == Flat) TC_FF(GData, Flat); if (C == ThreeD) TC_FF(GData, ThreeD); TC_FF(GData, Sphere);
#undef TC_FF
}

static int set_err_bits(const char* phase, int bits) {
    std::string m = std::string(phase) + " stopped:";
    if (bits & E_POOL) m += " work pool overflow;";
    if (bits & E_SPLITSTACK) m += " split stack overflow;";
    if (bits & E_UNITS) m += " unit list overflow;";
    if (bits & E_ARENA) m += " list arena overflow;";
    if (bits & E_STACK) m += " depth-first stack overflow;";
    if (bits & E_INDEX) m += " inconsistent tree or list data;";
    if (bits & E_BUDGET) m += " stage budget exhausted;";
    return fail(m);
}

// Runs the multipole auto-correlation of one device tree; writes NOUT3*ntot reduced
// sums (zeta0..7, meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri) into
// out (host) and the work counters into cnt (4 values).
int corr3_run(int D, int C, const DNode* nodes, int nnodes, const int* dtop, int ntop,
              const C3Params& P, double* out, unsigned long long* cnt) {
    hipStream_t st = stream();
    const int nb = P.nbins, M = P.maxn;
    const int ntot = nb * nb * (2 * M + 1);
    ScratchLayout Lo;
    Lo.init(nb, M, D);
    const size_t shbytes = sizeof(double) * Lo.total;
    if (shbytes > 150 * 1024) return fail("nbins and max_n too large for the LDS scratch");
    FinishFn ffn = pick_finish(D, C);
    TC_HIP(hipFuncSetAttribute((const void*)ffn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shbytes));

    // node sums, bottom up: each launch finishes the nodes whose children are done
    DBuf<NodeSum> nsum;
    DBuf<int> doneA, doneB, changed;
    if (nsum.alloc(nnodes) || doneA.alloc(nnodes) || doneB.alloc(nnodes) || changed.alloc(1))
        return 2;
    TC_HIP(hipMemsetAsync(doneA.p, 0, sizeof(int) * nnodes, st));
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

    // phase 1: units.  The c2 lists of the split walk go to one pool; its size is
    // not known beforehand, so an overflowing walk is repeated with a larger pool.
    unsigned long long pool_cap = std::max<unsigned long long>(1ull << 22, 64ull * nnodes);
    pool_cap = std::min<unsigned long long>(pool_cap, (1ull << 31) - 1);
    unsigned int unit_cap = (unsigned int)std::min<long long>(std::max(nnodes, 1024), 1 << 30);
    DBuf<int> pool, err;
    DBuf<unsigned long long> pool_top, counters;
    DBuf<Unit> units;
    DBuf<unsigned int> nunits, next_unit;
    if (err.alloc(1) || pool_top.alloc(1) || counters.alloc(NC3) || units.alloc(unit_cap) ||
        nunits.alloc(1) || next_unit.alloc(1))
        return 2;
    unsigned int hn = 0;
    int herr = 0;
    for (;;) {
        if (pool.alloc(pool_cap)) return 2;
        TC_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
        TC_HIP(hipMemsetAsync(pool_top.p, 0, sizeof(unsigned long long), st));
        TC_HIP(hipMemsetAsync(nunits.p, 0, sizeof(unsigned int), st));
        hipLaunchKernelGGL(pick_split(C), dim3(ntop), dim3(MP_THREADS), 0, st, nodes, nnodes,
                           dtop, ntop, P, pool.p, pool_top.p, pool_cap, units.p, nunits.p,
                           unit_cap, err.p);
        TC_HIP(hipGetLastError());
        unsigned long long used = 0;
        TC_HIP(hipMemcpyAsync(&hn, nunits.p, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        TC_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        TC_HIP(hipMemcpyAsync(&used, pool_top.p, sizeof(used), hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (herr == E_POOL && pool_cap < (1ull << 31) - 1) {
            pool_cap = std::min<unsigned long long>(std::max(2 * pool_cap, 2 * used),
                                                    (1ull << 31) - 1);
            continue;
        }
        break;
    }
    if (herr) return set_err_bits("multipole split phase", herr);

    // phase 2, repeated with larger per-workgroup list arenas if one overflows
    const int nwg = (int)std::max<unsigned int>(1u, std::min<unsigned int>(hn, 512u));
    DBuf<int> arena, stk;
    DBuf<double> sstack, slots, red;
    DBuf<unsigned long long> stage_count;
    // No valid walk takes more stages than this: each stage accepts, splits or drops
    // at least one (c1, c2) pair, and a pair's cells are refined at most to the leaves.
    const unsigned long long stage_budget = 256ull * (unsigned long long)nnodes + (1ull << 20);
    if (stage_count.alloc(1) || stk.alloc((size_t)nwg * MP_STACK * 8) ||
        sstack.alloc((size_t)nwg * MP_STACK * Lo.total) ||
        slots.alloc((size_t)nwg * NOUT3 * ntot) || red.alloc((size_t)NOUT3 * ntot))
        return 2;
    KernelStat& ks = kernel_stat(KS_MULTIPOLE);
    const bool timing = timing_enabled();
    const long long nred = (long long)NOUT3 * ntot;
    long long arena_cap = MP_ARENA0;
    for (;;) {
        if (arena.alloc((size_t)nwg * arena_cap)) return 2;
        TC_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
        TC_HIP(hipMemsetAsync(counters.p, 0, sizeof(unsigned long long) * NC3, st));
        TC_HIP(hipMemsetAsync(next_unit.p, 0, sizeof(unsigned int), st));
        TC_HIP(hipMemsetAsync(stage_count.p, 0, sizeof(unsigned long long), st));
        TC_HIP(hipMemsetAsync(slots.p, 0, sizeof(double) * (size_t)nwg * NOUT3 * ntot, st));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing) {
            TC_HIP(hipEventCreate(&e0));
            TC_HIP(hipEventCreate(&e1));
            TC_HIP(hipEventRecord(e0, st));
        }
        if (hn > 0) {
            hipLaunchKernelGGL(ffn, dim3(nwg), dim3(MP_THREADS), shbytes, st, nodes, nnodes,
                               nsum.p, pool.p, pool_cap, units.p, hn, next_unit.p, stage_count.p,
                               stage_budget, P, arena.p, arena_cap, sstack.p, stk.p, slots.p, ntot,
                               counters.p, err.p);
            TC_HIP(hipGetLastError());
        }
        if (timing) TC_HIP(hipEventRecord(e1, st));
        TC_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (timing) {
            float ms = 0.f;
            TC_HIP(hipEventElapsedTime(&ms, e0, e1));
            ks.ms += ms;
            ks.launches += 1;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        if (herr == E_ARENA && arena_cap < (1ll << 27)) {
            arena_cap *= 4;
            continue;
        }
        break;
    }
    hipLaunchKernelGGL(k_mp_reduce, dim3((unsigned)((nred + 255) / 256)), dim3(256), 0, st,
                       slots.p, nwg, nred, red.p);
    TC_HIP(hipGetLastError());
    TC_HIP(hipMemcpyAsync(out, red.p, sizeof(double) * nred, hipMemcpyDeviceToHost, st));
    TC_HIP(hipMemcpyAsync(cnt, counters.p, sizeof(unsigned long long) * NC3,
                          hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    if (herr) return set_err_bits("multipole finish phase", herr);
    return 0;
}

}  // namespace tc
