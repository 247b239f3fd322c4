// treecorr_amd: the 2-point dual-tree traversal and pair binning on the GPU.
//
// TreeCorr recurses process2/process11 depth first on every CPU thread
// (src/Corr2.cpp:146-389).  Here the recursion becomes a work stack of node-pair
// items processed in large chunks: every launch pops up to CHUNK items, each lane
// applies the process11 decision (prune / accept into one bin / split 2 or 4
// ways), accepted pairs are binned on the spot (directProcess11 + finishProcess,
// src/Corr2.cpp:476-566) into a per-workgroup LDS histogram, and the children of
// split pairs are written in item order to the workgroup's own output region; the
// next launch reads them there.  Popping the newest items first keeps the stack
// depth bounded like the reference's recursion.
//
// Accumulation: npairs in exact u64, every other sum in FP64 -- LDS atomics
// within a launch, Neumaier-compensated per-workgroup slots across launches, and
// a fixed-order compensated reduction over slots at the end.
//
// Patch batching (the jobs of a patch-based covariance, treecorr/corr2base.py:
// 836-1050): the trees of all patches are concatenated, every item carries the
// index of its patch pair, and all patch pairs run in the same launches; each
// workgroup keeps LDS histograms for a few patch pairs at a time and adds them
// to per-pair accumulators in HBM.
#include "tc_engine.h"

#include <algorithm>
#include <cstring>
#include <atomic>
#include <cmath>
#include <type_traits>
#include <vector>

namespace tc {


enum { Q_XI0 = 0, Q_XI1, Q_XI2, Q_XI3, Q_MEANR, Q_MEANLOGR, Q_WEIGHT, NQ2 = 7 };
enum { CNT_VISITS = 0, CNT_ACCEPTED, CNT_LEAF, CNT_REFINED, NCNT = 4 };
enum { CL_NONE = 0, CL_SPLIT, CL_SINGLE, CL_REFINE };   // process11 outcome classes

constexpr int TRAV_THREADS = 256;
constexpr int TRAV_SLOTS = 2048;          // persistent workgroups = accumulation slots
// Items per launch (buffers are sized for it: two output buffers of 4 x TRAV_CHUNK
// items, 4.3 GB for 8-byte items).  2^26 measured 3.5% faster than 2^25 on the 2M GG
// workload (half the launches, less copying onto the stack); 2^27 gained 1% more.
constexpr long long TRAV_CHUNK = 1LL << 26;
// Persistent launch control words: job counter, overflow flag, the launch's counters.
constexpr int PCTL_WORDS = 2 + 4;
// Jobs a persistent launch wants to start from: the initial node pairs are split
// breadth first by chunked launches until there are this many, so that no job's subtree
// is more than a small part of one workgroup's share (measured on GG 2M: 2^20 jobs
// 158 ms, 2^22 78 ms, 2^24 43.4 ms, 2^26 41.2 ms of traversal; chunked launches 42.5).
constexpr long long PERSIST_MIN_JOBS = 1LL << 26;
static long long g_chunk = TRAV_CHUNK;        // tc_set_traversal_chunk, for tests and tuning

template <int C>
__device__ __forceinline__ double dist_sq(const DNode& a, const DNode& b) {
    // MetricHelper<Euclidean>::DistSq, include/Metric.h:143-147, :197-202
    double dx = a.x - b.x, dy = a.y - b.y;
    if (C == Flat) return dx * dx + dy * dy;
    double dz = a.z - b.z;
    return dx * dx + dy * dy + dz * dz;
}


#ifndef TC_FAST_LOG
#define TC_FAST_LOG 1
#endif


// Natural logarithm for the bin index and meanlogr: the published fdlibm method
// (Sun's e_log.c): x = 2^k (1+f) with 1+f in [sqrt(2)/2, sqrt(2)), s = f/(2+f),
// log(1+f) = f - f^2/2 + s (f^2/2 + R(s^2)) with R the degree-7 minimax polynomial of
// that method, and k ln2 split in two parts.  Error below 1 ulp; about a third of the
// instructions of the device library's double-double log.  A bin decision that a
// last-bit difference from the host libm's log could turn over is taken again with
// log_cr.  Only called with finite x > 0 (squared distances >= minsep^2).
__device__ __forceinline__ double tc_log(double x) {
#if TC_FAST_LOG
    int e;
    double m = frexp(x, &e);           // x = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752440) { m *= 2.; --e; }
    const double f = m - 1.;
    const double t = 2. + f;
    double inv = __builtin_amdgcn_rcp(t);   // t in [1.7, 2.5): two Newton steps
    inv = fma(fma(-t, inv, 1.), inv, inv);
    inv = fma(fma(-t, inv, 1.), inv, inv);
    const double sv = f * inv;
    const double z = sv * sv, w = z * z;
    const double t1 = w * fma(w, fma(w, 1.531383769920937332e-01, 2.222219843214978396e-01),
                              3.999999999940941908e-01);
    const double t2 = z * fma(w, fma(w, fma(w, 1.479819860511658591e-01,
                                            1.818357216161805012e-01),
                                     2.857142874366239149e-01),
                              6.666666666666735130e-01);
    const double R = t2 + t1;
    const double hfsq = 0.5 * f * f;
    const double dk = double(e);
    return dk * 6.93147180369123816490e-01 -
           ((hfsq - (sv * (hfsq + R) + dk * 1.90821492927058770002e-10)) - f);
#else
    return log(x);
#endif
}

// ProjectHelper<Sphere>::calculate_direction, include/ProjectHelper.h:272-345.  dsq is
// |p1 - p2|^2 of the (unit) positions, which the caller has already computed as the
// pair's rsq with the same expression.
__device__ __forceinline__ void sphere_direction(double x1, double y1, double z1, double x2,
                                                 double y2, double z2, double dsq, double& re,
                                                 double& im) {
    double cosA = (z1 - z2) + 0.5 * z2 * dsq;
    double sinA = y1 * x2 - x1 * y2;
    re = sinA;
    im = -cosA;
}

// sqrt(x) for meanr (finite x > 0, not denormal): the hardware reciprocal square root
// estimate, two Goldschmidt steps and a final Newton correction, within an ulp of the
// correctly rounded root.  No decision depends on it.
__device__ __forceinline__ double sqrt_fast(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g); h = fma(h, e, h);
    e = fma(-g, h, 0.5);
    g = fma(g, e, g); h = fma(h, e, h);
    const double d = fma(-g, g, x);
    return fma(d, h, g);
}

// 1/x for the projection factors (finite x > 0): hardware reciprocal estimate + two
// Newton steps, within an ulp of the division.
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.), r, r);
    return fma(fma(-x, r, 1.), r, r);
}

// calculate_expmsialpha<2> (include/ProjectHelper.h:45-47) for the shear projections:
// conj(r^2)/|r|^2 with a reciprocal instead of two divisions.  The projected shears
// then differ from the reference's by an ulp or two, inside the xi tolerance.
__device__ __forceinline__ void expm2ialpha_proj(double re, double im, double& er, double& ei) {
    double n = re * re + im * im;
    n = n > 0. ? n : 1.;
    const double inv = recip(n);
    er = (re * re - im * im) * inv;
    ei = -(re * im + im * re) * inv;
}

// The two projection factors of a pair with one reciprocal.
__device__ __forceinline__ void expm2ialpha_proj2(double r1, double i1, double r2, double i2,
                                                  double& e1r, double& e1i, double& e2r,
                                                  double& e2i) {
    double n1 = r1 * r1 + i1 * i1, n2 = r2 * r2 + i2 * i2;
    n1 = n1 > 0. ? n1 : 1.;
    n2 = n2 > 0. ? n2 : 1.;
    const double inv = recip(n1 * n2);
    const double inv1 = n2 * inv, inv2 = n1 * inv;
    e1r = (r1 * r1 - i1 * i1) * inv1;
    e1i = -(r1 * i1 + i1 * r1) * inv1;
    e2r = (r2 * r2 - i2 * i2) * inv2;
    e2i = -(r2 * i2 + i2 * r2) * inv2;
}

__device__ __forceinline__ void cmul(double& zr, double& zi, double er, double ei) {
    double r = zr * er - zi * ei;
    double i = zr * ei + zi * er;
    zr = r; zi = i;
}

// BaseCorr2::directProcess11 + Corr2::finishProcess + DirectHelper<0,1,3>
// (src/Corr2.cpp:394-566) for one accepted pair, accumulated into LDS.
// hist and hnp point at this lane's copy of the workgroup histogram; copies are
// interleaved with stride R (entry e of copy c lives at e*R + c).
// GM: rsq comes from a metric other than Euclidean (the Arc metric's angle squared), so
// the sphere projections take the chord length from the positions.
template <int D, int C, bool GM>
__device__ __forceinline__ void bin_pair(const DNode& c1, const DNode& c2, double rsq, int k,
                                         double r, double logr, const C2Params& P,
                                         double* hist, unsigned long long* hnp, int R) {
    const int nb = P.nbins;
    if (k == nb) --k;
    double ww = double(c1.w) * double(c2.w);
    atomicAdd(&hist[(Q_WEIGHT * nb + k) * R], ww);
    if (!P.quick) {
        atomicAdd(&hnp[k * R], (unsigned long long)c1.n * (unsigned long long)c2.n);
        atomicAdd(&hist[(Q_MEANR * nb + k) * R], ww * r);
        atomicAdd(&hist[(Q_MEANLOGR * nb + k) * R], ww * logr);
    }
    if (D == KData) {
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], double(c1.a) * double(c2.a));
    } else if (D == GData) {
        double g1r = c1.a, g1i = c1.b, g2r = c2.a, g2i = c2.b;
        double er, ei;
        if (C == Flat) {
            // ProjectHelper<Flat>::Project, include/ProjectHelper.h:126-136
            expm2ialpha_proj(c2.x - c1.x, c2.y - c1.y, er, ei);
            cmul(g1r, g1i, er, ei);
            cmul(g2r, g2i, er, ei);
        } else {
            // ProjectHelper<Sphere>::Project :370-388; ThreeD normalises first :573-589
            double x1 = c1.x, y1 = c1.y, z1 = c1.z, x2 = c2.x, y2 = c2.y, z2 = c2.z;
            // Sphere positions are unit vectors already: the chord is their separation
            double dsq = GM ? dist_sq<Sphere>(c1, c2) : rsq;
            if (C == ThreeD) {
                normalize_recip(x1, y1, z1);
                normalize_recip(x2, y2, z2);
                const double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                dsq = dx * dx + dy * dy + dz * dz;
            }
            double d2r, d2i, d1r, d1i, e1r, e1i;
            sphere_direction(x1, y1, z1, x2, y2, z2, dsq, d2r, d2i);
            sphere_direction(x2, y2, z2, x1, y1, z1, dsq, d1r, d1i);
            expm2ialpha_proj2(d2r, d2i, d1r, d1i, er, ei, e1r, e1i);
            cmul(g2r, g2i, er, ei);
            cmul(g1r, g1i, e1r, e1i);
        }
        // both_complex_prod(g2, g1), include/ProjectHelper.h:84-101
        double rr = g2r * g1r, ri = g2r * g1i, ir = g2i * g1r, ii = g2i * g1i;
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], rr + ii);   // xip
        atomicAdd(&hist[(Q_XI1 * nb + k) * R], ri - ir);   // xip_im
        atomicAdd(&hist[(Q_XI2 * nb + k) * R], rr - ii);   // xim
        atomicAdd(&hist[(Q_XI3 * nb + k) * R], ri + ir);   // xim_im
    } else if (D == VData || D == TData || D == QData) {
        // DirectHelper<3> for spin 1, 3, 4 (the shear above has the fast form of spin 2):
        // ProjectHelper<C>::Project(c1,c2,z1,z2), include/ProjectHelper.h:138-147,
        // :370-388, :573-589; odd spins turn z1 by pi (p1 on the left)
        constexpr int S = spin_of(D);
        double g1r = c1.a, g1i = c1.b, g2r = c2.a, g2i = c2.b;
        if (C == Flat) {
            double er, ei;
            expmsialpha<S>(c2.x - c1.x, c2.y - c1.y, er, ei);
            cmul(g1r, g1i, er, ei);
            cmul(g2r, g2i, er, ei);
        } else {
            double x1 = c1.x, y1 = c1.y, z1 = c1.z, x2 = c2.x, y2 = c2.y, z2 = c2.z;
            double dsq = GM ? dist_sq<Sphere>(c1, c2) : rsq;
            if (C == ThreeD) {
                normalize_recip(x1, y1, z1);
                normalize_recip(x2, y2, z2);
                const double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                dsq = dx * dx + dy * dy + dz * dz;
            }
            double d2r, d2i, d1r, d1i, e2r, e2i, e1r, e1i;
            sphere_direction(x1, y1, z1, x2, y2, z2, dsq, d2r, d2i);
            sphere_direction(x2, y2, z2, x1, y1, z1, dsq, d1r, d1i);
            expmsialpha<S>(d2r, d2i, e2r, e2i);
            expmsialpha<S>(d1r, d1i, e1r, e1i);
            cmul(g2r, g2i, e2r, e2i);
            cmul(g1r, g1i, e1r, e1i);
            if (S == 1 || S == 3) { g1r = -g1r; g1i = -g1i; }
        }
        double rr = g2r * g1r, ri = g2r * g1i, ir = g2i * g1r, ii = g2i * g1i;
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], rr + ii);   // xip
        atomicAdd(&hist[(Q_XI1 * nb + k) * R], ri - ir);   // xip_im
        atomicAdd(&hist[(Q_XI2 * nb + k) * R], rr - ii);   // xim
        atomicAdd(&hist[(Q_XI3 * nb + k) * R], ri + ir);   // xim_im
    } else if (D == XK_NV || D == XK_NT || D == XK_NQ || D == XK_KV || D == XK_KT ||
               D == XK_KQ) {
        // DirectHelper<2> with a spin 1, 3 or 4 second field: z2 projected onto the line
        // to c1, times +getWK(c1)
        constexpr int S = spin_of(mixed_d2(D));
        double g2r = c2.a, g2i = c2.b;
        double er, ei;
        if (C == Flat) {
            expmsialpha<S>(c2.x - c1.x, c2.y - c1.y, er, ei);
        } else {
            double x1 = c1.x, y1 = c1.y, z1 = c1.z, x2 = c2.x, y2 = c2.y, z2 = c2.z;
            double dsq = GM ? dist_sq<Sphere>(c1, c2) : rsq;
            if (C == ThreeD) {
                normalize_recip(x1, y1, z1);
                normalize_recip(x2, y2, z2);
                const double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                dsq = dx * dx + dy * dy + dz * dz;
            }
            double d2r, d2i;
            sphere_direction(x1, y1, z1, x2, y2, z2, dsq, d2r, d2i);
            expmsialpha<S>(d2r, d2i, er, ei);
        }
        cmul(g2r, g2i, er, ei);
        const double f = mixed_k1(D) ? double(c1.a) : double(c1.w);
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], g2r * f);
        atomicAdd(&hist[(Q_XI1 * nb + k) * R], g2i * f);
    } else if (D == ZData) {
        // DirectHelper<3> for spin 0: no projection (exp(-i 0 alpha) = 1)
        const double g1r = c1.a, g1i = c1.b, g2r = c2.a, g2i = c2.b;
        double rr = g2r * g1r, ri = g2r * g1i, ir = g2i * g1r, ii = g2i * g1i;
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], rr + ii);   // xip
        atomicAdd(&hist[(Q_XI1 * nb + k) * R], ri - ir);   // xip_im
        atomicAdd(&hist[(Q_XI2 * nb + k) * R], rr - ii);   // xim
        atomicAdd(&hist[(Q_XI3 * nb + k) * R], ri + ir);   // xim_im
    } else if (D == XK_NZ || D == XK_KZ) {
        // DirectHelper<2> for a spin-0 second field: z2 unprojected, times +getWK(c1)
        const double f = D == XK_NZ ? double(c1.w) : double(c1.a);
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], double(c2.a) * f);
        atomicAdd(&hist[(Q_XI1 * nb + k) * R], double(c2.b) * f);
    } else if (D == XK_NK) {
        // DirectHelper<1>, src/Corr2.cpp:408-420; getWK of an N cell is its w
        // (include/ProjectHelper.h:75-80)
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], double(c1.w) * double(c2.a));
    } else if (D == XK_NG || D == XK_KG) {
        // DirectHelper<2>, src/Corr2.cpp:423-442: c2's shear projected onto the line to
        // c1 (ProjectHelper<C>::Project(c1,c2,z2), include/ProjectHelper.h:107-113,
        // :348-355, :545-555), times -getWK(c1)
        double g2r = c2.a, g2i = c2.b;
        double er, ei;
        if (C == Flat) {
            expm2ialpha_proj(c2.x - c1.x, c2.y - c1.y, er, ei);
        } else {
            double x1 = c1.x, y1 = c1.y, z1 = c1.z, x2 = c2.x, y2 = c2.y, z2 = c2.z;
            double dsq = GM ? dist_sq<Sphere>(c1, c2) : rsq;
            if (C == ThreeD) {
                normalize_recip(x1, y1, z1);
                normalize_recip(x2, y2, z2);
                const double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                dsq = dx * dx + dy * dy + dz * dz;
            }
            double d2r, d2i;
            sphere_direction(x1, y1, z1, x2, y2, z2, dsq, d2r, d2i);
            expm2ialpha_proj(d2r, d2i, er, ei);
        }
        cmul(g2r, g2i, er, ei);
        const double f = -(D == XK_NG ? double(c1.w) : double(c1.a));
        atomicAdd(&hist[(Q_XI0 * nb + k) * R], g2r * f);
        atomicAdd(&hist[(Q_XI1 * nb + k) * R], g2i * f);
    }
}

// A node pair of one patch pair in a batched run: p indexes the patch-pair list.
struct alignas(16) ItemP { int i, j, p, pad; };

template <class IT> struct ItemOps;
template <> struct ItemOps<Item> {
    static constexpr bool batched = false;
    __device__ static Item make(int i, int j, const Item&) { return Item{i, j}; }
    __device__ static int pair(const Item&) { return 0; }
};
template <> struct ItemOps<ItemP> {
    static constexpr bool batched = true;
    __device__ static ItemP make(int i, int j, const ItemP& parent) {
        return ItemP{i, j, parent.p, 0};
    }
    __device__ static int pair(const ItemP& it) { return it.p; }
};

// A node pair of a logged run: job is the index of the job (initial node pair) whose
// subtree it belongs to, so that every job's visits can be counted.
struct alignas(16) ItemJ { int i, j, job, pad; };
template <> struct ItemOps<ItemJ> {
    static constexpr bool batched = false;
    __device__ static ItemJ make(int i, int j, const ItemJ& parent) {
        return ItemJ{i, j, parent.job, 0};
    }
    __device__ static int pair(const ItemJ&) { return 0; }
};
template <class IT> struct IsLogged { static constexpr bool value = false; };
template <> struct IsLogged<ItemJ> { static constexpr bool value = true; };

constexpr int PB_SLOTS = 4;   // patch pairs with an LDS histogram per workgroup

// Where a launch reads its items.  Items [0, nhead) come from the output regions of
// the previous launch, read in place: region r holds off[r+1]-off[r] items at
// in + r*stride, and the launch takes their concatenated items [first, first+nhead).
// Items [nhead, nin) come from the contiguous run tail (the top of the stack).
template <class IT>
struct InView {
    const IT* in;
    const unsigned long long* off;
    long long stride, first, nhead;
    int nreg;
    const IT* tail;
};

// LDS histogram slot of patch pair p in this workgroup (claimed on first use), or -1
// when all PB_SLOTS are taken by other pairs.
__device__ __forceinline__ int pair_slot(int p, int* tags) {
#pragma unroll
    for (int s = 0; s < PB_SLOTS; ++s)
        if (tags[s] == p) return s;
#pragma unroll
    for (int s = 0; s < PB_SLOTS; ++s) {
        const int old = atomicCAS(&tags[s], -1, p);
        if (old == -1 || old == p) return s;
    }
    return -1;
}

// The per-workgroup state of a traversal round: this lane's copy of the LDS histogram
// (and of the patch-pair histograms of a batched run), and the visit counters.
template <class IT>
struct VisitCtx {
    double* hist;
    unsigned long long* hnp;
    int R;
    size_t per_slot;
    int* s_tag;
    double* pair_hist;
    unsigned long long* pair_np;
    unsigned int c_vis, c_acc, c_leaf, c_ref;
    int lane;
    unsigned int* job_visits = nullptr;   // a logged run: items processed per job
};

// One round of the traversal for one lane: the process2 / process11 decision for the
// lane's item (valid: the lane has one), the accepted pair binned into the LDS
// histogram, and the children of a split pair written with the rest of the wavefront's
// children to bout[*s_pos ...] (s_pos an LDS counter the wavefront claims room from).
template <int D, int C, class IT, bool GM>
__device__ __forceinline__ void visit_round(VisitCtx<IT>& X, const C2Params& P, const IT it,
                                            const DNode& c1, const DNode& c2, bool valid,
                                            IT* __restrict__ bout, unsigned int* s_pos) {
    using Ops = ItemOps<IT>;
    IT o0{}, o1{}, o2{}, o3{};
    int no = 0;
    if constexpr (IsLogged<IT>::value)
        if (valid) atomicAdd(&X.job_visits[it.job], 1u);
    if (valid) {
        if (it.j < 0) {
            // BaseCorr2::process2, src/Corr2.cpp:290-306
            const DNode& c = c1;
            if (c.w != 0.f && double(c.size) > P.halfminsep) {
                const int L = c.child;
                o0 = Ops::make(L, -1, it); o1 = Ops::make(L + 1, -1, it);
                o2 = Ops::make(L, L + 1, it);
                no = 3;
            }
        } else {
            // BaseCorr2::process11, src/Corr2.cpp:308-389.  The decision is taken in
            // three steps so that every lane needs at most one sqrt, one log and one
            // bin-index division, evaluated once per wavefront (the reference's
            // expressions, unchanged: 0.5*log(rsq) for singleBin's refined test,
            // log(sqrt(rsq)) for directProcess11).
            ++X.c_vis;
            // step 1: pruning and singleBin's cheap tests (include/BinType.h:93-151)
            int cls = CL_NONE;
            double rsq = 0., s1ps2 = 0., s1ps2sq = 0.;
            double s1 = c1.size, s2 = c2.size;    // some metrics rescale them
            if (c1.w != 0.f && c2.w != 0.f) {
                rsq = GM ? metric_dist_sq(P.metric, C, c1.x, c1.y, c1.z, c2.x, c2.y,
                                          c2.z, P.xp, P.yp, P.zp, s1, s2)
                         : dist_sq<C>(c1, c2);
                s1ps2 = s1 + s2;
                bool too_small = rsq < P.minsepsq && s1ps2 < P.minsep &&
                                 rsq < (P.minsep - s1ps2) * (P.minsep - s1ps2);
                bool too_large = rsq >= P.maxsepsq &&
                                 rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
                bool rin = true;
                if (GM && (P.rpar_on || los_metric(P.metric))) {
                    too_small = metric_prune(P.metric, P.rpar_on, P.minrpar, P.maxrpar,
                                             c1.x, c1.y, c1.z, c2.x, c2.y, c2.z, rsq,
                                             s1ps2, too_small, too_large, P.minsep,
                                             P.minsepsq, P.maxsep, P.maxsepsq, rin);
                    too_large = false;
                }
                if (!too_small && !too_large && !rin) {
                    cls = CL_SPLIT;     // singleBin needs r_par inside the range
                } else if (!too_small && !too_large) {
                    s1ps2sq = s1ps2 * s1ps2;
                    if (s1ps2 == 0.) cls = CL_SINGLE;
                    else if (s1ps2sq > P.asq * rsq) cls = CL_SPLIT;
                    else if (s1ps2sq <= P.bsq * rsq) cls = CL_SINGLE;
                    else if (s1ps2sq > 0.25 * ((P.binsize + P.b) * (P.binsize + P.b)) * rsq)
                        cls = CL_SPLIT;
                    else cls = CL_REFINE;
                }
            }
            const bool inrange = rsq >= P.minsepsq && rsq < P.maxsepsq;
            // step 2: the log and the bin coordinate kk for the lanes that need them.
            // Both classes take 0.5 log(rsq) times the rounded 1/binsize; away from an
            // edge and from the refined test's thresholds that decides as the
            // reference's expressions do (singleBin's 0.5*log(rsq), directProcess11's
            // log(sqrt(rsq)), each divided by binsize).
            double r = 0., logr = 0., kk = 0.;
            int k = 0;
            if (cls == CL_REFINE || (cls == CL_SINGLE && inrange)) {
                const double L = tc_log(rsq);
                logr = 0.5 * L;
                kk = (logr - P.logminsep) * P.inv_binsize;
                // step 3: the refined test, then accept or split.  A decision within
                // a few ulp of a bin edge or of the refined test's threshold is taken
                // again by the reference's own expression with the correctly rounded
                // log, as the host libm rounds it, and the division.
                const double sl = kk_slack(logr, P.logminsep, P.inv_binsize, kk);
                bool near = near_int(kk, sl), fits = true;
                if (cls == CL_REFINE)
                    fits = refined_fits(kk, s1ps2sq, rsq, P.binsize, P.b, sl, near);
                if (near) {
                    logr = cls == CL_REFINE ? 0.5 * log_cr(rsq, L)
                                            : log_cr(sqrt(rsq), logr);
                    kk = (logr - P.logminsep) / P.binsize;
                    if (cls == CL_REFINE)
                        fits = refined_fits(kk, s1ps2sq, rsq, P.binsize, P.b, 0., near);
                }
                k = int(kk);
                if (cls == CL_REFINE) {
                    ++X.c_ref;
                    cls = fits ? CL_SINGLE : CL_SPLIT;
                }
            }
            if (cls == CL_SINGLE && inrange) {
                r = sqrt_fast(rsq);
                ++X.c_acc;
                if (c1.child < 0 && c2.child < 0) ++X.c_leaf;
                if (!Ops::batched) {
                    bin_pair<D, C, GM>(c1, c2, rsq, k, r, logr, P, X.hist, X.hnp, X.R);
                } else {
                    const int pp = Ops::pair(it);
                    const int sl = pair_slot(pp, X.s_tag);
                    if (sl >= 0)
                        bin_pair<D, C, GM>(c1, c2, rsq, k, r, logr, P, X.hist + sl * X.per_slot,
                                       X.hnp + sl * X.per_slot, X.R);
                    else   // all LDS slots taken: straight into the pair's HBM sums
                        bin_pair<D, C, GM>(c1, c2, rsq, k, r, logr, P,
                                       X.pair_hist + (size_t)pp * 2 * NQ2 * P.nbins,
                                       X.pair_np + (size_t)pp * P.nbins, 1);
                }
            } else if (cls == CL_SPLIT) {
                // CalcSplitSq, include/Split.h:82-101
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
                if (sp1 && sp2) {
                    const int L1 = c1.child, L2 = c2.child;
                    o0 = Ops::make(L1, L2, it); o1 = Ops::make(L1, L2 + 1, it);
                    o2 = Ops::make(L1 + 1, L2, it); o3 = Ops::make(L1 + 1, L2 + 1, it);
                    no = 4;
                } else if (sp1) {
                    o0 = Ops::make(c1.child, it.j, it); o1 = Ops::make(c1.child + 1, it.j, it);
                    no = 2;
                } else {
                    o0 = Ops::make(it.i, c2.child, it); o1 = Ops::make(it.i, c2.child + 1, it);
                    no = 2;
                }
            }
        }
    }
    // Wavefront compaction of the children: no <= 4 is three bits, so the lane's
    // exclusive prefix is sum_k 2^k (lanes below with bit k set), from three votes.
    const unsigned long long m0 = __ballot(no & 1), m1 = __ballot(no & 2),
                             m2 = __ballot(no & 4);
    const unsigned int wtot = __popcll(m0) + 2 * __popcll(m1) + 4 * __popcll(m2);
    if (wtot != 0) {      // uniform over the wavefront
        unsigned int wbase = 0;
        if (X.lane == 0) wbase = atomicAdd(s_pos, wtot);
        wbase = (unsigned int)__builtin_amdgcn_readlane((int)wbase, 0);
        auto below = [](unsigned long long m) {
            return __builtin_amdgcn_mbcnt_hi((unsigned int)(m >> 32),
                                             __builtin_amdgcn_mbcnt_lo((unsigned int)m, 0u));
        };
        const unsigned int pos = wbase + below(m0) + 2 * below(m1) + 4 * below(m2);
        if (no > 0) bout[pos] = o0;
        if (no > 1) bout[pos + 1] = o1;
        if (no > 2) bout[pos + 2] = o2;
        if (no > 3) bout[pos + 3] = o3;
    }
}

// Adds each used patch-pair histogram of the workgroup to that pair's sums in HBM.  A
// pair's sums are a (hi, lo) pair of arrays: the exact rounding error of every addition
// into hi (TwoSum of the value it replaced and the addend) goes into lo.
__device__ __forceinline__ void flush_pair_hists(const double* hist_all,
                                                 const unsigned long long* hnp_all, int nh,
                                                 int nb, int R, size_t per_slot,
                                                 const int* s_tag, double* __restrict__ pair_hist,
                                                 unsigned long long* __restrict__ pair_np) {
    for (int sl = 0; sl < PB_SLOTS; ++sl) {
        const int pp = s_tag[sl];
        if (pp < 0) continue;
        const double* hs = hist_all + sl * per_slot;
        const unsigned long long* ns = hnp_all + sl * per_slot;
        double* ph = pair_hist + (size_t)pp * 2 * nh;
        for (int t = threadIdx.x; t < nh; t += blockDim.x) {
            double v = 0.;
            for (int c = 0; c < R; ++c) v += hs[(size_t)t * R + c];
            if (v == 0.) continue;
            const double old = atomicAdd(&ph[t], v);
            const double sum = old + v;
            const double bv = sum - old;
            const double err = (old - (sum - bv)) + (v - bv);
            if (err != 0.) atomicAdd(&ph[nh + t], err);
        }
        for (int t = threadIdx.x; t < nb; t += blockDim.x) {
            unsigned long long v = 0;
            for (int c = 0; c < R; ++c) v += ns[(size_t)t * R + c];
            if (v) atomicAdd(&pair_np[(size_t)pp * nb + t], v);
        }
    }
}

// End of a traversal launch: the workgroup's visit counters and LDS histograms go to
// its accumulation slot (a single correlation: Neumaier-compensated, slot blockIdx.x)
// or, in a batched run, to the patch pairs' sums in HBM.
template <class IT>
__device__ __forceinline__ void flush_sums(double* hist_all, unsigned long long* hnp_all,
                                           unsigned long long* hcnt, int nh, int nb, int R,
                                           size_t per_slot, const int* s_tag,
                                           double* __restrict__ slot_hi,
                                           double* __restrict__ slot_lo,
                                           unsigned long long* __restrict__ slot_np,
                                           unsigned long long* __restrict__ slot_cnt,
                                           double* __restrict__ pair_hist,
                                           unsigned long long* __restrict__ pair_np) {
    using Ops = ItemOps<IT>;
    if (Ops::batched) {
        if (threadIdx.x < NCNT) slot_cnt[(size_t)blockIdx.x * NCNT + threadIdx.x] += hcnt[threadIdx.x];
        flush_pair_hists(hist_all, hnp_all, nh, nb, R, per_slot, s_tag, pair_hist, pair_np);
        return;
    }    // Flush into this workgroup's slot with Neumaier compensation.
    const size_t so = (size_t)blockIdx.x * nh;
    for (int t = threadIdx.x; t < nh; t += blockDim.x) {
        double v = 0.;
        for (int c = 0; c < R; ++c) v += hist_all[(size_t)t * R + c];
        if (v == 0.) continue;
        const double s = slot_hi[so + t];
        const double tt = s + v;
        const double c = (fabs(s) >= fabs(v)) ? (s - tt) + v : (v - tt) + s;
        slot_hi[so + t] = tt;
        slot_lo[so + t] += c;
    }
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        unsigned long long v = 0;
        for (int c = 0; c < R; ++c) v += hnp_all[(size_t)t * R + c];
        slot_np[(size_t)blockIdx.x * nb + t] += v;
    }
    if (threadIdx.x < NCNT) slot_cnt[(size_t)blockIdx.x * NCNT + threadIdx.x] += hcnt[threadIdx.x];
}

// Workgroup b takes the items [b*per_block, (b+1)*per_block) of the chunk and writes
// the children of its split pairs to its own region out[b*4*per_block ...] (at most
// four children per item); bcount[b] gets their number.  Each wavefront claims room
// for its children with one LDS atomic, so the waves of a workgroup never wait for
// each other; the order of the children in a region does not matter.  No global
// atomics: the regions are concatenated afterwards.
// GM: the other metrics and the r_par range (metric_dist_sq, metric_prune); the Euclidean instantiations
// keep the plain distance, whose kernels need far fewer registers.
template <int D, int C, class IT, bool GM>
__global__ __launch_bounds__(TRAV_THREADS) void k_traverse(
    const DNode* __restrict__ n1, const DNode* __restrict__ n2, InView<IT> V,
    long long nin, long long per_block, IT* __restrict__ out,
    unsigned int* __restrict__ bcount, C2Params P, int R,
    double* __restrict__ slot_hi, double* __restrict__ slot_lo,
    unsigned long long* __restrict__ slot_np, unsigned long long* __restrict__ slot_cnt,
    double* __restrict__ pair_hist, unsigned long long* __restrict__ pair_np) {
    using Ops = ItemOps<IT>;
    HIP_DYNAMIC_SHARED(double, sh)
    __shared__ unsigned int s_pos;   // children written by this workgroup so far
    // region offsets of the in-place head (a launch's output stays below 2^32 items)
    __shared__ unsigned int s_off[TRAV_SLOTS + 1];
    __shared__ int s_tag[PB_SLOTS];
    const int nb = P.nbins;
    const int nh = NQ2 * nb;
    // R interleaved copies of the histogram (R a power of two): lanes that land in the
    // same bin in one instruction hit R different LDS banks instead of one address.
    // A batched run holds PB_SLOTS such histograms, one per patch pair.
    const int nslot = Ops::batched ? PB_SLOTS : 1;
    const size_t per_slot = (size_t)(nh + nb) * R;
    unsigned long long* hcnt = reinterpret_cast<unsigned long long*>(sh + per_slot * nslot);
    for (int t = threadIdx.x; t < (int)(per_slot * nslot) + NCNT; t += blockDim.x)
        reinterpret_cast<unsigned long long*>(sh)[t] = 0ull;
    if (V.nhead > 0)
        for (int t = threadIdx.x; t <= V.nreg; t += blockDim.x) s_off[t] = (unsigned int)V.off[t];
    if (threadIdx.x < PB_SLOTS) s_tag[threadIdx.x] = -1;
    if (threadIdx.x == 0) s_pos = 0;
    __syncthreads();

    const int lane = threadIdx.x & 63;
    double* hist_all = sh;
    unsigned long long* hnp_all = reinterpret_cast<unsigned long long*>(sh + (size_t)nh * R);
    double* hist = hist_all + (lane & (R - 1));
    unsigned long long* hnp = hnp_all + (lane & (R - 1));
    VisitCtx<IT> X{hist, hnp, R, per_slot, s_tag, pair_hist, pair_np, 0u, 0u, 0u, 0u, lane};
    const long long b0 = (long long)blockIdx.x * per_block;
    const long long b1 = b0 + per_block < nin ? b0 + per_block : nin;
    IT* __restrict__ bout = out + (size_t)blockIdx.x * 4 * per_block;
    // region cursor of this thread's next item (items of a thread only move forward)
    int reg = 0;
    if (b0 + threadIdx.x < b1 && b0 + threadIdx.x < V.nhead) {
        const unsigned int g = (unsigned int)(V.first + b0 + threadIdx.x);
        int lo = 0, hi = V.nreg - 1;      // last region with s_off[r] <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= g) lo = mid; else hi = mid - 1;
        }
        reg = lo;
    }
    auto fetch = [&](long long idx) {
        if (idx < V.nhead) {
            const unsigned int g = (unsigned int)(V.first + idx);
            while (s_off[reg + 1] <= g) ++reg;
            return V.in[(long long)reg * V.stride + (long long)(g - s_off[reg])];
        }
        return V.tail[idx - V.nhead];
    };
    // Software pipeline over two register sets: while a round works on one item and
    // its nodes, the nodes of the next round's item and the item two rounds ahead are
    // loaded into the other set, so their memory latency overlaps this round's work.
    // The loop is unrolled by two so that the sets alternate without register copies.
    // (A process2 item, j < 0, loads node 0 of the second tree in place of its unused
    // second node; items past the end load nothing.)
    const long long bd = blockDim.x;
    IT itA{}, itB{};
    DNode a1{}, a2{}, q1{}, q2{};
    const long long t0 = b0 + threadIdx.x;
    if (t0 < b1) {
        itA = fetch(t0);
        a1 = n1[itA.i];
        a2 = n2[itA.j >= 0 ? itA.j : 0];
    }
    if (t0 + bd < b1) itB = fetch(t0 + bd);
    for (long long base = b0; base < b1; base += 2 * bd) {
        const long long idx = base + threadIdx.x;
        // round on set A; loads for set B (the item of the next round is itB)
        const IT curA = itA;
        if (idx + bd < b1) {
            q1 = n1[itB.i];
            q2 = n2[itB.j >= 0 ? itB.j : 0];
        }
        if (idx + 2 * bd < b1) itA = fetch(idx + 2 * bd);
        visit_round<D, C, IT, GM>(X, P, curA, a1, a2, idx < b1, bout, &s_pos);
        if (base + bd >= b1) break;     // uniform over the workgroup
        // round on set B; loads for set A
        const IT curB = itB;
        if (idx + 2 * bd < b1) {
            a1 = n1[itA.i];
            a2 = n2[itA.j >= 0 ? itA.j : 0];
        }
        if (idx + 3 * bd < b1) itB = fetch(idx + 3 * bd);
        visit_round<D, C, IT, GM>(X, P, curB, q1, q2, idx + bd < b1, bout, &s_pos);
    }
    atomicAdd(&hcnt[CNT_VISITS], (unsigned long long)X.c_vis);
    atomicAdd(&hcnt[CNT_ACCEPTED], (unsigned long long)X.c_acc);
    atomicAdd(&hcnt[CNT_LEAF], (unsigned long long)X.c_leaf);
    atomicAdd(&hcnt[CNT_REFINED], (unsigned long long)X.c_ref);
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = s_pos;
    flush_sums<IT>(hist_all, hnp_all, hcnt, nh, nb, R, per_slot, s_tag, slot_hi, slot_lo,
                   slot_np, slot_cnt, pair_hist, pair_np);
}

// Persistent form of the traversal: one launch runs a whole correlation.  The grid is
// the number of workgroups the chip holds at once; each keeps a LIFO stack of items in
// its own HBM region stk[0, cap), pops up to TRAV_THREADS items a round (the newest
// first) and pushes the children of its split pairs back on top, so the stack stays
// shallow as the reference's recursion does.  When fewer than TRAV_THREADS items are
// left it takes more jobs (initial items): in dynamic mode the next ones of the job
// list by one atomic counter (as the reference's OpenMP schedule(dynamic) hands out
// top-level pairs, src/Corr2.cpp:158-203), in static mode the next ones of its own
// contiguous share of the list (equal counts, or the shares of equal predicted work
// that the cost model gives, bounds[b] to bounds[b+1]).  A workgroup leaves when its stack and its
// jobs are exhausted; no workgroup ever waits for another.  A round whose children
// might not fit the stack sets *flag and ends the workgroup instead; the caller then
// discards the sums and runs the chunked traversal.
template <int D, int C, class IT, bool GM>
__global__ __launch_bounds__(TRAV_THREADS) void k_traverse_persist(
    const DNode* __restrict__ n1, const DNode* __restrict__ n2, const IT* __restrict__ jobs,
    long long njobs, unsigned long long* __restrict__ job_next, int static_part, int grab_div,
    const long long* __restrict__ bounds, IT* __restrict__ stacks, long long cap,
    unsigned int* __restrict__ flag, C2Params P, int R, double* __restrict__ slot_hi,
    double* __restrict__ slot_lo, unsigned long long* __restrict__ slot_np,
    unsigned long long* __restrict__ slot_cnt, double* __restrict__ pair_hist,
    unsigned long long* __restrict__ pair_np, unsigned int* __restrict__ job_visits,
    unsigned long long* __restrict__ launch_cnt) {
    using Ops = ItemOps<IT>;
    HIP_DYNAMIC_SHARED(double, sh)
    __shared__ unsigned int s_pos;           // children written this round
    __shared__ long long s_size;             // items on the stack
    __shared__ long long s_g0, s_gn;         // jobs taken this round
    __shared__ long long s_snext, s_send;    // static mode: the rest of this share
    __shared__ int s_tag[PB_SLOTS];
    __shared__ int s_full;                   // a batched round found no free pair slot
    const int nb = P.nbins;
    const int nh = NQ2 * nb;
    const int nslot = Ops::batched ? PB_SLOTS : 1;
    const size_t per_slot = (size_t)(nh + nb) * R;
    unsigned long long* hcnt = reinterpret_cast<unsigned long long*>(sh + per_slot * nslot);
    for (int t = threadIdx.x; t < (int)(per_slot * nslot) + NCNT; t += blockDim.x)
        reinterpret_cast<unsigned long long*>(sh)[t] = 0ull;
    if (threadIdx.x < PB_SLOTS) s_tag[threadIdx.x] = -1;
    if (threadIdx.x == 0) {
        s_size = 0;
        // static shares: jobs [bounds[b], bounds[b+1]), or equal counts without bounds
        s_snext = bounds ? bounds[blockIdx.x] : njobs * (long long)blockIdx.x / gridDim.x;
        s_send = bounds ? bounds[blockIdx.x + 1] : njobs * (long long)(blockIdx.x + 1) / gridDim.x;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    double* hist_all = sh;
    unsigned long long* hnp_all = reinterpret_cast<unsigned long long*>(sh + (size_t)nh * R);
    VisitCtx<IT> X{hist_all + (lane & (R - 1)), hnp_all + (lane & (R - 1)), R, per_slot, s_tag,
                   pair_hist, pair_np, 0u, 0u, 0u, 0u, lane, job_visits};
    IT* __restrict__ stk = stacks + (size_t)blockIdx.x * cap;
    const long long bd = blockDim.x;
    for (;;) {
        if (threadIdx.x == 0) {
            long long g0 = 0, gn = 0;
            const long long need = bd - s_size;
            if (need > 0) {
                if (static_part) {
                    g0 = s_snext;
                    gn = need < s_send - s_snext ? need : s_send - s_snext;
                    s_snext += gn;
                } else {
                    // a batch of at most 1/grab_div of a fair share, so the jobs left
                    // near the end stay spread over the workgroups
                    long long want = njobs / ((long long)grab_div * gridDim.x);
                    want = want < 1 ? 1 : (want < need ? want : need);
                    const unsigned long long a = atomicAdd(job_next, (unsigned long long)want);
                    if (a < (unsigned long long)njobs) {
                        g0 = (long long)a;
                        gn = want < njobs - g0 ? want : njobs - g0;
                    }
                }
            }
            s_g0 = g0;
            s_gn = gn;
            s_pos = 0;
            s_full = 0;
        }
        __syncthreads();
        const long long size0 = s_size, gn = s_gn;
        if (threadIdx.x < gn) stk[size0 + threadIdx.x] = jobs[s_g0 + threadIdx.x];
        const long long size = size0 + gn;
        if (size == 0) break;                       // uniform: nothing left anywhere
        const long long n = size < bd ? size : bd;
        const long long base = size - n;
        if (base + 4 * n > cap) {                   // uniform
            if (threadIdx.x == 0) atomicOr(flag, 1u);
            break;
        }
        __syncthreads();                            // the jobs taken are on the stack
        const bool valid = threadIdx.x < n;
        IT it{};
        DNode c1{}, c2{};
        if (valid) {
            it = stk[base + threadIdx.x];
            c1 = n1[it.i];
            c2 = n2[it.j >= 0 ? it.j : 0];
        }
        __syncthreads();                            // every item read before children overwrite them
        if constexpr (Ops::batched) {
            // A workgroup meets many patch pairs over its life: when a round's items
            // find no free LDS histogram, the pairs' histograms go to HBM and the slots
            // are claimed anew (a lane still without one adds straight to HBM).
            if (valid && pair_slot(Ops::pair(it), s_tag) < 0) s_full = 1;
            __syncthreads();
            if (s_full) {                           // uniform
                flush_pair_hists(hist_all, hnp_all, nh, nb, R, per_slot, s_tag, pair_hist,
                                 pair_np);
                __syncthreads();
                for (int t = threadIdx.x; t < (int)(per_slot * nslot); t += blockDim.x)
                    reinterpret_cast<unsigned long long*>(sh)[t] = 0ull;
                if (threadIdx.x < PB_SLOTS) s_tag[threadIdx.x] = -1;
                __syncthreads();
            }
        }
        visit_round<D, C, IT, GM>(X, P, it, c1, c2, valid, stk + base, &s_pos);
        __syncthreads();
        if (threadIdx.x == 0) s_size = base + s_pos;
    }
    atomicAdd(&hcnt[CNT_VISITS], (unsigned long long)X.c_vis);
    atomicAdd(&hcnt[CNT_ACCEPTED], (unsigned long long)X.c_acc);
    atomicAdd(&hcnt[CNT_LEAF], (unsigned long long)X.c_leaf);
    atomicAdd(&hcnt[CNT_REFINED], (unsigned long long)X.c_ref);
    __syncthreads();
    // this launch's own counters too (the chunked split before it has its own)
    if (threadIdx.x < NCNT) atomicAdd(&launch_cnt[threadIdx.x], hcnt[threadIdx.x]);
    flush_sums<IT>(hist_all, hnp_all, hcnt, nh, nb, R, per_slot, s_tag, slot_hi, slot_lo,
                   slot_np, slot_cnt, pair_hist, pair_np);
}

// The cost model's features of job k (a node pair, or a cell of a process2 job), in the
// units of the binning: the two sizes and counts, the separation, the split ratio, the
// job's kind, the bin slop and width and whether the positions are 3-d (flat 0).
template <int C>
__global__ void k_job_features(const DNode* __restrict__ n1, const DNode* __restrict__ n2,
                               const Item* __restrict__ jobs, long long njobs, C2Params P,
                               float* __restrict__ feat) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= njobs) return;
    const Item it = jobs[k];
    const DNode c1 = n1[it.i];
    const bool p2 = it.j < 0;
    const DNode c2 = p2 ? c1 : n2[it.j];
    const double d = p2 ? 0. : sqrt(dist_sq<C>(c1, c2));
    const double s1 = c1.size, s2 = c2.size;
    const double dm = d > 1e-3 * P.minsep ? d : 1e-3 * P.minsep;
    float* f = feat + k * CM_NF;
    f[0] = (float)log2(1. + s1 / P.minsep);
    f[1] = (float)log2(1. + s2 / P.minsep);
    f[2] = (float)log2((double)(c1.n > 0 ? c1.n : 1));
    f[3] = (float)log2((double)(c2.n > 0 ? c2.n : 1));
    f[4] = (float)log2(dm / P.minsep);
    f[5] = (float)fmin((s1 + s2) / dm, 8.);
    f[6] = p2 ? 1.f : 0.f;
    f[7] = (float)log2(1. + d / P.maxsep);
    f[8] = (float)log2(P.b > 1e-6 ? P.b : 1e-6);
    f[9] = (float)log2(P.binsize);
    f[10] = C == Flat ? 0.f : 1.f;
}

__global__ void k_jobs_logged(const Item* __restrict__ jobs, long long njobs,
                              ItemJ* __restrict__ out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < njobs) out[k] = ItemJ{jobs[k].i, jobs[k].j, (int)k, 0};
}

// Every stride-th job of a logged run: its features and items processed.
__global__ void k_log_sample(const float* __restrict__ feat, const unsigned int* __restrict__ vis,
                             long long njobs, long long stride, long long nout,
                             float* __restrict__ fout, float* __restrict__ vout) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nout) return;
    const long long src = k * stride;
    for (int f = 0; f < CM_NF; ++f) fout[k * CM_NF + f] = feat[src * CM_NF + f];
    vout[k] = (float)vis[src];
}

// Predicted work of each job, 2^y - 1 (the model predicts y = log2(1 + visits)).
__global__ void k_pred_cost(const float* __restrict__ y, long long n, double* __restrict__ c) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        const double v = exp2((double)fminf(fmaxf(y[k], 0.f), 60.f)) - 1.;
        c[k] = v;
    }
}

// Predicted work per tile of COST_TILE jobs (a fixed-order tree sum per tile).
constexpr int COST_TILE = 4096;
__global__ __launch_bounds__(256) void k_cost_tiles(const double* __restrict__ c, long long n,
                                                    double* __restrict__ tiles) {
    __shared__ double s_sum[256];
    const long long t0 = (long long)blockIdx.x * COST_TILE;
    double a = 0.;
    for (int k = threadIdx.x; k < COST_TILE; k += 256)
        if (t0 + k < n) a += c[t0 + k];
    s_sum[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) tiles[blockIdx.x] = s_sum[0];
}

// bounds[w] for w = 1..G-1: the first job of tile tile[w] at which the running sum of
// predicted work within the tile exceeds rem[w] (the part of w's target in that tile);
// bounds[0] = 0 and bounds[G] = n.  Targets increase with w, so the bounds do too.
__global__ void k_cost_bounds(const double* __restrict__ c, long long n, int G,
                              const long long* __restrict__ tile, const double* __restrict__ rem,
                              long long* __restrict__ bounds) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w > G) return;
    if (w == 0) { bounds[0] = 0; return; }
    if (w == G) { bounds[G] = n; return; }
    const long long t0 = tile[w] * COST_TILE;
    const long long t1 = t0 + COST_TILE < n ? t0 + COST_TILE : n;
    double a = 0.;
    long long k = t0;
    while (k < t1) {
        a += c[k];
        if (a > rem[w]) break;
        ++k;
    }
    bounds[w] = k;
}

// Longest-first order of the jobs by predicted y in steps of 1/4 (256 classes): a
// counting sort.  Key 255 - 4y, so the largest predictions come first.
__device__ __forceinline__ int lpt_key(float y) {
    const float q = fminf(fmaxf(y * 4.f, 0.f), 255.f);
    return 255 - (int)q;
}
// Per-workgroup class counts (LDS) added to the global counts.
__global__ __launch_bounds__(256) void k_lpt_count(const float* __restrict__ y, long long n,
                                                   unsigned long long* __restrict__ cnt) {
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        atomicAdd(&h[lpt_key(y[k])], 1u);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}
// Scatter: cursor[key] starts at the class's offset; each workgroup reserves room for
// its jobs of every class with one atomic per class, then places them (the order
// within a class does not matter).
__global__ __launch_bounds__(256) void k_lpt_scatter(const float* __restrict__ y,
                                                     const Item* __restrict__ jobs, long long n,
                                                     unsigned long long* __restrict__ cursor,
                                                     Item* __restrict__ out) {
    __shared__ unsigned int h[256];
    __shared__ unsigned long long base[256];
    const long long k0 = (long long)blockIdx.x * 4096;
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < 4096; t += 256)
        if (k0 + t < n) atomicAdd(&h[lpt_key(y[k0 + t])], 1u);
    __syncthreads();
    base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], (unsigned long long)h[threadIdx.x]) : 0;
    h[threadIdx.x] = 0;
    __syncthreads();
    for (int t = threadIdx.x; t < 4096; t += 256)
        if (k0 + t < n) {
            const int key = lpt_key(y[k0 + t]);
            const unsigned int r = atomicAdd(&h[key], 1u);
            out[base[key] + r] = jobs[k0 + t];
        }
}


// Clears n words (a kernel, not a memset, so that a captured graph replays it reliably).
__global__ void k_fill_u64(unsigned long long* __restrict__ p, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0ull;
}

// Fixed-order compensated reduction of the per-workgroup slots.
// One wavefront per output entry (nh sums, nb npairs, NCNT counters, in that order):
// lane l adds slots l, l+64, ... with compensation, then the 64 partial sums are
// combined in a fixed tree order, so the result does not depend on timing.
__global__ __launch_bounds__(64) void k_reduce_slots(const double* slot_hi, const double* slot_lo,
                                                     const unsigned long long* slot_np,
                                                     const unsigned long long* slot_cnt,
                                                     int nslots, int nh, int nb, double* out,
                                                     unsigned long long* out_np,
                                                     unsigned long long* out_cnt) {
    const int e = blockIdx.x, lane = threadIdx.x;
    if (e < nh) {
        double s = 0., c = 0.;
        auto add = [&](double v) {
            const double tt = s + v;
            c += (fabs(s) >= fabs(v)) ? (s - tt) + v : (v - tt) + s;
            s = tt;
        };
        for (int k = lane; k < nslots; k += 64) {
            add(slot_hi[(size_t)k * nh + e]);
            add(slot_lo[(size_t)k * nh + e]);
        }
        for (int d = 32; d >= 1; d >>= 1) {
            const double os = __shfl(s, lane + d, 64), oc = __shfl(c, lane + d, 64);
            if (lane < d) { add(os); c += oc; }
        }
        if (lane == 0) out[e] = s + c;
        return;
    }
    const bool isnp = e < nh + nb;
    const int k = isnp ? e - nh : e - nh - nb;
    const int stride = isnp ? nb : NCNT;
    const unsigned long long* src = isnp ? slot_np : slot_cnt;
    unsigned long long v = 0;
    for (int j = lane; j < nslots; j += 64) v += src[(size_t)j * stride + k];
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = __shfl(v, lane + d, 64);
        if (lane < d) v += o;
    }
    if (lane == 0) (isnp ? out_np : out_cnt)[k] = v;
}

// Exclusive scan of the per-workgroup child counts (one workgroup of 1024 threads,
// nblk <= TRAV_SLOTS); boff[nblk] = total.
__global__ __launch_bounds__(1024) void k_scan_counts(const unsigned int* bcount, int nblk,
                                                      unsigned long long* boff,
                                                      unsigned long long* total) {
    __shared__ unsigned long long wsum[16];
    __shared__ unsigned long long carry_s;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nblk; c0 += 1024) {
        const int i = c0 + t;
        const unsigned long long v = i < nblk ? bcount[i] : 0ull;
        unsigned long long x = v;
        for (int d = 1; d < 64; d <<= 1) {
            unsigned long long y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        unsigned long long wb = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            if (k < wave) wb += wsum[k];
            tot += wsum[k];
        }
        const unsigned long long carry = carry_s;
        if (i < nblk) boff[i] = carry + wb + x - v;
        __syncthreads();
        if (t == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (t == 0) { boff[nblk] = carry_s; *total = carry_s; }
}

// Copies the concatenated items [0, limit) of the regions onto the stack (the rest is
// read in place by the next launch): workgroup b copies its share to dst[boff[b] ...].
template <class IT>
__global__ __launch_bounds__(256) void k_gather_children(const IT* __restrict__ out,
                                                         long long per_block,
                                                         const unsigned int* __restrict__ bcount,
                                                         const unsigned long long* __restrict__ boff,
                                                         unsigned long long limit,
                                                         IT* __restrict__ dst) {
    const int b = blockIdx.x;
    if (boff[b] >= limit) return;
    const unsigned long long n = bcount[b] < limit - boff[b] ? bcount[b] : limit - boff[b];
    const IT* src = out + (size_t)b * 4 * per_block;
    IT* d = dst + boff[b];
    for (unsigned long long i = threadIdx.x; i < n; i += blockDim.x) d[i] = src[i];
}

__global__ void k_items_auto(const int* top, int ntop, Item* out) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long npair = (long long)ntop * (ntop - 1) / 2;
    if (t < ntop) {
        out[t] = Item{top[t], -1};
    } else if (t < ntop + npair) {
        // map t - ntop to (i, j), i < j, row-major over i
        long long q = t - ntop;
        long long i = 0;
        // row i holds ntop-1-i pairs; solve directly then adjust
        double nn = double(ntop);
        i = (long long)((2 * nn - 1 - sqrt((2 * nn - 1) * (2 * nn - 1) - 8.0 * double(q))) / 2);
        if (i < 0) i = 0;
        auto row_start = [&](long long r) { return r * (2 * (long long)ntop - r - 1) / 2; };
        while (i > 0 && row_start(i) > q) --i;
        while (row_start(i + 1) <= q) ++i;
        long long j = i + 1 + (q - row_start(i));
        out[t] = Item{top[i], top[j]};
    }
}

__global__ void k_items_cross(const int* top1, int ntop1, const int* top2, int ntop2, Item* out) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (long long)ntop1 * ntop2) {
        long long i = t / ntop2, j = t % ntop2;
        out[t] = Item{top1[i], top2[j]};
    }
}

// Initial items of a batched run, one workgroup per patch pair.  Pair p (pi[p],
// pj[p]) owns the items [ioff[p], ioff[p+1]): for an auto pair (pauto[p] != 0) the
// process2 items of the patch's top cells then its top-cell pairs a<b, otherwise all
// top-cell pairs of the two patches.  Top cells are global node ids, patch q's at
// tops[toff[q] ...].
__global__ __launch_bounds__(256) void k_items_pairs(
    const int* __restrict__ pi, const int* __restrict__ pj, const int* __restrict__ pauto,
    const long long* __restrict__ ioff, const int* __restrict__ tops1,
    const int* __restrict__ toff1, const int* __restrict__ tops2,
    const int* __restrict__ toff2, ItemP* __restrict__ out) {
    const int p = blockIdx.x;
    const long long o0 = ioff[p], cnt = ioff[p + 1] - o0;
    const int i = pi[p], j = pj[p];
    const int* ta = tops1 + toff1[i];
    const long long na = toff1[i + 1] - toff1[i];
    const bool autop = pauto[p] != 0;
    const int* tb = tops2 + toff2[j];
    const long long nbt = toff2[j + 1] - toff2[j];
    auto row_start = [&](long long r) { return r * (2 * na - r - 1) / 2; };
    for (long long q = threadIdx.x; q < cnt; q += blockDim.x) {
        if (!autop) {
            out[o0 + q] = ItemP{ta[q / nbt], tb[q % nbt], p, 0};
        } else if (q < na) {
            out[o0 + q] = ItemP{ta[q], -1, p, 0};
        } else {
            // row-major over a < b: row a holds na-1-a pairs
            const long long u = q - na;
            long long a = (long long)((2. * na - 1 -
                                       sqrt((2. * na - 1) * (2. * na - 1) - 8. * double(u))) / 2);
            if (a < 0) a = 0;
            while (a > 0 && row_start(a) > u) --a;
            while (row_start(a + 1) <= u) ++a;
            out[o0 + q] = ItemP{ta[a], ta[a + 1 + (u - row_start(a))], p, 0};
        }
    }
}

// Copies a patch's nodes into a concatenated array, moving child links by the
// patch's offset.
__global__ void k_concat_nodes(const DNode* __restrict__ src, int n, int offset,
                               DNode* __restrict__ dst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    DNode d = src[t];
    if (d.child >= 0) d.child += offset;
    dst[offset + t] = d;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// The arguments of a persistent launch, as the key of its captured graph.
struct GraphKey {
    const void* fn;
    const void* ptr[13];
    long long num[7];
    unsigned char params[sizeof(C2Params)];
    bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(GraphKey)) == 0; }
};
// A few captured launches, the least recently used replaced first.
struct GraphCache {
    static constexpr int N = 8;
    GraphKey key[N];
    hipGraphExec_t exec[N] = {};
    long long used[N] = {};
    long long clock = 0;
    hipGraphExec_t find(const GraphKey& k) {
        for (int i = 0; i < N; ++i)
            if (exec[i] && key[i] == k) {
                used[i] = ++clock;
                return exec[i];
            }
        return nullptr;
    }
    void add(const GraphKey& k, hipGraphExec_t e) {
        int v = 0;
        for (int i = 0; i < N; ++i)
            if (!exec[i] || used[i] < used[v]) { v = i; if (!exec[i]) break; }
        if (exec[v]) (void)hipGraphExecDestroy(exec[v]);
        std::memset(&key[v], 0, sizeof(GraphKey));
        key[v] = k;
        exec[v] = e;
        used[v] = ++clock;
    }
    void clear() {
        for (int i = 0; i < N; ++i) {
            if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
            exec[i] = nullptr;
        }
    }
    ~GraphCache() { clear(); }
};

// Traversal buffers, kept by the library for reuse (sized in bytes: a batched run
// stores 16-byte items, a single correlation 8-byte ones).
struct Corr2Scratch {
    void* stack = nullptr;
    size_t stack_bytes = 0;
    void* outq[2] = {nullptr, nullptr};     // output regions of consecutive launches
    size_t outq_bytes = 0;
    double *slot_hi = nullptr, *slot_lo = nullptr;
    unsigned long long *slot_np = nullptr, *slot_cnt = nullptr;
    double* red = nullptr;
    unsigned long long *red_np = nullptr, *red_cnt = nullptr;
    unsigned long long* outcount = nullptr;
    unsigned long long* h_outcount = nullptr;
    unsigned int* bcount = nullptr;     // two count arrays (ping-pong), TRAV_SLOTS each
    void* pstack = nullptr;             // per-workgroup stacks of the persistent traversal
    size_t pstack_bytes = 0;
    unsigned long long* pctl = nullptr; // persistent traversal: job counter, overflow flag
    // the learned schedule and the job logs: features, predictions, predicted work and
    // its prefix, share bounds, sort keys, reordered jobs, logged jobs and their visits
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
        int get(size_t b) {
            if (b <= bytes) return 0;
            dev_free(p, bytes);
            bytes = 0;
            p = dev_alloc(b ? b : 1);
            if (!p) return fail("out of device memory for the traversal schedule");
            bytes = b;
            return 0;
        }
        template <class T> T* as() const { return static_cast<T*>(p); }
        ~Buf() { dev_free(p, bytes); }
    };
    Buf feat, pred, cost, prefix, bounds, keys, jobs2, jobsJ, jvis, logX, logV;
    GraphCache graphs;      // persistent launches captured as hipGraphs (graph mode)
    unsigned long long* boff = nullptr; // two offset arrays, TRAV_SLOTS+1 each
    int nb = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // sizes of the per-bin arrays (they depend on nb) and of the fixed ones
    static size_t hist_bytes(int nb) { return sizeof(double) * (size_t)NQ2 * nb * TRAV_SLOTS; }
    static size_t np_bytes(int nb) { return sizeof(unsigned long long) * (size_t)nb * TRAV_SLOTS; }
    static constexpr size_t kCntBytes = sizeof(unsigned long long) * NCNT * TRAV_SLOTS;
    static constexpr size_t kBcountBytes = sizeof(unsigned int) * 2 * TRAV_SLOTS;
    static constexpr size_t kBoffBytes = sizeof(unsigned long long) * 2 * (TRAV_SLOTS + 1);
    void free_bins() {
        dev_free(slot_hi, hist_bytes(nb));
        dev_free(slot_lo, hist_bytes(nb));
        dev_free(slot_np, np_bytes(nb));
        dev_free(red, sizeof(double) * (size_t)NQ2 * nb);
        dev_free(red_np, sizeof(unsigned long long) * nb);
        slot_hi = slot_lo = red = nullptr;
        slot_np = red_np = nullptr;
    }
    ~Corr2Scratch() {
        // the caching allocator keeps these for the next correlation
        dev_free(stack, stack_bytes);
        for (void* q : outq) dev_free(q, outq_bytes);
        free_bins();
        dev_free(slot_cnt, kCntBytes);
        dev_free(red_cnt, sizeof(unsigned long long) * NCNT);
        dev_free(outcount, sizeof(unsigned long long));
        dev_free(bcount, kBcountBytes);
        dev_free(boff, kBoffBytes);
        dev_free(pstack, pstack_bytes);
        dev_free(pctl, PCTL_WORDS * sizeof(unsigned long long));
        if (h_outcount) (void)hipHostFree(h_outcount);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

static int ensure_scratch(Corr2Scratch& S, int nb, size_t item_bytes) {
    auto get = [](auto*& p, size_t bytes) {
        p = static_cast<std::remove_reference_t<decltype(p)>>(dev_alloc(bytes));
        return p != nullptr;
    };
    const char* oom = "out of device memory for the traversal";
    if (S.nb != nb || !S.slot_hi) {
        const size_t nh = (size_t)NQ2 * nb;
        S.free_bins();
        S.nb = nb;
        if (!get(S.slot_hi, Corr2Scratch::hist_bytes(nb)) ||
            !get(S.slot_lo, Corr2Scratch::hist_bytes(nb)) ||
            !get(S.slot_np, Corr2Scratch::np_bytes(nb)) || !get(S.red, sizeof(double) * nh) ||
            !get(S.red_np, sizeof(unsigned long long) * nb))
            return fail(oom);
    }
    if (!S.slot_cnt) {
        if (!get(S.slot_cnt, Corr2Scratch::kCntBytes) ||
            !get(S.red_cnt, sizeof(unsigned long long) * NCNT) ||
            !get(S.outcount, sizeof(unsigned long long)) ||
            !get(S.bcount, Corr2Scratch::kBcountBytes) || !get(S.boff, Corr2Scratch::kBoffBytes))
            return fail(oom);
        TC_HIP(hipHostMalloc((void**)&S.h_outcount, sizeof(unsigned long long)));
        TC_HIP(hipEventCreate(&S.ev0));
        TC_HIP(hipEventCreate(&S.ev1));
    }
    const size_t qb = item_bytes * 4 * ((size_t)g_chunk + (size_t)TRAV_THREADS * TRAV_SLOTS);
    if (S.outq_bytes < qb) {
        for (void*& q : S.outq) {
            dev_free(q, S.outq_bytes);
            q = nullptr;
        }
        S.outq_bytes = 0;
        for (void*& q : S.outq)
            if (!(q = dev_alloc(qb))) return fail(oom);
        S.outq_bytes = qb;
    }
    return 0;
}

static int ensure_stack(Corr2Scratch& S, size_t need, size_t keep) {
    if (need <= S.stack_bytes) return 0;
    size_t cap = std::max(need, std::max<size_t>(2 * S.stack_bytes, size_t(32) << 20));
    void* ns = dev_alloc(cap);
    if (!ns) return fail("out of device memory for the traversal stack");
    if (S.stack && keep > 0)
        TC_HIP(hipMemcpyAsync(ns, S.stack, keep, hipMemcpyDeviceToDevice, stream()));
    TC_HIP(hipStreamSynchronize(stream()));
    dev_free(S.stack, S.stack_bytes);
    S.stack = ns;
    S.stack_bytes = cap;
    return 0;
}

template <class IT>
using TravFn = void (*)(const DNode*, const DNode*, InView<IT>, long long, long long, IT*,
                        unsigned int*, C2Params, int, double*, double*, unsigned long long*,
                        unsigned long long*, double*, unsigned long long*);

template <class IT>
static TravFn<IT> pick_traverse(int D, int C, bool gm) {
#define TC_TF(DD, CC) return gm ? &k_traverse<DD, CC, IT, true> : &k_traverse<DD, CC, IT, false>
    if (D == NData) { if (C == Flat) TC_TF(NData, Flat); if (C == ThreeD) TC_TF(NData, ThreeD); TC_TF(NData, Sphere); }
    if (D == KData) { if (C == Flat) TC_TF(KData, Flat); if (C == ThreeD) TC_TF(KData, ThreeD); TC_TF(KData, Sphere); }
    if (D == XK_NK) { if (C == Flat) TC_TF(XK_NK, Flat); if (C == ThreeD) TC_TF(XK_NK, ThreeD); TC_TF(XK_NK, Sphere); }
    if (D == XK_NG) { if (C == Flat) TC_TF(XK_NG, Flat); if (C == ThreeD) TC_TF(XK_NG, ThreeD); TC_TF(XK_NG, Sphere); }
    if (D == XK_KG) { if (C == Flat) TC_TF(XK_KG, Flat); if (C == ThreeD) TC_TF(XK_KG, ThreeD); TC_TF(XK_KG, Sphere); }
    if (D == ZData) { if (C == Flat) TC_TF(ZData, Flat); if (C == ThreeD) TC_TF(ZData, ThreeD); TC_TF(ZData, Sphere); }
    if (D == XK_NZ) { if (C == Flat) TC_TF(XK_NZ, Flat); if (C == ThreeD) TC_TF(XK_NZ, ThreeD); TC_TF(XK_NZ, Sphere); }
    if (D == XK_KZ) { if (C == Flat) TC_TF(XK_KZ, Flat); if (C == ThreeD) TC_TF(XK_KZ, ThreeD); TC_TF(XK_KZ, Sphere); }
#define TC_TF3(KK) if (D == KK) { if (C == Flat) TC_TF(KK, Flat); if (C == ThreeD) TC_TF(KK, ThreeD); TC_TF(KK, Sphere); }
    TC_TF3(VData) TC_TF3(TData) TC_TF3(QData)
    TC_TF3(XK_NV) TC_TF3(XK_NT) TC_TF3(XK_NQ) TC_TF3(XK_KV) TC_TF3(XK_KT) TC_TF3(XK_KQ)
#undef TC_TF3
    if (C == Flat) TC_TF(GData, Flat); if (C == ThreeD) TC_TF(GData, ThreeD); TC_TF(GData, Sphere);
#undef TC_TF
}

// LDS histogram copies and bytes per workgroup for nslot pair histograms.
static int hist_copies(int nb, int nslot, size_t& shbytes) {
    const size_t nh = (size_t)NQ2 * nb;
    int R = 8;     // as many copies as fit in 48 KB, at most 8
    while (R > 1 && sizeof(double) * ((nh + nb) * (size_t)R * nslot + NCNT) > 48 * 1024) R /= 2;
    shbytes = sizeof(double) * ((nh + nb) * (size_t)R * nslot + NCNT);
    return R;
}

// The LIFO work-stack loop over launches; the stack holds the ninit initial items.
template <class IT>
// With until > 0 the loop stops as soon as at least `until` items are pending, with all
// of them on the stack; *left receives their number (the jobs of a persistent launch).
static int traverse_stack(Corr2Scratch& S, TravFn<IT> fn, const DNode* n1, const DNode* n2,
                          long long ninit, const C2Params& P, int R, size_t shbytes,
                          double* pair_hist, unsigned long long* pair_np,
                          long long until = 0, long long* left = nullptr) {
    hipStream_t st = stream();
    IT* stack = static_cast<IT*>(S.stack);
    IT* outq[2] = {static_cast<IT*>(S.outq[0]), static_cast<IT*>(S.outq[1])};
    long long top = ninit;
    KernelStat& ks = kernel_stat(KS_TRAVERSE2);
    const bool timing = timing_enabled();
    // children of the last launch, still in its output regions (newest work)
    int cur = 0;                 // buffer the last launch wrote
    long long reg_n = 0, reg_pb = 0;
    int reg_blocks = 0;
    for (;;) {
        if (until > 0 && top + reg_n >= until) {
            // enough pending work: everything onto the stack, for the persistent launch
            if (reg_n > 0) {
                if (ensure_stack(S, sizeof(IT) * (top + reg_n), sizeof(IT) * top)) return 2;
                stack = static_cast<IT*>(S.stack);
                hipLaunchKernelGGL(k_gather_children<IT>, dim3(reg_blocks), dim3(256), 0, st,
                                   outq[cur], reg_pb, S.bcount + cur * TRAV_SLOTS,
                                   S.boff + cur * (TRAV_SLOTS + 1), (unsigned long long)reg_n,
                                   stack + top);
                TC_HIP(hipGetLastError());
                top += reg_n;
            }
            *left = top;
            return 0;
        }
        InView<IT> V{nullptr, nullptr, 0, 0, 0, 0, nullptr};
        long long n = 0;
        if (reg_n > 0) {
            // LIFO: the newest children are processed next, in place; the oldest part
            // beyond one chunk is moved onto the stack first.
            n = std::min(reg_n, g_chunk);
            const long long excess = reg_n - n;
            if (excess > 0) {
                if (ensure_stack(S, sizeof(IT) * (top + excess + 4 * TRAV_CHUNK),
                                 sizeof(IT) * top)) return 2;
                stack = static_cast<IT*>(S.stack);
                hipLaunchKernelGGL(k_gather_children<IT>, dim3(reg_blocks), dim3(256), 0, st,
                                   outq[cur], reg_pb, S.bcount + cur * TRAV_SLOTS,
                                   S.boff + cur * (TRAV_SLOTS + 1),
                                   (unsigned long long)excess, stack + top);
                TC_HIP(hipGetLastError());
                top += excess;
            }
            // a short head is topped up with the newest stack items
            const long long extra = excess > 0 ? 0 : std::min(top, g_chunk - n);
            V = InView<IT>{outq[cur], S.boff + cur * (TRAV_SLOTS + 1), 4 * reg_pb, excess, n,
                           reg_blocks, stack + (top - extra)};
            n += extra;
            top -= extra;
            reg_n = 0;
        } else if (top > 0) {
            n = std::min(top, g_chunk);
            V = InView<IT>{nullptr, nullptr, 0, 0, 0, 0, stack + (top - n)};
            top -= n;
        } else {
            break;
        }
        const int nxt = 1 - cur;
        // contiguous ranges of whole 256-item rounds, one per workgroup
        const long long rounds = (n + TRAV_THREADS - 1) / TRAV_THREADS;
        const long long per_block =
            ((rounds + TRAV_SLOTS - 1) / TRAV_SLOTS) * (long long)TRAV_THREADS;
        const int blocks = (int)((n + per_block - 1) / per_block);
        if (timing) TC_HIP(hipEventRecord(S.ev0, st));
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(TRAV_THREADS), shbytes, st, n1, n2, V, n,
                           per_block, outq[nxt], S.bcount + nxt * TRAV_SLOTS, P, R, S.slot_hi,
                           S.slot_lo, S.slot_np, S.slot_cnt, pair_hist, pair_np);
        TC_HIP(hipGetLastError());
        if (timing) TC_HIP(hipEventRecord(S.ev1, st));
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, S.bcount + nxt * TRAV_SLOTS,
                           blocks, S.boff + nxt * (TRAV_SLOTS + 1), S.outcount);
        TC_HIP(hipGetLastError());
        TC_HIP(hipMemcpyAsync(S.h_outcount, S.outcount, sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (timing) {
            float ms = 0.f;
            TC_HIP(hipEventElapsedTime(&ms, S.ev0, S.ev1));
            ks.ms += ms;
            ks.launches += 1;
        }
        reg_n = (long long)*S.h_outcount;
        reg_pb = per_block;
        reg_blocks = blocks;
        cur = nxt;
    }
    if (left) *left = 0;
    return 0;
}

template <class IT>
using PersistFn = void (*)(const DNode*, const DNode*, const IT*, long long, unsigned long long*,
                           int, int, const long long*, IT*, long long, unsigned int*, C2Params, int,
                           double*, double*, unsigned long long*, unsigned long long*, double*,
                           unsigned long long*, unsigned int*, unsigned long long*);

template <class IT>
static PersistFn<IT> pick_persist(int D, int C, bool gm) {
#define TC_TF(DD, CC) return gm ? &k_traverse_persist<DD, CC, IT, true> : &k_traverse_persist<DD, CC, IT, false>
#define TC_TF3(KK) if (D == KK) { if (C == Flat) TC_TF(KK, Flat); if (C == ThreeD) TC_TF(KK, ThreeD); TC_TF(KK, Sphere); }
    TC_TF3(NData) TC_TF3(KData) TC_TF3(XK_NK) TC_TF3(XK_NG) TC_TF3(XK_KG) TC_TF3(ZData)
    TC_TF3(XK_NZ) TC_TF3(XK_KZ) TC_TF3(VData) TC_TF3(TData) TC_TF3(QData)
    TC_TF3(XK_NV) TC_TF3(XK_NT) TC_TF3(XK_NQ) TC_TF3(XK_KV) TC_TF3(XK_KT) TC_TF3(XK_KQ)
    if (C == Flat) TC_TF(GData, Flat); if (C == ThreeD) TC_TF(GData, ThreeD); TC_TF(GData, Sphere);
#undef TC_TF3
#undef TC_TF
}

// The logged form (items carry their job) for the correlation kinds the cost model is
// trained on: NN, KK and GG with the Euclidean metric.
static PersistFn<ItemJ> pick_persist_logged(int D, int C) {
#define TC_TF(DD, CC) return &k_traverse_persist<DD, CC, ItemJ, false>
#define TC_TF3(KK) if (D == KK) { if (C == Flat) TC_TF(KK, Flat); if (C == ThreeD) TC_TF(KK, ThreeD); TC_TF(KK, Sphere); }
    TC_TF3(NData) TC_TF3(KData) TC_TF3(GData)
#undef TC_TF3
#undef TC_TF
    return nullptr;
}

// How a correlation's traversal runs (tc_set_traversal_mode): the persistent kernel
// taking jobs by a shared counter (default), the persistent kernel with a static
// equal-count partition of the jobs, or the chunked launches driven from the host.
static int g_trav_mode = TRAV_PERSIST_DYNAMIC;
std::atomic<long long> g_persist_fallbacks{0};   // persistent runs redone chunked
static int g_persist_levels = 0;    // stack levels per workgroup (0: from the tree depths)
static int g_persist_grab_div = 16; // a shared-counter grab takes 1/div of a fair share
static bool g_graph_mode = false;   // persistent launches captured in hipGraphs and replayed
static long long g_graph_captures = 0, g_graph_replays = 0;
int set_persist_grab(int div) {
    if (div < 1) return fail("the grab divisor must be >= 1");
    g_persist_grab_div = div;
    return 0;
}
void set_graph_mode(bool on) { g_graph_mode = on; }
void graph_stats(long long* captures, long long* replays) {
    *captures = g_graph_captures;
    *replays = g_graph_replays;
}
static long long g_persist_min_jobs = PERSIST_MIN_JOBS;
int set_persist_params(int levels, long long min_jobs) {
    if (levels < 0 || min_jobs < 0) return fail("levels and min_jobs must be >= 0");
    g_persist_levels = levels;
    g_persist_min_jobs = min_jobs == 0 ? PERSIST_MIN_JOBS : min_jobs;
    return 0;
}
int set_traversal_mode(int mode) {
    if (mode < TRAV_CHUNKED || mode > TRAV_PERSIST_LEARNED_DYNAMIC)
        return fail("unknown traversal mode");
    g_trav_mode = mode;
    return 0;
}
int traversal_mode() { return g_trav_mode; }

// The cost model of the learned modes, and the job log: when g_log_max > 0 the next
// persistent traversal counts every job's items and keeps the features and counts of
// an evenly strided sample of at most g_log_max jobs.
static const CostModel* g_cost_model = nullptr;
void set_cost_model(const CostModel* m) { g_cost_model = m; }
static long long g_persist_cnt[NCNT] = {0, 0, 0, 0};   // the last persistent launch's counters
void persist_counters(long long* c) {
    for (int k = 0; k < NCNT; ++k) c[k] = g_persist_cnt[k];
}
static long long g_log_max = 0;
static std::vector<float> g_log_X, g_log_V;
void set_job_log(long long max_samples) {
    g_log_max = max_samples;
    g_log_X.clear();
    g_log_V.clear();
}
long long job_log_fetch(float* X, float* V, long long cap) {
    const long long n = std::min<long long>(cap, (long long)g_log_V.size());
    if (X) std::copy(g_log_X.begin(), g_log_X.begin() + n * CM_NF, X);
    if (V) std::copy(g_log_V.begin(), g_log_V.begin() + n, V);
    return (long long)g_log_V.size();
}

// Workgroups of a persistent launch: as many as the chip holds at once.
static int persist_grid(const void* fn, size_t shbytes, int& grid) {
    int dev = 0, ncu = 0, per_cu = 0;
    TC_HIP(hipGetDevice(&dev));
    TC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    TC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, TRAV_THREADS, shbytes));
    grid = std::max(1, std::min(TRAV_SLOTS, std::max(1, per_cu) * ncu));
    return 0;
}

// One launch of the persistent traversal over the njobs items at jobs.  depth_sum bounds
// the chain of splits below a job (the depths of the two trees); each workgroup's stack
// holds 4 * TRAV_THREADS items per level of it.  Sets overflow when a workgroup's stack
// would not have held its children (the sums are then incomplete).
template <class IT>
static int traverse_persist(Corr2Scratch& S, PersistFn<IT> fn, const DNode* n1, const DNode* n2,
                            const IT* jobs, long long njobs, int depth_sum, bool static_part,
                            const long long* bounds, const C2Params& P, int R, size_t shbytes,
                            double* pair_hist, unsigned long long* pair_np,
                            unsigned int* job_visits, bool& overflow) {
    hipStream_t st = stream();
    overflow = false;
    if (njobs <= 0) return 0;
    int grid = 0;
    if (int rc = persist_grid(reinterpret_cast<const void*>(fn), shbytes, grid)) return rc;
    const long long cap = 4LL * TRAV_THREADS * (g_persist_levels > 0 ? g_persist_levels : depth_sum + 4);
    const size_t bytes = sizeof(IT) * (size_t)cap * grid;
    if (S.pstack_bytes < bytes) {
        dev_free(S.pstack, S.pstack_bytes);
        S.pstack_bytes = 0;
        S.pstack = dev_alloc(bytes);
        if (!S.pstack) return fail("out of device memory for the traversal stacks");
        S.pstack_bytes = bytes;
    }
    if (!S.pctl) {
        S.pctl = static_cast<unsigned long long*>(dev_alloc(PCTL_WORDS * sizeof(unsigned long long)));
        if (!S.pctl) return fail("out of device memory for the traversal");
    }
    KernelStat& ks = kernel_stat(KS_TRAVERSE2);
    KernelStat& kp = kernel_stat(KS_PERSIST);
    const bool timing = timing_enabled();
    // the launch: the control words cleared, then the persistent kernel
    auto enqueue = [&]() -> int {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, S.pctl, (long long)PCTL_WORDS);
        hipLaunchKernelGGL(fn, dim3(grid), dim3(TRAV_THREADS), shbytes, st, n1, n2, jobs, njobs,
                           S.pctl, static_part ? 1 : 0, g_persist_grab_div, bounds,
                           static_cast<IT*>(S.pstack), cap,
                           reinterpret_cast<unsigned int*>(S.pctl + 1), P, R, S.slot_hi,
                           S.slot_lo, S.slot_np, S.slot_cnt, pair_hist, pair_np, job_visits,
                           S.pctl + 2);
        TC_HIP(hipGetLastError());
        return 0;
    };
    if (timing) TC_HIP(hipEventRecord(S.ev0, st));
    if (!g_graph_mode) {
        if (int rc = enqueue()) return rc;
    } else {
        // Captured once per launch signature (every argument the kernels take) and
        // replayed: the arguments of a graph's kernel nodes are fixed at capture.
        GraphKey key{};
        key.fn = reinterpret_cast<const void*>(fn);
        key.ptr[0] = n1; key.ptr[1] = n2; key.ptr[2] = jobs; key.ptr[3] = bounds;
        key.ptr[4] = S.pstack; key.ptr[5] = S.pctl; key.ptr[6] = S.slot_hi;
        key.ptr[7] = S.slot_lo; key.ptr[8] = S.slot_np; key.ptr[9] = S.slot_cnt;
        key.ptr[10] = pair_hist; key.ptr[11] = pair_np; key.ptr[12] = job_visits;
        key.num[0] = njobs; key.num[1] = cap; key.num[2] = grid; key.num[3] = (long long)shbytes;
        key.num[4] = R; key.num[5] = static_part; key.num[6] = g_persist_grab_div;
        std::memcpy(key.params, &P, sizeof(C2Params));
        hipGraphExec_t exec = S.graphs.find(key);
        if (!exec) {
            hipGraph_t graph = nullptr;
            TC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue();
            const hipError_t e = hipStreamEndCapture(st, &graph);
            if (rc) return rc;
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) return hip_fail(ei, "hipGraphInstantiate");
            S.graphs.add(key, exec);
            ++g_graph_captures;
        } else {
            ++g_graph_replays;
        }
        TC_HIP(hipGraphLaunch(exec, st));
    }
    if (timing) TC_HIP(hipEventRecord(S.ev1, st));
    unsigned long long ctl[PCTL_WORDS];
    TC_HIP(hipMemcpyAsync(ctl, S.pctl, sizeof ctl, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    if (timing) {
        float ms = 0.f;
        TC_HIP(hipEventElapsedTime(&ms, S.ev0, S.ev1));
        ks.ms += ms;
        ks.launches += 1;
        kp.ms += ms;
        kp.launches += 1;
    }
    overflow = (ctl[1] & 0xffffffffull) != 0;
    for (int k = 0; k < NCNT; ++k) g_persist_cnt[k] = (long long)ctl[2 + k];
    return 0;
}

// The learned schedules and the job log start from the jobs' features (S.feat).  The
// learned static mode gives workgroup w the jobs [bounds[w], bounds[w+1]) of equal
// predicted work; the learned dynamic mode hands out the jobs in order of decreasing
// predicted work (jobs then points at the reordered list).
static int schedule_jobs(Corr2Scratch& S, int D, int C, bool gm, const DNode* n1,
                         const DNode* n2, const Item*& jobs, long long njobs, const C2Params& P,
                         size_t shbytes, const long long*& bounds) {
    hipStream_t st = stream();
    const bool learned = g_trav_mode == TRAV_PERSIST_LEARNED_STATIC ||
                         g_trav_mode == TRAV_PERSIST_LEARNED_DYNAMIC;
    if (learned && !g_cost_model)
        return fail("a learned traversal mode needs a cost model (tc_set_cost_model)");
    if ((!learned && g_log_max <= 0) || njobs <= 0) return 0;
    if (S.feat.get(sizeof(float) * CM_NF * njobs)) return 2;
    const unsigned g = (unsigned)((njobs + 255) / 256);
    if (C == Flat) hipLaunchKernelGGL(k_job_features<Flat>, dim3(g), dim3(256), 0, st, n1, n2, jobs, njobs, P, S.feat.as<float>());
    else if (C == ThreeD) hipLaunchKernelGGL(k_job_features<ThreeD>, dim3(g), dim3(256), 0, st, n1, n2, jobs, njobs, P, S.feat.as<float>());
    else hipLaunchKernelGGL(k_job_features<Sphere>, dim3(g), dim3(256), 0, st, n1, n2, jobs, njobs, P, S.feat.as<float>());
    TC_HIP(hipGetLastError());
    if (!learned) return 0;
    if (S.pred.get(sizeof(float) * njobs)) return 2;
    if (int rc = cost_model_predict_device(g_cost_model, S.feat.as<float>(), njobs,
                                           S.pred.as<float>(), st))
        return rc;
    if (g_trav_mode == TRAV_PERSIST_LEARNED_STATIC) {
        int grid = 0;
        if (int rc = persist_grid(reinterpret_cast<const void*>(pick_persist<Item>(D, C, gm)),
                                  shbytes, grid))
            return rc;
        const long long ntile = (njobs + COST_TILE - 1) / COST_TILE;
        if (S.cost.get(sizeof(double) * njobs) || S.prefix.get(sizeof(double) * ntile) ||
            S.bounds.get(sizeof(long long) * (grid + 1)) ||
            S.keys.get((sizeof(long long) + sizeof(double)) * (grid + 1)))
            return 2;
        hipLaunchKernelGGL(k_pred_cost, dim3(g), dim3(256), 0, st, S.pred.as<float>(), njobs,
                           S.cost.as<double>());
        hipLaunchKernelGGL(k_cost_tiles, dim3((unsigned)ntile), dim3(256), 0, st,
                           S.cost.as<double>(), njobs, S.prefix.as<double>());
        TC_HIP(hipGetLastError());
        std::vector<double> tiles(ntile);
        TC_HIP(hipMemcpyAsync(tiles.data(), S.prefix.p, sizeof(double) * ntile,
                              hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        // the tile of each share's end and the rest of its target inside that tile
        std::vector<double> pre(ntile + 1, 0.);
        for (long long t = 0; t < ntile; ++t) pre[t + 1] = pre[t] + tiles[t];
        std::vector<long long> tile(grid + 1, 0);
        std::vector<double> rem(grid + 1, 0.);
        for (int w = 1; w < grid; ++w) {
            const double target = pre[ntile] * (double)w / (double)grid;
            long long t = (long long)(std::upper_bound(pre.begin() + 1, pre.end(), target) -
                                      (pre.begin() + 1));
            if (t >= ntile) t = ntile - 1;
            tile[w] = t;
            rem[w] = target - pre[t];
        }
        long long* dtile = S.keys.as<long long>();
        double* drem = reinterpret_cast<double*>(dtile + (grid + 1));
        TC_HIP(hipMemcpyAsync(dtile, tile.data(), sizeof(long long) * (grid + 1),
                              hipMemcpyHostToDevice, st));
        TC_HIP(hipMemcpyAsync(drem, rem.data(), sizeof(double) * (grid + 1),
                              hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_cost_bounds, dim3((grid + 256) / 256), dim3(256), 0, st,
                           S.cost.as<double>(), njobs, grid, dtile, drem,
                           S.bounds.as<long long>());
        TC_HIP(hipGetLastError());
        TC_HIP(hipStreamSynchronize(st));     // tile and rem are host arrays the copies read
        bounds = S.bounds.as<long long>();
        return 0;
    }
    if (S.keys.get(sizeof(unsigned long long) * 256) || S.jobs2.get(sizeof(Item) * njobs))
        return 2;
    unsigned long long* cnt = S.keys.as<unsigned long long>();
    TC_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * 256, st));
    hipLaunchKernelGGL(k_lpt_count, dim3((unsigned)std::min<long long>(g, 4096)), dim3(256), 0, st,
                       S.pred.as<float>(), njobs, cnt);
    TC_HIP(hipGetLastError());
    unsigned long long h[256];
    TC_HIP(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    unsigned long long off = 0;
    for (int k = 0; k < 256; ++k) {
        const unsigned long long c = h[k];
        h[k] = off;
        off += c;
    }
    TC_HIP(hipMemcpyAsync(cnt, h, sizeof h, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_lpt_scatter, dim3((unsigned)((njobs + 4095) / 4096)), dim3(256), 0, st,
                       S.pred.as<float>(), jobs, njobs, cnt, S.jobs2.as<Item>());
    TC_HIP(hipGetLastError());
    TC_HIP(hipStreamSynchronize(st));     // h is a host array the copy above reads
    jobs = S.jobs2.as<Item>();
    return 0;
}

// An evenly strided sample of at most g_log_max jobs of a logged run, to the host.
static int keep_job_log(Corr2Scratch& S, long long njobs) {
    hipStream_t st = stream();
    if (njobs <= 0) return 0;
    const long long stride = std::max<long long>(1, (njobs + g_log_max - 1) / g_log_max);
    const long long nout = (njobs + stride - 1) / stride;
    if (S.logX.get(sizeof(float) * CM_NF * nout) || S.logV.get(sizeof(float) * nout)) return 2;
    hipLaunchKernelGGL(k_log_sample, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st,
                       S.feat.as<float>(), S.jvis.as<unsigned int>(), njobs, stride, nout,
                       S.logX.as<float>(), S.logV.as<float>());
    TC_HIP(hipGetLastError());
    const size_t o = g_log_V.size();
    g_log_X.resize((o + nout) * CM_NF);
    g_log_V.resize(o + nout);
    TC_HIP(hipMemcpyAsync(g_log_X.data() + o * CM_NF, S.logX.p, sizeof(float) * CM_NF * nout,
                          hipMemcpyDeviceToHost, st));
    TC_HIP(hipMemcpyAsync(g_log_V.data() + o, S.logV.p, sizeof(float) * nout,
                          hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    return 0;
}

static int zero_slots(Corr2Scratch& S, int nb) {
    hipStream_t st = stream();
    const size_t nh = (size_t)NQ2 * nb;
    TC_HIP(hipMemsetAsync(S.slot_hi, 0, sizeof(double) * nh * TRAV_SLOTS, st));
    TC_HIP(hipMemsetAsync(S.slot_lo, 0, sizeof(double) * nh * TRAV_SLOTS, st));
    TC_HIP(hipMemsetAsync(S.slot_np, 0, sizeof(unsigned long long) * nb * TRAV_SLOTS, st));
    TC_HIP(hipMemsetAsync(S.slot_cnt, 0, sizeof(unsigned long long) * NCNT * TRAV_SLOTS, st));
    return 0;
}

// Runs the whole traversal for one auto or cross correlation and writes the
// reduced sums: out[NQ2*nb] doubles, np[nb], cnt[NCNT] (host arrays).
int corr2_run(Corr2Scratch& S, int D, int C, const DNode* nodes1, const int* dtop1, int ntop1,
              int depth1, const DNode* nodes2, const int* dtop2, int ntop2, int depth2,
              const C2Params& P, double* out, unsigned long long* np, unsigned long long* cnt) {
    hipStream_t st = stream();
    const int nb = P.nbins;
    const size_t nh = (size_t)NQ2 * nb;
    if (sizeof(double) * (nh + nb + NCNT) > 64 * 1024)
        return fail("too many bins for the LDS histogram (nbins <= 1000)");
    size_t shbytes = 0;
    const int R = hist_copies(nb, 1, shbytes);
    if (ensure_scratch(S, nb, sizeof(Item))) return 2;
    if (zero_slots(S, nb)) return 2;
    const bool autoc = (nodes2 == nullptr);
    long long ninit = autoc ? (long long)ntop1 + (long long)ntop1 * (ntop1 - 1) / 2
                            : (long long)ntop1 * ntop2;
    if (ensure_stack(S, sizeof(Item) * (ninit + 4 * TRAV_CHUNK), 0)) return 2;
    if (ninit > 0) {
        dim3 g((unsigned)((ninit + 255) / 256));
        Item* stack = static_cast<Item*>(S.stack);
        if (autoc) hipLaunchKernelGGL(k_items_auto, g, dim3(256), 0, st, dtop1, ntop1, stack);
        else hipLaunchKernelGGL(k_items_cross, g, dim3(256), 0, st, dtop1, ntop1, dtop2, ntop2, stack);
        TC_HIP(hipGetLastError());
    }
    const bool gm = P.metric != MetEuclidean || P.rpar_on;
    bool chunked = g_trav_mode == TRAV_CHUNKED;
    if (!chunked) {
        // The initial node pairs are first split by chunked launches (breadth first,
        // while fewer than a chunk are pending) until there are g_persist_min_jobs jobs;
        // then one persistent launch takes them all.
        long long njobs = ninit;
        if (ninit < g_persist_min_jobs)
            if (int rc = traverse_stack<Item>(S, pick_traverse<Item>(D, C, gm), nodes1,
                                              autoc ? nodes1 : nodes2, ninit, P, R, shbytes,
                                              nullptr, nullptr, g_persist_min_jobs, &njobs))
                return rc;
        bool overflow = false;
        const DNode* n2 = autoc ? nodes1 : nodes2;
        const int dsum = depth1 + (autoc ? depth1 : depth2);
        const bool stat = g_trav_mode == TRAV_PERSIST_STATIC ||
                          g_trav_mode == TRAV_PERSIST_LEARNED_STATIC;
        const Item* jobs = static_cast<const Item*>(S.stack);
        const long long* bounds = nullptr;
        if (int rc = schedule_jobs(S, D, C, gm, nodes1, n2, jobs, njobs, P, shbytes, bounds))
            return rc;
        if (g_log_max > 0) {
            PersistFn<ItemJ> fl = gm ? nullptr : pick_persist_logged(D, C);
            if (!fl) return fail("job logs cover NN, KK and GG with the Euclidean metric");
            if (S.jobsJ.get(sizeof(ItemJ) * std::max<long long>(njobs, 1)) ||
                S.jvis.get(sizeof(unsigned int) * std::max<long long>(njobs, 1)))
                return 2;
            const unsigned g = (unsigned)((njobs + 255) / 256);
            if (njobs > 0)
                hipLaunchKernelGGL(k_jobs_logged, dim3(g), dim3(256), 0, st, jobs, njobs,
                                   S.jobsJ.as<ItemJ>());
            TC_HIP(hipMemsetAsync(S.jvis.p, 0, sizeof(unsigned int) * njobs, st));
            if (int rc = traverse_persist<ItemJ>(S, fl, nodes1, n2, S.jobsJ.as<ItemJ>(), njobs,
                                                 dsum, stat, bounds, P, R, shbytes, nullptr,
                                                 nullptr, S.jvis.as<unsigned int>(), overflow))
                return rc;
            if (!overflow)
                if (int rc = keep_job_log(S, njobs)) return rc;
        } else {
            if (int rc = traverse_persist<Item>(S, pick_persist<Item>(D, C, gm), nodes1, n2, jobs,
                                                njobs, dsum, stat, bounds, P, R, shbytes, nullptr,
                                                nullptr, nullptr, overflow))
                return rc;
        }
        if (overflow) {      // a stack would have overflowed: start again, chunked
            g_persist_fallbacks.fetch_add(1);
            if (zero_slots(S, nb)) return 2;
            if (ninit > 0) {
                dim3 g((unsigned)((ninit + 255) / 256));
                Item* stack = static_cast<Item*>(S.stack);
                if (autoc) hipLaunchKernelGGL(k_items_auto, g, dim3(256), 0, st, dtop1, ntop1, stack);
                else hipLaunchKernelGGL(k_items_cross, g, dim3(256), 0, st, dtop1, ntop1, dtop2, ntop2, stack);
                TC_HIP(hipGetLastError());
            }
            chunked = true;
        }
    }
    if (chunked)
        if (int rc = traverse_stack<Item>(S, pick_traverse<Item>(D, C, gm), nodes1,
                                          autoc ? nodes1 : nodes2, ninit, P, R, shbytes, nullptr,
                                          nullptr))
            return rc;
    const int tot = (int)(nh + nb + NCNT);
    hipLaunchKernelGGL(k_reduce_slots, dim3(tot), dim3(64), 0, st, S.slot_hi,
                       S.slot_lo, S.slot_np, S.slot_cnt, TRAV_SLOTS, (int)nh, nb, S.red, S.red_np,
                       S.red_cnt);
    TC_HIP(hipGetLastError());
    TC_HIP(hipMemcpyAsync(out, S.red, sizeof(double) * nh, hipMemcpyDeviceToHost, st));
    TC_HIP(hipMemcpyAsync(np, S.red_np, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost, st));
    TC_HIP(hipMemcpyAsync(cnt, S.red_cnt, sizeof(unsigned long long) * NCNT, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    return 0;
}

// All patch pairs of a patch-based correlation in one traversal.  Side 1 has
// npatch1 trees (side 2 npatch2, or side 1 again when trees2 is empty); pair p
// correlates patch pi[p] of side 1 with patch pj[p] of side 2 (an auto pair when
// pauto[p]).  Writes NQ2*nb sums and nb npairs per pair (host arrays, overwritten).
int corr2_run_pairs(Corr2Scratch& S, int D, int C, const std::vector<const DeviceTree*>& trees1,
                    const std::vector<const DeviceTree*>& trees2, const std::vector<int>& pi,
                    const std::vector<int>& pj, const std::vector<int>& pauto, const C2Params& P,
                    double* pair_out, unsigned long long* pair_npairs, unsigned long long* cnt) {
    hipStream_t st = stream();
    const int nb = P.nbins;
    const size_t nh = (size_t)NQ2 * nb;
    const int npair = (int)pi.size();
    if (npair == 0) { for (int k = 0; k < NCNT; ++k) cnt[k] = 0; return 0; }
    size_t shbytes = 0;
    const int R = hist_copies(nb, PB_SLOTS, shbytes);
    if (sizeof(double) * ((nh + nb) * PB_SLOTS + NCNT) > 64 * 1024)
        return fail("too many bins for batched patch pairs (nbins <= 200)");
    const bool same = trees2.empty();
    const auto& t2 = same ? trees1 : trees2;
    // concatenated nodes and top cells of each side
    auto concat = [&](const std::vector<const DeviceTree*>& ts, DNode*& nodes, int*& tops,
                      std::vector<int>& toff, size_t& nbytes, size_t& tbytes) -> int {
        long long nn = 0, nt = 0;
        for (auto* t : ts) { nn += t->nnodes; nt += (long long)t->top.size(); }
        // Trees of one multi-root build share their node array: use it as it is.
        bool shared = true;
        for (auto* t : ts) shared = shared && t->nodes == ts[0]->nodes;
        if (shared) {
            nbytes = 0;
            nodes = ts[0]->nodes;
            tbytes = sizeof(int) * std::max<long long>(nt, 1);
            tops = static_cast<int*>(dev_alloc(tbytes));
            if (!tops) return fail("out of device memory for the top cells");
            std::vector<int> htops;
            htops.reserve(nt);
            toff.assign(1, 0);
            for (auto* t : ts) {
                htops.insert(htops.end(), t->top.begin(), t->top.end());
                toff.push_back((int)htops.size());
            }
            if (!htops.empty())
                TC_HIP(hipMemcpyAsync(tops, htops.data(), sizeof(int) * htops.size(),
                                      hipMemcpyHostToDevice, st));
            TC_HIP(hipStreamSynchronize(st));
            return 0;
        }
        if (nn >= (1LL << 31) - 1) return fail("patches hold too many tree nodes together");
        nbytes = sizeof(DNode) * std::max<long long>(nn, 1);
        tbytes = sizeof(int) * std::max<long long>(nt, 1);
        nodes = static_cast<DNode*>(dev_alloc(nbytes));
        tops = static_cast<int*>(dev_alloc(tbytes));
        if (!nodes || !tops) return fail("out of device memory concatenating patch trees");
        std::vector<int> htops;
        htops.reserve(nt);
        toff.assign(1, 0);
        int off = 0;
        for (auto* t : ts) {
            if (t->nnodes > 0)
                hipLaunchKernelGGL(k_concat_nodes, dim3((t->nnodes + 255) / 256), dim3(256), 0,
                                   st, t->nodes, t->nnodes, off, nodes);
            for (int v : t->top) htops.push_back(v + off);
            toff.push_back((int)htops.size());
            off += t->nnodes;
        }
        TC_HIP(hipGetLastError());
        if (!htops.empty())
            TC_HIP(hipMemcpyAsync(tops, htops.data(), sizeof(int) * htops.size(),
                                  hipMemcpyHostToDevice, st));
        TC_HIP(hipStreamSynchronize(st));
        return 0;
    };
    DNode *nodes1 = nullptr, *nodes2 = nullptr;
    int *tops1 = nullptr, *tops2 = nullptr;
    std::vector<int> toff1, toff2;
    size_t nb1 = 0, tb1 = 0, nb2 = 0, tb2 = 0;
    struct Freer {
        std::vector<std::pair<void*, size_t>> v;
        ~Freer() { for (auto& b : v) dev_free(b.first, b.second); }
    } fr;
    if (int rc = concat(trees1, nodes1, tops1, toff1, nb1, tb1)) return rc;
    if (nb1) fr.v.push_back({nodes1, nb1});     // nb == 0: a shared array, not ours
    fr.v.push_back({tops1, tb1});
    if (same) { nodes2 = nodes1; tops2 = tops1; toff2 = toff1; }
    else {
        if (int rc = concat(t2, nodes2, tops2, toff2, nb2, tb2)) return rc;
        if (nb2) fr.v.push_back({nodes2, nb2});
        fr.v.push_back({tops2, tb2});
    }
    // item offsets of the pairs
    std::vector<long long> ioff(npair + 1, 0);
    for (int p = 0; p < npair; ++p) {
        const long long na = toff1[pi[p] + 1] - toff1[pi[p]];
        const long long nbt = toff2[pj[p] + 1] - toff2[pj[p]];
        ioff[p + 1] = ioff[p] + (pauto[p] ? na + na * (na - 1) / 2 : na * nbt);
    }
    const long long ninit = ioff[npair];
    // small device tables
    const size_t pb = sizeof(int) * npair, ob = sizeof(long long) * (npair + 1);
    const size_t f1b = sizeof(int) * toff1.size(), f2b = sizeof(int) * toff2.size();
    int* dpi = static_cast<int*>(dev_alloc(pb));
    int* dpj = static_cast<int*>(dev_alloc(pb));
    int* dpa = static_cast<int*>(dev_alloc(pb));
    long long* dio = static_cast<long long*>(dev_alloc(ob));
    int* dt1 = static_cast<int*>(dev_alloc(f1b));
    int* dt2 = static_cast<int*>(dev_alloc(f2b));
    fr.v.push_back({dpi, pb}); fr.v.push_back({dpj, pb}); fr.v.push_back({dpa, pb});
    fr.v.push_back({dio, ob}); fr.v.push_back({dt1, f1b}); fr.v.push_back({dt2, f2b});
    if (!dpi || !dpj || !dpa || !dio || !dt1 || !dt2) return fail("out of device memory");
    TC_HIP(hipMemcpyAsync(dpi, pi.data(), pb, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(dpj, pj.data(), pb, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(dpa, pauto.data(), pb, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(dio, ioff.data(), ob, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(dt1, toff1.data(), f1b, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(dt2, toff2.data(), f2b, hipMemcpyHostToDevice, st));
    // per-pair accumulators
    const size_t hb = sizeof(double) * 2 * nh * npair, npb = sizeof(unsigned long long) * nb * npair;
    double* ph = static_cast<double*>(dev_alloc(hb));
    unsigned long long* pn = static_cast<unsigned long long*>(dev_alloc(npb));
    fr.v.push_back({ph, hb}); fr.v.push_back({pn, npb});
    if (!ph || !pn) return fail("out of device memory for the patch-pair sums");
    TC_HIP(hipMemsetAsync(ph, 0, hb, st));
    TC_HIP(hipMemsetAsync(pn, 0, npb, st));
    if (ensure_scratch(S, nb, sizeof(ItemP))) return 2;
    if (zero_slots(S, nb)) return 2;
    if (ensure_stack(S, sizeof(ItemP) * (ninit + 4 * TRAV_CHUNK), 0)) return 2;
    if (ninit > 0) {
        hipLaunchKernelGGL(k_items_pairs, dim3((unsigned)npair), dim3(256), 0, st,
                           dpi, dpj, dpa, dio, tops1, dt1, tops2, dt2,
                           static_cast<ItemP*>(S.stack));
        TC_HIP(hipGetLastError());
    }
    const bool gm = P.metric != MetEuclidean || P.rpar_on;
    bool chunked = g_trav_mode == TRAV_CHUNKED;
    if (!chunked) {
        // as in corr2_run: a breadth-first chunked split into enough jobs (half as many:
        // the items are twice as large), then every patch pair's jobs in one persistent
        // launch (the learned modes are for single correlations; a batched run takes its
        // jobs by the shared counter in them)
        long long njobs = ninit;
        if (ninit < g_persist_min_jobs / 2)
            if (int rc = traverse_stack<ItemP>(S, pick_traverse<ItemP>(D, C, gm), nodes1, nodes2,
                                               ninit, P, R, shbytes, ph, pn,
                                               g_persist_min_jobs / 2, &njobs))
                return rc;
        int depth = 0;
        for (auto* t : trees1) depth = std::max(depth, t->depth);
        int depth2 = 0;
        for (auto* t : t2) depth2 = std::max(depth2, t->depth);
        bool overflow = false;
        if (int rc = traverse_persist<ItemP>(S, pick_persist<ItemP>(D, C, gm), nodes1, nodes2,
                                             static_cast<const ItemP*>(S.stack), njobs,
                                             depth + depth2, g_trav_mode == TRAV_PERSIST_STATIC,
                                             nullptr, P, R, shbytes, ph, pn, nullptr, overflow))
            return rc;
        if (overflow) {      // a stack would have overflowed: start again, chunked
            g_persist_fallbacks.fetch_add(1);
            if (zero_slots(S, nb)) return 2;
            TC_HIP(hipMemsetAsync(ph, 0, hb, st));
            TC_HIP(hipMemsetAsync(pn, 0, npb, st));
            if (ninit > 0) {
                hipLaunchKernelGGL(k_items_pairs, dim3((unsigned)npair), dim3(256), 0, st,
                                   dpi, dpj, dpa, dio, tops1, dt1, tops2, dt2,
                                   static_cast<ItemP*>(S.stack));
                TC_HIP(hipGetLastError());
            }
            chunked = true;
        }
    }
    if (chunked)
        if (int rc = traverse_stack<ItemP>(S, pick_traverse<ItemP>(D, C, gm), nodes1, nodes2,
                                           ninit, P, R, shbytes, ph, pn))
            return rc;
    // counters from the workgroup slots
    const int tot = (int)(nh + nb + NCNT);
    hipLaunchKernelGGL(k_reduce_slots, dim3(tot), dim3(64), 0, st, S.slot_hi,
                       S.slot_lo, S.slot_np, S.slot_cnt, TRAV_SLOTS, (int)nh, nb, S.red, S.red_np,
                       S.red_cnt);
    TC_HIP(hipGetLastError());
    std::vector<double> hl(2 * nh * (size_t)npair);
    TC_HIP(hipMemcpyAsync(hl.data(), ph, hb, hipMemcpyDeviceToHost, st));
    TC_HIP(hipMemcpyAsync(pair_npairs, pn, npb, hipMemcpyDeviceToHost, st));
    TC_HIP(hipMemcpyAsync(cnt, S.red_cnt, sizeof(unsigned long long) * NCNT, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    for (int p = 0; p < npair; ++p)
        for (size_t t = 0; t < nh; ++t)
            pair_out[p * nh + t] = hl[2 * nh * p + t] + hl[2 * nh * p + nh + t];
    return 0;
}

long long persist_fallbacks() { return g_persist_fallbacks.load(); }

int set_traversal_chunk(long long items) {
    if (items < 0) return fail("chunk must be >= 0");
    g_chunk = items == 0 ? TRAV_CHUNK : std::min<long long>(std::max<long long>(items, 256), TRAV_CHUNK);
    return 0;
}

Corr2Scratch* new_corr2_scratch() { return new Corr2Scratch(); }
void delete_corr2_scratch(Corr2Scratch* s) { delete s; }

}  // namespace tc
