// treecorr_amd: shared definitions for the device ball-tree engine (gfx950).
//
// Device layout in HBM
//   DNode  48 B per tree node, children of node i at child(i), child(i)+1.
//   Elements of a catalog live as structure-of-arrays (x, y, z, wpos: double;
//   w, a, b: float payloads) and are physically permuted level by level while
//   the tree is built.
//
// All floating-point code here is compiled with -ffp-contract=off: the engine
// reproduces the arithmetic of TreeCorr's C++ (which runs without FMA) bit for
// bit where results depend on it (tree payloads, bin decisions).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace tc {

enum Coord { Flat = 1, ThreeD = 2, Sphere = 3 };       // include/Position.h:25
enum DType { NData = 0, KData = 1, ZData = 2, VData = 3, GData = 4, TData = 5, QData = 6 };
                                                            // include/Cell.h:34
// the complex field types and their spins (include/ProjectHelper.h:63-72)
__host__ __device__ constexpr bool is_complex(int d) {
    return d == ZData || d == VData || d == GData || d == TData || d == QData;
}
__host__ __device__ constexpr int spin_of(int d) {
    return d == VData ? 1 : d == GData ? 2 : d == TData ? 3 : d == QData ? 4 : 0;
}
// The mixed 2-point correlations (cross only) as codes beside the DType values:
// NK (src/Corr2.cpp DirectHelper<1>), N or K with a complex field (DirectHelper<2>).
enum Corr2Kind { XK_NK = 8, XK_NG = 9, XK_KG = 10, XK_NZ = 11, XK_KZ = 12, XK_NV = 13,
                 XK_NT = 14, XK_NQ = 15, XK_KV = 16, XK_KT = 17, XK_KQ = 18 };
// the second field's type of a mixed kind with a complex second field (-1 otherwise)
__host__ __device__ constexpr int mixed_d2(int kind) {
    return kind == XK_NG || kind == XK_KG ? GData : kind == XK_NZ || kind == XK_KZ ? ZData
         : kind == XK_NV || kind == XK_KV ? VData : kind == XK_NT || kind == XK_KT ? TData
         : kind == XK_NQ || kind == XK_KQ ? QData : -1;
}
// whether the first field of a mixed kind is K (else N)
__host__ __device__ constexpr bool mixed_k1(int kind) {
    return kind == XK_KG || kind == XK_KZ || kind == XK_KV || kind == XK_KT || kind == XK_KQ;
}
__host__ __device__ inline int corr2_kind(int d1, int d2) {
    if (d1 == d2) return (d1 == NData || d1 == KData || is_complex(d1)) ? d1 : -1;
    if (d1 == NData && d2 == KData) return XK_NK;
    if (d1 == NData && d2 == GData) return XK_NG;
    if (d1 == KData && d2 == GData) return XK_KG;
    if (d1 == NData && d2 == ZData) return XK_NZ;
    if (d1 == KData && d2 == ZData) return XK_KZ;
    if (d1 == NData && d2 == VData) return XK_NV;
    if (d1 == NData && d2 == TData) return XK_NT;
    if (d1 == NData && d2 == QData) return XK_NQ;
    if (d1 == KData && d2 == VData) return XK_KV;
    if (d1 == KData && d2 == TData) return XK_KT;
    if (d1 == KData && d2 == QData) return XK_KQ;
    return -1;
}
// xi arrays of a 2-point correlation: NN none; KK, NK xi; NG, KG, NZ, KZ xi and xi_im;
// GG, ZZ xip, xip_im, xim, xim_im
__host__ __device__ inline int corr2_nxi(int kind) {
    return kind == NData ? 0 : is_complex(kind) ? 4 : mixed_d2(kind) >= 0 ? 2 : 1;
}
enum SplitMethod { Middle = 0, Median = 1, Mean = 2, Random = 3 };  // include/Cell.h:26

// One tree node.  Mirrors BaseCell/CellData (include/Cell.h:79-353): float size,
// float w, float payload (wk, or the real/imag parts of wz), integer count.
struct alignas(16) DNode {
    double x, y, z;
    float size, w;
    float a, b;     // KData: a = w*k sum; GData: (a,b) = w*g sum (parallel transported)
    int child;      // first child index (second child is child+1); -1 for a leaf
    int n;          // number of catalog objects below this node
};
static_assert(sizeof(DNode) == 48, "DNode must stay 48 bytes");

// Traversal work item: a node pair (i, j) of the two trees, or j < 0 for the
// self-pairs of node i (BaseCorr2::process2).
struct alignas(8) Item { int i, j; };

// Record kept per device tree.
struct FieldInfo {
    double cx, cy, cz;   // Field::_center
    double sizesq;       // Field::_sizesq
};

void set_error(const std::string& msg);
int fail(const std::string& msg);     // sets the error, returns TC_ERR_INVALID
int hip_fail(hipError_t e, const char* what);  // sets the error, returns TC_ERR_HIP

hipStream_t stream();   // the library's stream (default priority)
hipStream_t stream2();  // second and third streams for work that overlaps the first
hipStream_t stream3();  // (default priority)

// Device memory for temporaries, kept for reuse across calls (a caching allocator:
// blocks are returned to a pool instead of hipFree).  The pool holds at most a fixed
// number of bytes; beyond that, freed blocks are released.
void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);

// Timing of the dominant kernels, recorded with HIP events on stream() when enabled.
struct KernelStat { double ms = 0; long long launches = 0; };
bool timing_enabled();
KernelStat& kernel_stat(int which);
enum { KS_TRAVERSE2 = 0, KS_TREE = 1, KS_MULTIPOLE = 2, KS_ZETA = 3, KS_PERSIST = 4, KS_COUNT = 5 };

}  // namespace tc

#define TC_HIP(call)                                                     \
    do {                                                                 \
        hipError_t _e = (call);                                          \
        if (_e != hipSuccess) return ::tc::hip_fail(_e, #call);          \
    } while (0)

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
namespace tc {

__device__ __forceinline__ double bcast_lane(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// ProjectHelper.h:25-32, :45-47: conj(r*r)/|r|^2 with |r|^2 -> 1 when zero.
__device__ __forceinline__ void expm2ialpha(double re, double im, double& er, double& ei) {
    double n = re * re + im * im;
    n = n > 0. ? n : 1.;
    double rr = re * re - im * im;
    double ri = re * im + im * re;
    er = rr / n;
    ei = -ri / n;
}

// Position::wrap, include/Position.h:106-112, :201-209: a separation brought into
// [-p/2, p/2] by whole periods.  Inside |d| <= 1.5 p (every pair of a catalog that lies
// in its box) this is the reference's subtraction loop, bit for bit.  A separation
// further out is first reduced exactly to |d| < p by fmod, and both loops are bounded,
// so a period below ulp(d) or positions far outside the box cannot make a kernel spin.
__host__ __device__ inline double wrap_period(double d, double p) {
    if (fabs(d) > 1.5 * p) d = fmod(d, p);
    for (int i = 0; i < 4 && d > p / 2.; ++i) d -= p;
    for (int i = 0; i < 4 && d < -p / 2.; ++i) d += p;
    return d;
}

// ---------------------------------------------------------------------------
// Bin-edge guard: the correctly rounded log for decisions close to an edge
// ---------------------------------------------------------------------------
// A pair whose bin coordinate kk = (log r - log minsep) / binsize lies within a few ulp
// of an integer, or whose singleBin refined test (include/BinType.h:128-147) lies within
// a few ulp of its threshold, lands differently for two logs that differ in the last
// bit.  The reference takes std::log from the host libm, which rounds correctly for all
// but about 1.5e-4 of its arguments; the device logs (the fdlibm-method tc_log, the
// device library's log) are within one ulp.  Such pairs therefore take their log again
// from log_cr: the approximation y0 corrected by one Newton step on exp(y) = x,
// log x = y0 + (x - e^y0) / e^y0 + O(2^-104 y0^2), with e^y0 in double-double
// arithmetic (reduction by ln 2 held in three parts, then a 24-term Taylor series on
// |r| <= ln2/2), so the sum carries about 100 bits before its final rounding.
struct DD { double hi, lo; };
__device__ __forceinline__ DD dd_fast(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ DD dd_two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD dd_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = dd_two_sum(a.hi, b.hi);
    const DD t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast(s.hi, s.lo);
}
__device__ __forceinline__ DD dd_mul(DD a, DD b) {
    DD p = dd_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return dd_fast(p.hi, p.lo);
}
__device__ __forceinline__ DD dd_div(DD a, double b) {
    const double q = a.hi / b;
    const DD p = dd_prod(q, b);
    return dd_fast(q, ((a.hi - p.hi) - p.lo + a.lo) / b);
}
// e^y in double-double for |y| < 700
__device__ inline DD exp_dd(double y) {
    const double n = rint(y * 0x1.71547652b82fep+0);
    const DD a = dd_prod(n, 0x1.62e42fefa39efp-1), b = dd_prod(n, 0x1.abc9e3b39803fp-56);
    DD r = dd_add(DD{y, 0.}, DD{-a.hi, -a.lo});
    r = dd_add(r, DD{-b.hi, -b.lo});
    r = dd_add(r, DD{-n * 0x1.7b57a079a1934p-111, 0.});
    DD s{1., 0.}, t{1., 0.};
    for (int i = 1; i <= 24; ++i) {
        t = dd_div(dd_mul(t, r), double(i));
        s = dd_add(s, t);
    }
    const int e = int(n);
    return {ldexp(s.hi, e), ldexp(s.lo, e)};
}
// log x rounded correctly (to within 2^-50 ulp of a rounding midpoint), from y0, any
// approximation of log x within a few ulp
__device__ inline double log_cr(double x, double y0) {
    const DD e = exp_dd(y0);
    const double d = x - e.hi;                    // exact: x and e.hi are within 2x
    return y0 + (d - e.lo) / e.hi;
}
// The change of kk that a two-ulp change of logr can make, with the rounding of the
// subtraction and of the division (or of the product with a rounded 1/binsize) on top,
// an absolute 4e-16 of logr for 0.5 log(rsq) in place of log(sqrt(rsq)) (the rounding of
// the sqrt moves its log by at most 2^-53), and a factor of two to spare.
__device__ __forceinline__ double kk_slack(double logr, double logminsep, double inv_binsize,
                                           double kk) {
    return (2e-15 * (fabs(logr) + fabs(logminsep)) + 4e-16) * inv_binsize +
           2e-15 * (fabs(kk) + 1.);
}
__device__ __forceinline__ bool near_int(double kk, double slack) {
    return fabs(kk - rint(kk)) <= slack;
}
// directProcess11's bin index int((log x - logminsep) / binsize) from lx, a device log of
// x, with the guard; logr receives the log that decided it
__device__ __forceinline__ int guarded_bin(double x, double lx, double logminsep,
                                           double binsize, double& logr) {
    double kk = (lx - logminsep) / binsize;
    if (near_int(kk, kk_slack(lx, logminsep, 1. / binsize, kk))) {
        lx = log_cr(x, lx);
        kk = (lx - logminsep) / binsize;
    }
    logr = lx;
    return int(kk);
}
// singleBin's refined test at bin coordinate kk (include/BinType.h:128-147), the
// reference's expressions unchanged: true when the pair fits one bin.  near is set when
// a comparison is within what a change of kk by dkk could turn over.
__device__ __forceinline__ bool refined_fits(double kk, double s1ps2sq, double rsq,
                                             double binsize, double b, double dkk,
                                             bool& near) {
    const int ik = int(kk);
    const double frackk = kk - ik;
    const double omf = 1. - frackk;
    const double f = (omf < frackk) ? omf : frackk;
    const double dt = dkk * binsize;
    const double t1 = f * binsize + b;
    const double m1 = (t1 * t1) * rsq;
    near = near || fabs(s1ps2sq - m1) <= (2. * fabs(t1) + dt) * dt * rsq + 1e-15 * m1;
    if (s1ps2sq > m1) return false;
    const double t2 = frackk * binsize + b - s1ps2sq / rsq;
    const double m2 = (t2 * t2) * rsq;
    near = near || fabs(s1ps2sq - m2) <= (2. * fabs(t2) + dt) * dt * rsq + 1e-15 * m2;
    return !(s1ps2sq > m2);
}

// The metrics of a 2-point run, values of the reference's enum (include/Metric.h:29).
enum Metric { MetEuclidean = 1, MetRperp = 2, MetRlens = 3, MetArc = 4, MetOldRperp = 5,
              MetPeriodic = 6 };
// the metrics measured across the line of sight (3-d only)
__host__ __device__ inline bool los_metric(int m) {
    return m == MetRperp || m == MetRlens || m == MetOldRperp;
}

// MetricHelper<M>::DistSq (include/Metric.h) for the metrics above:
//   Euclidean  the squared norm of the separation (:143-147, :197-202);
//   Arc        on the sphere (2 asin(L/2))^2, L the chord (:666-680); in 3-d
//              atan2(|a x b|, a.b)^2, the sizes divided by |a|, |b| to put them on the
//              unit sphere (:701-721);
//   Periodic   each separation wrapped by its period, then the squared norm (:831-838).
__host__ __device__ __forceinline__ double metric_dist_sq(int metric, int coords, double ax,
                                                          double ay, double az, double bx,
                                                          double by, double bz, double xp,
                                                          double yp, double zp, double& s1,
                                                          double& s2) {
    if (metric == MetRperp) {
        // MetricHelper<Rperp>::DistSq, include/Metric.h:416-434: |a x b|^2 / |L|^2 with L
        // the midpoint; each size scaled by the farther object's distance over |L|
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double lx = (ax + bx) * 0.5, ly = (ay + by) * 0.5, lz = (az + bz) * 0.5;
        const double nl = lx * lx + ly * ly + lz * lz;
        const double rperpsq = nl > 0. ? (cx * cx + cy * cy + cz * cz) / nl
                                       : 4. * (ax * ax + ay * ay + az * az);
        const double r1sq = ax * ax + ay * ay + az * az, r2sq = bx * bx + by * by + bz * bz;
        if (r2sq > nl && s1 != 0) s1 *= sqrt(r2sq / nl);
        if (r1sq > nl && s2 != 0) s2 *= sqrt(r1sq / nl);
        return rperpsq;
    }
    if (metric == MetOldRperp) {
        // MetricHelper<OldRperp>::DistSq, include/Metric.h:277-293: |a-b|^2 - (r1-r2)^2,
        // the nearer object's size grown by (r2^2-r1^2)/(4 r1^2)
        const double dx = ax - bx, dy = ay - by, dz = az - bz;
        const double d3sq = dx * dx + dy * dy + dz * dz;
        const double r1sq = ax * ax + ay * ay + az * az, r2sq = bx * bx + by * by + bz * bz;
        const double rparsq = (r1sq - r2sq) * (r1sq - r2sq) / (r1sq + r2sq + 2. * sqrt(r1sq * r2sq));
        if (r1sq < r2sq) {
            if (s1 != 0. && s1 < 1.7976931348623157e308) s1 *= (1. + 0.25 * (r2sq - r1sq) / r1sq);
        } else {
            if (s2 != 0. && s2 < 1.7976931348623157e308) s2 *= (1. + 0.25 * (r1sq - r2sq) / r2sq);
        }
        return fabs(d3sq - rparsq);
    }
    if (metric == MetRlens) {
        // MetricHelper<Rlens>::DistSq, include/Metric.h:582-589: the separation at the
        // lens (the first object's) distance, the second size scaled to it
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double r2sq = bx * bx + by * by + bz * bz;
        const double rsq = (cx * cx + cy * cy + cz * cz) / r2sq;
        s2 *= sqrt((ax * ax + ay * ay + az * az) / r2sq);
        return rsq;
    }
    if (metric == MetArc) {
        if (coords == Sphere) {
            const double dx = ax - bx, dy = ay - by, dz = az - bz;
            const double th = 2. * asin(sqrt(dx * dx + dy * dy + dz * dz) / 2.);
            return th * th;
        }
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double th = atan2(sqrt(cx * cx + cy * cy + cz * cz), ax * bx + ay * by + az * bz);
        s1 /= sqrt(ax * ax + ay * ay + az * az);
        s2 /= sqrt(bx * bx + by * by + bz * bz);
        return th * th;
    }
    double dx = ax - bx, dy = ay - by, dz = az - bz;
    if (metric == MetPeriodic) {
        dx = wrap_period(dx, xp); dy = wrap_period(dy, yp);
        if (coords != Flat) dz = wrap_period(dz, zp);
    }
    return coords == Flat ? dx * dx + dy * dy : dx * dx + dy * dy + dz * dz;
}

// BaseCorr2::process11's pruning, src/Corr2.cpp:326-346, for the line-of-sight metrics
// and the r_par range (3-d): true when the pair is dropped, because r_par is outside
// [minrpar, maxrpar] by more than s1+s2 (ParHelper<1>: r along the midpoint's direction;
// ParHelper<2> for OldRperp: r2 - r1, include/Metric.h:37-106), or because the bin
// type's too-small/too-large test (bt_small, bt_large) holds and the metric's own
// (Rperp :458-476, OldRperp :300-318; the others always agree) confirms it.  rin: r_par
// lies inside the range by s1+s2, which singleBin requires.
__host__ __device__ inline bool metric_prune(int metric, int rpar_on, double minrpar,
                                             double maxrpar, double ax, double ay, double az,
                                             double bx, double by, double bz, double rsq,
                                             double s1ps2, bool bt_small, bool bt_large,
                                             double minsep, double minsepsq, double maxsep,
                                             double maxsepsq, bool& rin) {
    rin = true;
    double rpar = 0.;
    auto calc_rpar = [&]() {
        if (metric == MetOldRperp)
            return sqrt(bx * bx + by * by + bz * bz) - sqrt(ax * ax + ay * ay + az * az);
        const double rx = bx - ax, ry = by - ay, rz = bz - az;
        const double lx = (ax + bx) * 0.5, ly = (ay + by) * 0.5, lz = (az + bz) * 0.5;
        return (rx * lx + ry * ly + rz * lz) / sqrt(lx * lx + ly * ly + lz * lz);
    };
    if (rpar_on) {
        rpar = calc_rpar();
        if (rpar + s1ps2 < minrpar || rpar - s1ps2 > maxrpar) return true;
    }
    bool small = bt_small, large = bt_large;
    if (metric == MetRperp && (small || large)) {
        const double lx = (ax + bx) * 0.5, ly = (ay + by) * 0.5, lz = (az + bz) * 0.5;
        const double nl = lx * lx + ly * ly + lz * lz;
        if (small && !(rsq < nl)) {
            if (s1ps2 * s1ps2 > 4. * nl) small = false;
            else {
                const double t = minsep * (1 - s1ps2 / (2. * sqrt(nl))) - s1ps2;
                small = rsq < t * t;
            }
        }
        if (large && !(rsq < nl)) {
            const double t = maxsep * (1 + s1ps2 / (2. * sqrt(nl))) + s1ps2;
            large = rsq > t * t;
        }
    } else if (metric == MetOldRperp && (small || large)) {
        if (rpar == 0.) rpar = calc_rpar();
        const double d3 = sqrt(rpar * rpar + rsq);
        small = small && rsq + 2. * (d3 + fabs(rpar)) * s1ps2 < minsepsq;
        large = large && rsq - 2. * (d3 + fabs(rpar)) * s1ps2 > maxsepsq;
    }
    if (small || large) return true;
    if (rpar_on) rin = !(rpar - s1ps2 < minrpar) && !(rpar + s1ps2 > maxrpar);
    return false;
}

// calculate_expmsialpha<S>, include/ProjectHelper.h:37-61: exp(-i S alpha) of the
// direction r for spins 0..4 (S = 2 above).
template <int S>
__device__ __forceinline__ void expmsialpha(double re, double im, double& er, double& ei) {
    if (S == 0) { er = 1.; ei = 0.; return; }
    if (S == 2) { expm2ialpha(re, im, er, ei); return; }
    if (S == 1) {            // conj(r) / sqrt(|r|^2)
        double n = re * re + im * im;
        n = n > 0. ? n : 1.;
        const double q = sqrt(n);
        er = re / q; ei = -im / q;
        return;
    }
    const double r2r = re * re - im * im, r2i = re * im + im * re;
    if (S == 3) {            // conj(r^3) / sqrt(|r^3|^2)
        const double r3r = r2r * re - r2i * im, r3i = r2r * im + r2i * re;
        double n = r3r * r3r + r3i * r3i;
        n = n > 0. ? n : 1.;
        const double q = sqrt(n);
        er = r3r / q; ei = -r3i / q;
        return;
    }
    // S == 4: conj(r^4) / |r^2|^2
    const double r4r = r2r * r2r - r2i * r2i, r4i = r2r * r2i + r2i * r2r;
    double n = r2r * r2r + r2i * r2i;
    n = n > 0. ? n : 1.;
    er = r4r / n; ei = -r4i / n;
}

// Position<Sphere>::normalize as compiled where a Position<Sphere> is built
// from coordinates (reciprocal multiply) -- see the oracle's notes.
__device__ __forceinline__ void normalize_recip(double& x, double& y, double& z) {
    double n = sqrt(x * x + y * y + z * z);
    if (n == 0.) { x = 1.; y = 0.; z = 0.; return; }
    double r = 1. / n;
    x *= r; y *= r; z *= r;
}
__device__ __forceinline__ void normalize_div(double& x, double& y, double& z) {
    double n = sqrt(x * x + y * y + z * z);
    if (n == 0.) { x = 1.; y = 0.; z = 0.; return; }
    x /= n; y /= n; z /= n;
}

}  // namespace tc
