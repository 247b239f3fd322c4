// TEST INFRASTRUCTURE ONLY -- the CPU oracle for the treecorr_amd device engine.
//
// This header is shared by the oracle translation units (tc_oracle.cpp for the
// ball tree and the 2-point traversal, tc_oracle3.cpp for the 3-point multipole
// algorithm).  It restates, in plain C++17 without -ffast-math, the arithmetic of
// the TreeCorr 5.1.3 C++ engine that the device kernels must reproduce.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load the
// library built from these sources; the product never does.
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

namespace tco {

enum Coord { Flat = 1, ThreeD = 2, Sphere = 3 };
enum DType { NData = 0, KData = 1, ZData = 2, VData = 3, GData = 4, TData = 5, QData = 6 };
// spin of a complex field type (include/ProjectHelper.h:63-72)
inline int spin_of(int d) { return d == VData ? 1 : d == GData ? 2 : d == TData ? 3 : d == QData ? 4 : 0; }
inline bool is_complex(int d) { return d == ZData || d == VData || d == GData || d == TData || d == QData; }
// The mixed 2-point correlations (cross only), as codes beside the DType values:
// NK (src/Corr2.cpp DirectHelper<1>), NG, KG, NZ, KZ (DirectHelper<2>).
enum Corr2Kind { XK_NK = 8, XK_NG = 9, XK_KG = 10, XK_NZ = 11, XK_KZ = 12, XK_NV = 13,
                 XK_NT = 14, XK_NQ = 15, XK_KV = 16, XK_KT = 17, XK_KQ = 18 };
// the second field's type of a mixed kind
inline int mixed_d2(int kind) {
    switch (kind) {
        case XK_NK: return KData;
        case XK_NG: case XK_KG: return GData;
        case XK_NZ: case XK_KZ: return ZData;
        case XK_NV: case XK_KV: return VData;
        case XK_NT: case XK_KT: return TData;
        case XK_NQ: case XK_KQ: return QData;
        default: return -1;
    }
}
inline bool mixed_k1(int kind) { return kind == XK_KG || kind == XK_KZ || kind == XK_KV || kind == XK_KT || kind == XK_KQ; }
enum SplitMethod { Middle = 0, Median = 1, Mean = 2, Random = 3 };

struct Pt { double x = 0, y = 0, z = 0; };

// Position<C>::normSq, include/Position.h:66-70 (Flat), :171-175 (ThreeD)
inline double normsq(int C, const Pt& p) {
    return C == Flat ? p.x * p.x + p.y * p.y : p.x * p.x + p.y * p.y + p.z * p.z;
}
inline double coord_get(int C, const Pt& p, int split) {
    return split == 2 ? p.z : split == 1 ? p.y : p.x;
}
// Position<Sphere>::normalize, include/Position.h:251-256.  The reference binary
// (built with -O3 -ffast-math) divides by the norm at the centroid site
// (BuildCellData, src/Cell.cpp:190-193) but multiplies by the reciprocal of the
// norm where a Position<Sphere> is constructed from three coordinates or from a
// Position<ThreeD> (leaf objects, src/Field.cpp:127-130; ProjectHelper<ThreeD>).
inline void sphere_normalize_recip(Pt& p) {
    double n = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (n == 0.) { p.x = 1; p.y = 0; p.z = 0; return; }
    double r = 1. / n;
    p.x *= r; p.y *= r; p.z *= r;
}
inline void sphere_normalize(Pt& p) {
    double n = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    if (n == 0.) { p.x = 1; p.y = 0; p.z = 0; }
    else { p.x /= n; p.y /= n; p.z /= n; }
}

// One catalog object after the leaf CellData constructors (include/Cell.h:79-81,
// 131-133, 167-169): float w, float w*k, complex<float> w*g.
struct Item {
    Pt p;
    float w = 0, wk = 0, gr = 0, gi = 0;
    double wpos = 0;
    long idx = 0;
};

// One tree node.  left == -1 marks a leaf (single object or list-leaf).
struct Node {
    Pt p;
    float size = 0, w = 0, wk = 0, gr = 0, gi = 0;
    long n = 0;
    int left = -1, right = -1;
};

struct Tree {
    int D = 0, C = 0;
    std::vector<Node> nodes;   // pre-order per top cell, top cells in order
    std::vector<int> top;      // indices of top-level cells
    Pt center; double sizesq = 0;  // Field::_center/_sizesq, src/Field.cpp:199-203
    long nobj = 0;
};

Tree* build_tree(int D, int C, const double* x, const double* y, const double* z,
                 const double* d1, const double* d2, const double* w, const double* wpos,
                 long n, double minsize, double maxsize, int sm, bool brute,
                 int mintop, int maxtop);

// Neumaier-compensated accumulator used by the "exact" accumulation mode.
struct KSum {
    double s = 0, c = 0;
    void add(double v) {
        double t = s + v;
        if (std::fabs(s) >= std::fabs(v)) c += (s - t) + v; else c += (v - t) + s;
        s = t;
    }
    double get() const { return s + c; }
};

// ProjectHelper<Sphere>::calculate_direction, include/ProjectHelper.h:272-345
inline std::complex<double> sphere_direction(const Pt& p1, const Pt& p2) {
    double z1 = p1.z, z2 = p2.z;
    double dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
    double dsq = dx * dx + dy * dy + dz * dz;
    double cosA = (z1 - z2) + 0.5 * z2 * dsq;
    double sinA = p1.y * p2.x - p1.x * p2.y;
    return std::complex<double>(sinA, -cosA);
}
inline double safe_norm(const std::complex<double>& z) {  // ProjectHelper.h:25-32
    double n = z.real() * z.real() + z.imag() * z.imag();
    return n > 0. ? n : 1.;
}
// calculate_expmsialpha<2>, ProjectHelper.h:45-47: conj(r*r)/|r|^2
inline std::complex<double> expm2ialpha(const std::complex<double>& r) {
    double a = r.real(), b = r.imag();
    double rr = a * a - b * b, ri = a * b + b * a;
    double n = safe_norm(r);
    return std::complex<double>(rr / n, -ri / n);
}
inline std::complex<double> cmul(const std::complex<double>& u, const std::complex<double>& v) {
    return std::complex<double>(u.real() * v.real() - u.imag() * v.imag(),
                                u.real() * v.imag() + u.imag() * v.real());
}
// calculate_expmsialpha<s>, ProjectHelper.h:37-61, for s = 0..4
inline std::complex<double> expmsialpha(int s, const std::complex<double>& r) {
    if (s == 0) return std::complex<double>(1., 0.);
    if (s == 2) return expm2ialpha(r);
    if (s == 1) {                       // conj(r) / sqrt(|r|^2)
        const double q = std::sqrt(safe_norm(r));
        return std::complex<double>(r.real() / q, -r.imag() / q);
    }
    const std::complex<double> r2 = cmul(r, r);
    if (s == 3) {                       // conj(r^3) / sqrt(|r^3|^2)
        const std::complex<double> r3 = cmul(r2, r);
        const double q = std::sqrt(safe_norm(r3));
        return std::complex<double>(r3.real() / q, -r3.imag() / q);
    }
    const std::complex<double> r4 = cmul(r2, r2);   // s = 4: conj(r^4) / |r^2|^2
    const double q = safe_norm(r2);
    return std::complex<double>(r4.real() / q, -r4.imag() / q);
}

}  // namespace tco
