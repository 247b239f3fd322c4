// TEST INFRASTRUCTURE ONLY -- the CPU oracle for the treecorr_amd device engine.
//
// A plain C++17 + OpenMP restatement of the TreeCorr 5.1.3 ball-tree engine for the
// 2-point path: the Field/Cell mean-split tree build and the process2/process11
// dual-tree recursion with Log binning and the Euclidean metric for NN, KK and GG
// (Flat, ThreeD, Sphere).  Each function names the reference lines it follows.
// It is used as the parity checker by tests/ and as the timed CPU baseline in
// bench.py; it is never linked into or called by the product library.
//
// Accumulation modes: ref (plain double, the reference's semantics) and exact
// (Neumaier-compensated sums, the precision oracle).
#include "tc_oracle.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace tco {

// ---------------------------------------------------------------------------
// Tree build
// ---------------------------------------------------------------------------

struct Builder {
    int D, C, sm;
    bool brute;
    double minsizesq;
    std::vector<Item>& items;
    std::vector<int>& perm;   // the reference's vdata order, as indices into items
    Builder(int D_, int C_, int sm_, bool brute_, double minsizesq_,
            std::vector<Item>& it, std::vector<int>& pm)
        : D(D_), C(C_), sm(sm_), brute(brute_), minsizesq(minsizesq_), items(it), perm(pm) {}

    const Item& at(size_t i) const { return items[perm[i]]; }

    // BuildCellData, src/Cell.cpp:170-200: sequential weighted centroid + float(sum w).
    void cell_data(size_t start, size_t end, Node& nd) const {
        const Item& f = at(start);
        double wp = f.wpos;
        Pt pos{f.p.x * wp, f.p.y * wp, C == Flat ? 0. : f.p.z * wp};
        double ww = f.w;
        double sumwp = wp;
        for (size_t i = start + 1; i != end; ++i) {
            const Item& d = at(i);
            wp = d.wpos;
            pos.x += d.p.x * wp;
            pos.y += d.p.y * wp;
            if (C != Flat) pos.z += d.p.z * wp;
            sumwp += wp;
            ww += d.w;
        }
        nd.w = float(ww);
        if (sumwp != 0.) {
            pos.x /= sumwp; pos.y /= sumwp;
            if (C != Flat) pos.z /= sumwp;
            if (C == Sphere) sphere_normalize(pos);
        } else {
            pos = f.p;
        }
        nd.p = pos;
        nd.n = long(end - start);
    }

    // finishAverages: KData src/Cell.cpp:208-219; Flat SimpleSum :221-250;
    // ThreeD/Sphere ParallelTransportSum :276-354.
    void finish(size_t start, size_t end, Node& nd) const {
        if (D == KData) {
            double s = 0.;
            for (size_t i = start; i < end; ++i) s += double(at(i).wk);
            nd.wk = float(s);
        } else if (is_complex(D)) {
            // spin 0 (Z) is never rotated: ParallelTransportSum multiplies by
            // calculate_expmsialpha<0> = 1 (include/ProjectHelper.h:37-39)
            if (C == Flat || D == ZData) {
                double sr = 0., si = 0.;
                for (size_t i = start; i < end; ++i) { sr += double(at(i).gr); si += double(at(i).gi); }
                nd.gr = float(sr); nd.gi = float(si);
            } else {
                // The compiled reference evaluates this loop as below: for Sphere the
                // Position<Sphere> copies are not renormalised and sin(beta) is
                // factored as (cosA+cosB)*sinA; for ThreeD both positions are
                // normalised by reciprocal multiplication, dsq sums dy,dz first and
                // sin(beta) = ((z1-z2) + (0.5*z2*dsq + cosB))*sinA.
                const Pt& center = nd.p;
                Pt cen = center;
                if (C == ThreeD) sphere_normalize_recip(cen);
                double sr = 0., si = 0.;
                for (size_t i = start; i < end; ++i) {
                    const Item& d = at(i);
                    Pt pi = d.p;
                    if (C == ThreeD) sphere_normalize_recip(pi);
                    double z1 = center.z, z2 = pi.z;
                    double dx = cen.x - pi.x, dy = cen.y - pi.y, dz = cen.z - pi.z;
                    double dsq = C == Sphere ? (dx * dx + dy * dy) + dz * dz
                                             : (dy * dy + dz * dz) + dx * dx;
                    if (dsq < 1.e-12) {
                        sr += double(d.gr); si += double(d.gi);
                    } else {
                        double cosA = (z1 - z2) + 0.5 * z2 * dsq;
                        double sinA = cen.y * pi.x - cen.x * pi.y;
                        double cosB = (z2 - z1) + 0.5 * z1 * dsq;
                        double cosbeta = -cosA * cosB + sinA * sinA;
                        double sinbeta = C == Sphere ? (cosA + cosB) * sinA
                                                     : ((z1 - z2) + (0.5 * z2 * dsq + cosB)) * sinA;
                        std::complex<double> e = expmsialpha(spin_of(D), std::complex<double>(cosbeta, sinbeta));
                        std::complex<double> t = cmul(std::complex<double>(d.gr, d.gi), e);
                        sr += t.real(); si += t.imag();
                    }
                }
                nd.gr = float(sr); nd.gi = float(si);
            }
        }
    }

    // CalculateSizeSq, src/Cell.cpp:89-100
    double size_sq(const Pt& cen, size_t start, size_t end) const {
        double s = 0.;
        for (size_t i = start; i < end; ++i) {
            const Pt& p = at(i).p;
            Pt d{cen.x - p.x, cen.y - p.y, cen.z - p.z};
            double dsq = normsq(C, d);
            if (dsq > s) s = dsq;
        }
        return s;
    }

    // Bounds<C>::getSplit, include/Bounds.h:65-66 (Flat) and :133-141 (ThreeD/Sphere)
    int split_dim(size_t start, size_t end, double* middle) const {
        Pt mn = at(start).p, mx = at(start).p;
        for (size_t i = start; i < end; ++i) {
            const Pt& p = at(i).p;
            mn.x = std::min(mn.x, p.x); mx.x = std::max(mx.x, p.x);
            mn.y = std::min(mn.y, p.y); mx.y = std::max(mx.y, p.y);
            mn.z = std::min(mn.z, p.z); mx.z = std::max(mx.z, p.z);
        }
        int split;
        if (C == Flat) {
            split = (mx.y - mn.y) > (mx.x - mn.x) ? 1 : 0;
        } else {
            double xr = mx.x - mn.x, yr = mx.y - mn.y, zr = mx.z - mn.z;
            split = yr > xr ? (zr > yr ? 2 : 1) : (zr > xr ? 2 : 0);
        }
        // Bounds::getMiddle, include/Bounds.h:67-68, :143-144
        *middle = split == 2 ? (mx.z + mn.z) / 2. : split == 1 ? (mx.y + mn.y) / 2. : (mx.x + mn.x) / 2.;
        return split;
    }

    // SplitData<C,SM>, src/Cell.cpp:520-554 with SplitDataCore Middle/Median/Mean
    // (:445-495).  std::partition / std::nth_element are the same libstdc++ calls
    // the reference makes, applied to the same sequence order.
    size_t split(size_t start, size_t end, const Pt& meanpos, int method) {
        double middle;
        int sd = split_dim(start, end, &middle);
        size_t mid;
        if (method == Median) {
            mid = (start + end) / 2;
            auto cmp = [&](int a, int b) { return coord_get(C, items[a].p, sd) < coord_get(C, items[b].p, sd); };
            std::nth_element(perm.begin() + start, perm.begin() + mid, perm.begin() + end, cmp);
        } else {
            double v = method == Middle ? middle : coord_get(C, meanpos, sd);
            auto pred = [&](int a) { return coord_get(C, items[a].p, sd) < v; };
            mid = std::partition(perm.begin() + start, perm.begin() + end, pred) - perm.begin();
        }
        if (mid == start || mid == end) return split(start, end, meanpos, Median);
        return mid;
    }

    // BuildCell, src/Cell.cpp:556-615.  Appends the subtree in pre-order.
    int build(std::vector<Node>& out, size_t start, size_t end, const Node* data, double sizesq) {
        int me = int(out.size());
        out.emplace_back();
        if (end - start == 1) {
            Node nd;
            if (data) nd = *data;
            else {
                const Item& it = at(start);
                nd.p = it.p; nd.w = it.w; nd.wk = it.wk; nd.gr = it.gr; nd.gi = it.gi; nd.n = 1;
            }
            nd.size = 0; nd.left = nd.right = -1;
            out[me] = nd;
            return me;
        }
        Node nd;
        if (data) nd = *data;
        else {
            cell_data(start, end, nd);
            finish(start, end, nd);
            sizesq = size_sq(nd.p, start, end);
        }
        if (sizesq > minsizesq) {
            double size = brute ? std::numeric_limits<double>::max() : std::sqrt(sizesq);
            nd.size = float(size);
            size_t mid = split(start, end, nd.p, sm);
            int l = build(out, start, mid, nullptr, 0.);
            int r = build(out, mid, end, nullptr, 0.);
            nd.left = l; nd.right = r;
        } else {
            nd.size = 0; nd.left = nd.right = -1;
        }
        out[me] = nd;
        return me;
    }
};

struct TopCell { size_t start, end; Node data; double sizesq; };

// SetupTopLevelCells, src/Field.cpp:27-81
static void setup_top(Builder& bd, double maxsizesq, size_t start, size_t end, int mintop,
                      int maxtop, std::vector<TopCell>& tops) {
    Node ave;
    double sizesq;
    if (end - start == 1) {
        const Item& it = bd.at(start);
        ave.p = it.p; ave.w = it.w; ave.wk = it.wk; ave.gr = it.gr; ave.gi = it.gi; ave.n = 1;
        sizesq = 0.;
    } else {
        bd.cell_data(start, end, ave);
        sizesq = bd.size_sq(ave.p, start, end);
    }
    if (sizesq == 0 || (sizesq <= maxsizesq && mintop <= 0) || maxtop <= 0) {
        if (end - start > 1) bd.finish(start, end, ave);
        tops.push_back({start, end, ave, sizesq});
    } else {
        size_t mid = bd.split(start, end, ave.p, bd.sm);
        setup_top(bd, maxsizesq, start, mid, mintop - 1, maxtop - 1, tops);
        setup_top(bd, maxsizesq, mid, end, mintop - 1, maxtop - 1, tops);
    }
}

Tree* build_tree(int D, int C, const double* x, const double* y, const double* z,
                 const double* d1, const double* d2, const double* w, const double* wpos,
                 long n, double minsize, double maxsize, int sm, bool brute,
                 int mintop, int maxtop) {
    // Field<D,C>::Field, src/Field.cpp:163-204, with CellDataHelper :83-144
    std::vector<Item> items(n);
    for (long i = 0; i < n; ++i) {
        Item& it = items[i];
        it.p.x = x[i]; it.p.y = y[i]; it.p.z = (C == Flat || !z) ? 0. : z[i];
        if (C == Sphere) sphere_normalize_recip(it.p);
        it.w = float(w[i]);
        it.wpos = wpos ? wpos[i] : w[i];
        it.idx = i;
        if (D == KData) it.wk = float(w[i] * d1[i]);
        if (is_complex(D)) { it.gr = float(w[i] * d1[i]); it.gi = float(w[i] * d2[i]); }
    }
    std::vector<int> perm(n);
    for (long i = 0; i < n; ++i) perm[i] = int(i);
    Tree* t = new Tree;
    t->D = D; t->C = C; t->nobj = n;
    Builder bd(D, C, sm, brute, minsize * minsize, items, perm);
    {
        Node root;
        bd.cell_data(0, n, root);
        bd.finish(0, n, root);
        t->center = root.p;
        t->sizesq = bd.size_sq(root.p, 0, n);
    }
    // DoBuildCells, src/Field.cpp:230-277
    std::vector<TopCell> tops;
    setup_top(bd, maxsize * maxsize, 0, n, mintop, maxtop, tops);
    std::vector<std::vector<Node>> sub(tops.size());
#pragma omp parallel for schedule(dynamic)
    for (long i = 0; i < long(tops.size()); ++i) {
        Builder b2(D, C, sm, brute, minsize * minsize, items, perm);
        b2.build(sub[i], tops[i].start, tops[i].end, &tops[i].data, tops[i].sizesq);
    }
    for (size_t i = 0; i < tops.size(); ++i) {
        int off = int(t->nodes.size());
        t->top.push_back(off);
        for (Node nd : sub[i]) {
            if (nd.left >= 0) { nd.left += off; nd.right += off; }
            t->nodes.push_back(nd);
        }
    }
    return t;
}

// ---------------------------------------------------------------------------
// 2-point dual-tree traversal (Log binning, Euclidean metric)
// ---------------------------------------------------------------------------

struct Params2 {
    int nbins;
    double minsep, maxsep, binsize, b, a;
    double logminsep, halfminsep, minsepsq, maxsepsq, bsq, asq;
};

// The Periodic metric's periods for the next 2-point runs (0: Euclidean); set by
// tco_set_periodic.  MetricHelper<Periodic>::DistSq, include/Metric.h:831-838, wraps
// each separation by Position::wrap, include/Position.h:106-112, :201-209.
static double g_period[3] = {0., 0., 0.};
// The Arc metric for the next 2-point runs (tco_set_arc); MetricHelper<Arc>,
// include/Metric.h:656-810.
static int g_arc = 0;
// The line-of-sight metrics for the next 2-point runs (tco_set_metric3): Rperp (2),
// Rlens (3), OldRperp (5), include/Metric.h:268-651; and the r_par range (ParHelper,
// include/Metric.h:37-106), used when g_rpar_on.
static int g_metric3 = 0;
static int g_rpar_on = 0;
static double g_minrpar = 0., g_maxrpar = 0.;
static inline double dot3(const Pt& a, const Pt& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline Pt cross3(const Pt& a, const Pt& b) {
    return Pt{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// ParHelper<1>::calculateRPar (r along the midpoint's direction); ParHelper<2> (OldRperp)
// the difference of the distances
static double calc_rpar(const Pt& a, const Pt& b) {
    if (g_metric3 == 5) return std::sqrt(dot3(b, b)) - std::sqrt(dot3(a, a));
    Pt r{b.x - a.x, b.y - a.y, b.z - a.z};
    Pt L{(a.x + b.x) * 0.5, (a.y + b.y) * 0.5, (a.z + b.z) * 0.5};
    return dot3(r, L) / std::sqrt(dot3(L, L));
}
struct PruneOut { bool skip; bool rin; };
// BaseCorr2::process11's pruning, src/Corr2.cpp:326-346: the r_par range, then the bin
// type's too-small/too-large tests (given) refined by the metric's own.
static PruneOut metric_prune(const Pt& a, const Pt& b, double rsq, double s1ps2, bool bt_small,
                             bool bt_large, double minsep, double minsepsq, double maxsep,
                             double maxsepsq) {
    PruneOut o{false, true};
    double rpar = 0.;
    if (g_rpar_on) {
        rpar = calc_rpar(a, b);
        if (rpar + s1ps2 < g_minrpar || rpar - s1ps2 > g_maxrpar) { o.skip = true; return o; }
    }
    bool small = bt_small, large = bt_large;
    if (g_metric3 == 2) {                  // MetricHelper<Rperp>, Metric.h:458-476
        Pt L{(a.x + b.x) * 0.5, (a.y + b.y) * 0.5, (a.z + b.z) * 0.5};
        double nL = dot3(L, L);
        if (small) {
            if (rsq < nL) small = true;
            else if (s1ps2 * s1ps2 > 4. * nL) small = false;
            else {
                double twoL = 2. * std::sqrt(nL);
                double t = minsep * (1 - s1ps2 / twoL) - s1ps2;
                small = rsq < t * t;
            }
        }
        if (large) {
            if (rsq < nL) large = true;
            else {
                double twoL = 2. * std::sqrt(nL);
                double t = maxsep * (1 + s1ps2 / twoL) + s1ps2;
                large = rsq > t * t;
            }
        }
    } else if (g_metric3 == 5 && (small || large)) {   // MetricHelper<OldRperp>, :300-318
        if (rpar == 0.) rpar = std::sqrt(dot3(b, b)) - std::sqrt(dot3(a, a));
        double d3 = std::sqrt(rpar * rpar + rsq);
        small = small && rsq + 2. * (d3 + std::fabs(rpar)) * s1ps2 < minsepsq;
        large = large && rsq - 2. * (d3 + std::fabs(rpar)) * s1ps2 > maxsepsq;
    }
    if (small || large) { o.skip = true; return o; }
    if (g_rpar_on) o.rin = !(rpar - s1ps2 < g_minrpar) && !(rpar + s1ps2 > g_maxrpar);
    return o;
}
static inline double wrap1(double d, double p) {
    while (d > p / 2.) d -= p;
    while (d < -p / 2.) d += p;
    return d;
}

struct Acc2 {
    int nb; bool exact;
    std::vector<KSum> v;   // 8 quantities x nbins: xi0..xi3, meanr, meanlogr, weight, npairs
    long long visits = 0, accepted = 0, leafpairs = 0, refined = 0;
    Acc2(int nbins, bool ex) : nb(nbins), exact(ex), v(8 * nbins) {}
    void add(int q, int k, double x) {
        KSum& s = v[q * nb + k];
        if (exact) s.add(x); else s.s += x;
    }
};

template <int D, int C>
struct Corr2Engine {
    const Params2& P;
    const Tree& T1;
    const Tree& T2;
    bool quick;
    Corr2Engine(const Params2& p, const Tree& t1, const Tree& t2, bool q) : P(p), T1(t1), T2(t2), quick(q) {}

    // MetricHelper<M>::DistSq; the Arc metric's 3-d form also scales the sizes to the
    // unit sphere
    static double dist_sq(const Pt& a, const Pt& b, double& s1, double& s2) {
        if (g_metric3 == 2) {     // MetricHelper<Rperp>::DistSq, include/Metric.h:416-434
            Pt c = cross3(a, b);
            Pt L{(a.x + b.x) * 0.5, (a.y + b.y) * 0.5, (a.z + b.z) * 0.5};
            double nL = dot3(L, L);
            double rperpsq = nL > 0. ? dot3(c, c) / nL : 4. * dot3(a, a);
            double r1sq = dot3(a, a), r2sq = dot3(b, b);
            if (r2sq > nL && s1 != 0) s1 *= std::sqrt(r2sq / nL);
            if (r1sq > nL && s2 != 0) s2 *= std::sqrt(r1sq / nL);
            return rperpsq;
        }
        if (g_metric3 == 5) {     // MetricHelper<OldRperp>::DistSq, include/Metric.h:277-293
            Pt r{a.x - b.x, a.y - b.y, a.z - b.z};
            double d3sq = dot3(r, r), r1sq = dot3(a, a), r2sq = dot3(b, b);
            double rparsq = (r1sq - r2sq) * (r1sq - r2sq) / (r1sq + r2sq + 2. * std::sqrt(r1sq * r2sq));
            const double big = std::numeric_limits<double>::max();
            if (r1sq < r2sq) {
                if (s1 != 0. && s1 < big) s1 *= (1. + 0.25 * (r2sq - r1sq) / r1sq);
            } else {
                if (s2 != 0. && s2 < big) s2 *= (1. + 0.25 * (r1sq - r2sq) / r2sq);
            }
            return std::fabs(d3sq - rparsq);
        }
        if (g_metric3 == 3) {     // MetricHelper<Rlens>::DistSq, include/Metric.h:582-589
            Pt c = cross3(a, b);
            double r2sq = dot3(b, b);
            double rsq = dot3(c, c) / r2sq;
            s2 *= std::sqrt(dot3(a, a) / r2sq);
            return rsq;
        }
        if (g_arc) {
            if (C == Sphere) {      // Metric.h:666-680: theta = 2 asin(L/2), L the chord
                Pt r{a.x - b.x, a.y - b.y, a.z - b.z};
                double th = 2. * std::asin(std::sqrt(normsq(ThreeD, r)) / 2.);
                return th * th;
            }
            // Metric.h:701-721: theta = atan2(|a x b|, a.b); s /= |p|
            Pt c{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
            double th = std::atan2(std::sqrt(normsq(ThreeD, c)),
                                   a.x * b.x + a.y * b.y + a.z * b.z);
            s1 /= std::sqrt(normsq(ThreeD, a));
            s2 /= std::sqrt(normsq(ThreeD, b));
            return th * th;
        }
        return dist_sq(a, b);
    }
    static double dist_sq(const Pt& a, const Pt& b) {   // Metric.h:143-147, :197-202
        Pt r{a.x - b.x, a.y - b.y, a.z - b.z};
        if (g_period[0] > 0.) {
            r.x = wrap1(r.x, g_period[0]); r.y = wrap1(r.y, g_period[1]);
            if (C != Flat) r.z = wrap1(r.z, g_period[2]);
        }
        return normsq(C, r);
    }

    // BinTypeHelper<Log>::singleBin, include/BinType.h:93-151
    bool single_bin(double rsq, double s1ps2, int& ik, double& r, double& logr, Acc2& acc) const {
        if (s1ps2 == 0.) return true;
        double s1ps2sq = s1ps2 * s1ps2;
        if (s1ps2sq > P.asq * rsq) return false;
        if (s1ps2sq <= P.bsq * rsq) return true;
        if (s1ps2sq > 0.25 * ((P.binsize + P.b) * (P.binsize + P.b)) * rsq) return false;
        acc.refined++;
        logr = 0.5 * std::log(rsq);
        double kk = (logr - P.logminsep) / P.binsize;
        ik = int(kk);
        double frackk = kk - ik;
        double f = std::min(frackk, 1. - frackk);
        double t1 = f * P.binsize + P.b;
        if (s1ps2sq > (t1 * t1) * rsq) return false;
        double t2 = frackk * P.binsize + P.b - s1ps2sq / rsq;
        if (s1ps2sq > (t2 * t2) * rsq) return false;
        r = std::sqrt(rsq);
        return true;
    }

    // BaseCorr2::directProcess11 src/Corr2.cpp:476-525 + finishProcess :527-566
    // + DirectHelper<0,1,3> :397-474.
    void direct(const Node& c1, const Node& c2, double rsq, int k, double r, double logr, Acc2& acc) const {
        acc.accepted++;
        if (c1.left < 0 && c2.left < 0) acc.leafpairs++;
        if (k < 0) {
            r = std::sqrt(rsq);
            logr = std::log(r);
            k = int((logr - P.logminsep) / P.binsize);
        }
        if (k == P.nbins) --k;
        double ww = double(c1.w) * double(c2.w);
        acc.add(6, k, ww);
        if (!quick) {
            acc.add(7, k, double(c1.n) * double(c2.n));
            acc.add(4, k, ww * r);
            acc.add(5, k, ww * logr);
        }
        if (D == KData) {
            acc.add(0, k, double(c1.wk) * double(c2.wk));
        } else if (is_complex(D)) {
            // DirectHelper<3>, src/Corr2.cpp:445-474, for a spin-s field with itself
            const int S = spin_of(D);
            std::complex<double> g1(c1.gr, c1.gi), g2(c2.gr, c2.gi);
            if (S == 0) {
            } else if (C == Flat) {   // ProjectHelper<Flat>::Project, ProjectHelper.h:126-147
                std::complex<double> rr(c2.p.x - c1.p.x, c2.p.y - c1.p.y);
                std::complex<double> e = expmsialpha(S, rr);
                g1 = cmul(g1, e); g2 = cmul(g2, e);
            } else {           // ProjectHelper<Sphere/ThreeD>::Project, :370-388, :573-589
                Pt p1 = c1.p, p2 = c2.p;
                if (C == ThreeD) { sphere_normalize_recip(p1); sphere_normalize_recip(p2); }
                g2 = cmul(g2, expmsialpha(S, sphere_direction(p1, p2)));
                g1 = cmul(g1, expmsialpha(S, sphere_direction(p2, p1)));
                if (S == 1 || S == 3) g1 = -g1;   // odd spins: p1 on the left
            }
            // both_complex_prod(g2, g1), ProjectHelper.h:84-101
            double z1rz2r = g2.real() * g1.real(), z1rz2i = g2.real() * g1.imag();
            double z1iz2r = g2.imag() * g1.real(), z1iz2i = g2.imag() * g1.imag();
            acc.add(2, k, z1rz2r - z1iz2i);   // xim
            acc.add(3, k, z1rz2i + z1iz2r);   // xim_im
            acc.add(0, k, z1rz2r + z1iz2i);   // xip
            acc.add(1, k, z1rz2i - z1iz2r);   // xip_im
        } else if (D == XK_NK) {
            // DirectHelper<1>, src/Corr2.cpp:408-420: getWK of an N cell is its w
            // (include/ProjectHelper.h:75-80)
            acc.add(0, k, double(c1.w) * double(c2.wk));
        } else if (mixed_d2(D) >= 0) {
            // DirectHelper<2>, src/Corr2.cpp:423-442: z2 projected onto the line to c1
            // (ProjectHelper<C>::Project(c1,c2,z2), include/ProjectHelper.h:107-113,
            // :348-355, :545-555), times getWK(c1), negated for a shear (tangential)
            const int d2 = mixed_d2(D), S = spin_of(d2);
            std::complex<double> g2(c2.gr, c2.gi);
            if (S == 0) {
            } else if (C == Flat) {
                g2 = cmul(g2, expmsialpha(S, std::complex<double>(c2.p.x - c1.p.x, c2.p.y - c1.p.y)));
            } else {
                Pt p1 = c1.p, p2 = c2.p;
                if (C == ThreeD) { sphere_normalize_recip(p1); sphere_normalize_recip(p2); }
                g2 = cmul(g2, expmsialpha(S, sphere_direction(p1, p2)));
            }
            double f = mixed_k1(D) ? double(c1.wk) : double(c1.w);
            if (d2 == GData) f = -f;
            acc.add(0, k, g2.real() * f);
            acc.add(1, k, g2.imag() * f);
        }
    }

    // BaseCorr2::process11, src/Corr2.cpp:308-389
    void process11(const Tree& ta, int i1, const Tree& tb, int i2, Acc2& acc) const {
        const Node& c1 = ta.nodes[i1];
        const Node& c2 = tb.nodes[i2];
        acc.visits++;
        if (c1.w == 0.f || c2.w == 0.f) return;
        double s1 = c1.size, s2 = c2.size;
        double rsq = dist_sq(c1.p, c2.p, s1, s2);
        double s1ps2 = s1 + s2;
        // BinTypeHelper<Log>::tooSmallDist / tooLargeDist, BinType.h:70-82
        const bool bt_small = rsq < P.minsepsq && s1ps2 < P.minsep &&
                              rsq < (P.minsep - s1ps2) * (P.minsep - s1ps2);
        const bool bt_large = rsq >= P.maxsepsq && rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
        bool rin = true;
        if (g_metric3 || g_rpar_on) {
            PruneOut po = metric_prune(c1.p, c2.p, rsq, s1ps2, bt_small, bt_large, P.minsep,
                                       P.minsepsq, P.maxsep, P.maxsepsq);
            if (po.skip) return;
            rin = po.rin;
        } else if (bt_small || bt_large) {
            return;
        }
        int k = -1; double r = 0, logr = 0;
        if (rin && single_bin(rsq, s1ps2, k, r, logr, acc)) {
            if (rsq >= P.minsepsq && rsq < P.maxsepsq) direct(c1, c2, rsq, k, r, logr, acc);
            return;
        }
        // CalcSplitSq, include/Split.h:82-101, with bsq_eff = rsq*min(bsq,asq)
        double bsq_eff = rsq * std::min(P.bsq, P.asq);
        bool split1 = false, split2 = false;
        calc_split_sq(split1, split2, s1, s2, bsq_eff);
        if (split1 && split2) {
            process11(ta, c1.left, tb, c2.left, acc);
            process11(ta, c1.left, tb, c2.right, acc);
            process11(ta, c1.right, tb, c2.left, acc);
            process11(ta, c1.right, tb, c2.right, acc);
        } else if (split1) {
            process11(ta, c1.left, tb, i2, acc);
            process11(ta, c1.right, tb, i2, acc);
        } else {
            process11(ta, i1, tb, c2.left, acc);
            process11(ta, i1, tb, c2.right, acc);
        }
    }

    static void calc_split_sq(bool& split1, bool& split2, double s1, double s2, double bsq) {
        const double splitfactorsq = 0.3422;
        if (split1 && split2) return;
        if (s2 > s1) calc_split_sq(split2, split1, s2, s1, bsq);
        else if (s1 > 2. * s2) split1 = true;
        else { split1 = true; split2 = s2 * s2 > splitfactorsq * bsq; }
    }

    // BaseCorr2::process2, src/Corr2.cpp:290-306
    void process2(const Tree& t, int i, Acc2& acc) const {
        const Node& c = t.nodes[i];
        if (c.w == 0.f) return;
        if (double(c.size) <= P.halfminsep) return;
        process2(t, c.left, acc);
        process2(t, c.right, acc);
        process11(t, c.left, t, c.right, acc);
    }
};

template <int D, int C>
static void run_corr2(const Params2& P, const Tree& t1, const Tree* t2, bool quick, bool exact,
                      int nthreads, double* out[8], long long* counters, int row_stride) {
    Corr2Engine<D, C> eng(P, t1, t2 ? *t2 : t1, quick);
    const long n1 = long(t1.top.size());
    std::vector<std::unique_ptr<Acc2>> accs;
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    for (int i = 0; i < nt; ++i) accs.emplace_back(new Acc2(P.nbins, exact));
    if (t2) {
        // BaseCorr2::process(f1,f2) early exit, src/Corr2.cpp:215-231
        double s1 = std::sqrt(t1.sizesq), s2 = std::sqrt(t2->sizesq);
        double rsq = Corr2Engine<D, C>::dist_sq(t1.center, t2->center, s1, s2);
        double s1ps2 = s1 + s2;
        bool too_small = rsq < P.minsepsq && s1ps2 < P.minsep && rsq < (P.minsep - s1ps2) * (P.minsep - s1ps2);
        bool too_large = rsq >= P.maxsepsq && rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
        if (g_metric3 || g_rpar_on) {
            too_small = metric_prune(t1.center, t2->center, rsq, s1ps2, too_small, too_large,
                                     P.minsep, P.minsepsq, P.maxsep, P.maxsepsq).skip;
            too_large = false;
        }
        if (!(too_small || too_large)) {
            const long n2 = long(t2->top.size());
#pragma omp parallel for schedule(dynamic) num_threads(nt)
            for (long i = 0; i < n1; ++i) {
                int tid = 0;
#ifdef _OPENMP
                tid = omp_get_thread_num();
#endif
                if (i % row_stride != 0) continue;
                for (long j = 0; j < n2; ++j) eng.process11(t1, t1.top[i], *t2, t2->top[j], *accs[tid]);
            }
        }
    } else {
        // BaseCorr2::process(field), src/Corr2.cpp:146-205
#pragma omp parallel for schedule(dynamic) num_threads(nt)
        for (long i = 0; i < n1; ++i) {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            if (i % row_stride != 0) continue;
            eng.process2(t1, t1.top[i], *accs[tid]);
            for (long j = i + 1; j < n1; ++j) eng.process11(t1, t1.top[i], t1, t1.top[j], *accs[tid]);
        }
    }
    // Merge per-thread accumulators (the reference merges under omp critical,
    // src/Corr2.cpp:197-201) and add into the caller's arrays.
    for (int q = 0; q < 8; ++q) {
        if (!out[q]) continue;
        for (int k = 0; k < P.nbins; ++k) {
            KSum s;
            for (auto& a : accs) {
                const KSum& x = a->v[q * P.nbins + k];
                if (exact) { s.add(x.s); s.add(x.c); } else s.s += x.s;
            }
            out[q][k] += exact ? s.get() : s.s;
        }
    }
    if (counters) {
        for (auto& a : accs) {
            counters[0] += a->visits; counters[1] += a->accepted;
            counters[2] += a->leafpairs; counters[3] += a->refined;
        }
    }
}

}  // namespace tco

using namespace tco;

extern "C" {

void* tco_field_new(int d, int coords, const double* x, const double* y, const double* z,
                    const double* d1, const double* d2, const double* w, const double* wpos,
                    long n, double minsize, double maxsize, int sm, int brute, int mintop,
                    int maxtop) {
    return build_tree(d, coords, x, y, z, d1, d2, w, wpos, n, minsize, maxsize, sm, brute != 0,
                      mintop, maxtop);
}

void tco_field_free(void* h) { delete static_cast<Tree*>(h); }
void tco_set_periodic(double xp, double yp, double zp) {
    g_period[0] = xp; g_period[1] = yp; g_period[2] = zp;
}
void tco_set_arc(int on) { g_arc = on; }
void tco_set_metric3(int metric, int rpar_on, double minrpar, double maxrpar) {
    g_metric3 = metric; g_rpar_on = rpar_on; g_minrpar = minrpar; g_maxrpar = maxrpar;
}
long tco_field_nnodes(void* h) { return long(static_cast<Tree*>(h)->nodes.size()); }
long tco_field_ntop(void* h) { return long(static_cast<Tree*>(h)->top.size()); }
double tco_field_sizesq(void* h) { return static_cast<Tree*>(h)->sizesq; }
void tco_field_center(void* h, double* c) {
    Tree* t = static_cast<Tree*>(h);
    c[0] = t->center.x; c[1] = t->center.y; c[2] = t->center.z;
}

// Dump the tree in pre-order: pos (3 per node), size, w, wk, g (2 per node), n, left, right.
void tco_field_dump(void* h, double* pos, float* size, float* w, float* wk, float* g, long* n,
                    int* left, int* right, int* top) {
    Tree* t = static_cast<Tree*>(h);
    for (size_t i = 0; i < t->nodes.size(); ++i) {
        const Node& nd = t->nodes[i];
        pos[3 * i] = nd.p.x; pos[3 * i + 1] = nd.p.y; pos[3 * i + 2] = nd.p.z;
        size[i] = nd.size; w[i] = nd.w; wk[i] = nd.wk; g[2 * i] = nd.gr; g[2 * i + 1] = nd.gi;
        n[i] = nd.n; left[i] = nd.left; right[i] = nd.right;
    }
    for (size_t i = 0; i < t->top.size(); ++i) top[i] = t->top[i];
}

// 2-point correlation over one tree (auto, f2 == NULL) or two trees (cross).
// d: a DType for NN, KK, ZZ, VV, GG, TT, QQ; an XK_* code for the mixed cross
// correlations (f1 the N or K field, f2 the other).
// out arrays (nbins each, NULL allowed): xi0..xi3, meanr, meanlogr, weight, npairs.
// Values are ADDED to the arrays.  counters[4]: node-pair visits, accepted pairs,
// accepted leaf-leaf pairs, refined singleBin evaluations.  row_stride > 1 keeps
// only the top-level rows i with i % row_stride == 0 (a bounded timing sample).
int tco_corr2_sampled(void* f1, void* f2, int d, int coords, int nbins, double minsep,
                      double maxsep, double binsize, double b, double a, int quick, int exact,
                      int nthreads, double* xi0, double* xi1, double* xi2, double* xi3,
                      double* meanr, double* meanlogr, double* weight, double* npairs,
                      long long* counters, int row_stride) {
    Params2 P;
    P.nbins = nbins; P.minsep = minsep; P.maxsep = maxsep; P.binsize = binsize; P.b = b; P.a = a;
    // BaseCorr2::BaseCorr2, src/Corr2.cpp:51-77
    P.logminsep = std::log(minsep); P.halfminsep = 0.5 * minsep;
    P.minsepsq = minsep * minsep; P.maxsepsq = maxsep * maxsep;
    P.bsq = b * b; P.asq = a * a;
    double* out[8] = {xi0, xi1, xi2, xi3, meanr, meanlogr, weight, npairs};
    Tree* t1 = static_cast<Tree*>(f1);
    Tree* t2 = static_cast<Tree*>(f2);
    if (row_stride < 1) row_stride = 1;
#define TCO_RUN(DD, CC) run_corr2<DD, CC>(P, *t1, t2, quick != 0, exact != 0, nthreads, out, counters, row_stride)
    if (d == NData) {
        if (coords == Flat) TCO_RUN(NData, Flat); else if (coords == ThreeD) TCO_RUN(NData, ThreeD); else TCO_RUN(NData, Sphere);
    } else if (d == KData) {
        if (coords == Flat) TCO_RUN(KData, Flat); else if (coords == ThreeD) TCO_RUN(KData, ThreeD); else TCO_RUN(KData, Sphere);
    } else if (d == GData) {
        if (coords == Flat) TCO_RUN(GData, Flat); else if (coords == ThreeD) TCO_RUN(GData, ThreeD); else TCO_RUN(GData, Sphere);
    } else if (d == ZData) {
        if (coords == Flat) TCO_RUN(ZData, Flat); else if (coords == ThreeD) TCO_RUN(ZData, ThreeD); else TCO_RUN(ZData, Sphere);
    } else if (d == VData) {
        if (coords == Flat) TCO_RUN(VData, Flat); else if (coords == ThreeD) TCO_RUN(VData, ThreeD); else TCO_RUN(VData, Sphere);
    } else if (d == TData) {
        if (coords == Flat) TCO_RUN(TData, Flat); else if (coords == ThreeD) TCO_RUN(TData, ThreeD); else TCO_RUN(TData, Sphere);
    } else if (d == QData) {
        if (coords == Flat) TCO_RUN(QData, Flat); else if (coords == ThreeD) TCO_RUN(QData, ThreeD); else TCO_RUN(QData, Sphere);
#define TCO_MX(KK) } else if (d == KK && t2) { \
        if (coords == Flat) TCO_RUN(KK, Flat); else if (coords == ThreeD) TCO_RUN(KK, ThreeD); else TCO_RUN(KK, Sphere);
    TCO_MX(XK_NV) TCO_MX(XK_NT) TCO_MX(XK_NQ) TCO_MX(XK_KV) TCO_MX(XK_KT) TCO_MX(XK_KQ)
#undef TCO_MX
    } else if (d == XK_NZ && t2) {
        if (coords == Flat) TCO_RUN(XK_NZ, Flat); else if (coords == ThreeD) TCO_RUN(XK_NZ, ThreeD); else TCO_RUN(XK_NZ, Sphere);
    } else if (d == XK_KZ && t2) {
        if (coords == Flat) TCO_RUN(XK_KZ, Flat); else if (coords == ThreeD) TCO_RUN(XK_KZ, ThreeD); else TCO_RUN(XK_KZ, Sphere);
    } else if (d == XK_NK && t2) {
        if (coords == Flat) TCO_RUN(XK_NK, Flat); else if (coords == ThreeD) TCO_RUN(XK_NK, ThreeD); else TCO_RUN(XK_NK, Sphere);
    } else if (d == XK_NG && t2) {
        if (coords == Flat) TCO_RUN(XK_NG, Flat); else if (coords == ThreeD) TCO_RUN(XK_NG, ThreeD); else TCO_RUN(XK_NG, Sphere);
    } else if (d == XK_KG && t2) {
        if (coords == Flat) TCO_RUN(XK_KG, Flat); else if (coords == ThreeD) TCO_RUN(XK_KG, ThreeD); else TCO_RUN(XK_KG, Sphere);
    } else {
        return 1;
    }
#undef TCO_RUN
    return 0;
}

int tco_corr2(void* f1, void* f2, int d, int coords, int nbins, double minsep, double maxsep,
              double binsize, double b, double a, int quick, int exact, int nthreads,
              double* xi0, double* xi1, double* xi2, double* xi3, double* meanr,
              double* meanlogr, double* weight, double* npairs, long long* counters) {
    return tco_corr2_sampled(f1, f2, d, coords, nbins, minsep, maxsep, binsize, b, a, quick,
                             exact, nthreads, xi0, xi1, xi2, xi3, meanr, meanlogr, weight,
                             npairs, counters, 1);
}

// Brute-force O(N1*N2) pair loop over the objects of two catalogs (auto when
// the second catalog is NULL: i<j), with the reference's float leaf payloads and
// the binning of directProcess11.  Reference tests restate the same loops
// (tests/test_nn.py:448-463, tests/test_kk.py, tests/test_gg.py:54-84).
int tco_corr2_brute(int d, int coords, long n1, const double* x1, const double* y1,
                    const double* z1, const double* a1, const double* b1, const double* w1,
                    long n2, const double* x2, const double* y2, const double* z2,
                    const double* a2, const double* b2, const double* w2, int nbins,
                    double minsep, double maxsep, double binsize, int exact, int nthreads,
                    double* xi0, double* xi1, double* xi2, double* xi3, double* meanr,
                    double* meanlogr, double* weight, double* npairs) {
    bool autoc = (x2 == nullptr);
    auto mk = [&](long n, const double* x, const double* y, const double* z, const double* a,
                  const double* b, const double* w) {
        std::vector<Node> v(n);
        for (long i = 0; i < n; ++i) {
            Node& nd = v[i];
            nd.p.x = x[i]; nd.p.y = y[i]; nd.p.z = (coords == Flat || !z) ? 0. : z[i];
            if (coords == Sphere) sphere_normalize_recip(nd.p);
            nd.w = float(w[i]); nd.n = 1;
            if (d == KData) nd.wk = float(w[i] * a[i]);
            if (d == GData) { nd.gr = float(w[i] * a[i]); nd.gi = float(w[i] * b[i]); }
        }
        return v;
    };
    std::vector<Node> A = mk(n1, x1, y1, z1, a1, b1, w1);
    std::vector<Node> B = autoc ? std::vector<Node>() : mk(n2, x2, y2, z2, a2, b2, w2);
    const std::vector<Node>& Bref = autoc ? A : B;
    long nb2 = autoc ? n1 : n2;
    Params2 P;
    P.nbins = nbins; P.minsep = minsep; P.maxsep = maxsep; P.binsize = binsize; P.b = 0; P.a = 0;
    P.logminsep = std::log(minsep); P.halfminsep = 0.5 * minsep;
    P.minsepsq = minsep * minsep; P.maxsepsq = maxsep * maxsep; P.bsq = 0; P.asq = 0;
    Tree dummy;
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    std::vector<std::unique_ptr<Acc2>> accs;
    for (int i = 0; i < nt; ++i) accs.emplace_back(new Acc2(nbins, exact != 0));
    auto body = [&](auto eng) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
        for (long i = 0; i < n1; ++i) {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            for (long j = autoc ? i + 1 : 0; j < nb2; ++j) {
                double rsq = eng.dist_sq(A[i].p, Bref[j].p);
                if (rsq >= P.minsepsq && rsq < P.maxsepsq) eng.direct(A[i], Bref[j], rsq, -1, 0, 0, *accs[tid]);
            }
        }
    };
#define TCO_BR(DD, CC) body(Corr2Engine<DD, CC>(P, dummy, dummy, false))
    if (d == NData) {
        if (coords == Flat) TCO_BR(NData, Flat); else if (coords == ThreeD) TCO_BR(NData, ThreeD); else TCO_BR(NData, Sphere);
    } else if (d == KData) {
        if (coords == Flat) TCO_BR(KData, Flat); else if (coords == ThreeD) TCO_BR(KData, ThreeD); else TCO_BR(KData, Sphere);
    } else if (d == GData) {
        if (coords == Flat) TCO_BR(GData, Flat); else if (coords == ThreeD) TCO_BR(GData, ThreeD); else TCO_BR(GData, Sphere);
    } else return 1;
#undef TCO_BR
    double* out[8] = {xi0, xi1, xi2, xi3, meanr, meanlogr, weight, npairs};
    for (int q = 0; q < 8; ++q) {
        if (!out[q]) continue;
        for (int k = 0; k < nbins; ++k) {
            KSum s;
            for (auto& a : accs) { const KSum& x = a->v[q * nbins + k]; s.add(x.s); s.add(x.c); }
            out[q][k] += s.get();
        }
    }
    return 0;
}

}  // extern "C"
