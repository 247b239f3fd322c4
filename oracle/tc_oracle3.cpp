// TEST INFRASTRUCTURE ONLY -- the CPU oracle for the treecorr_amd multipole engine.
//
// Restates TreeCorr 5.1.3's 3-point multipole algorithm (Porth et al.) for the
// auto-correlation of one field, NNN and GGG, with the Log two-point tests it uses
// (ValidMPB, src/Corr3.cpp:3647-3652):
//   BaseCorr3::multipole          src/Corr3.cpp:1124-1175
//   splitC2Cells                  src/Corr3.cpp:1323-1372
//   multipoleSplit1               src/Corr3.cpp:1374-1408
//   splitC2CellsOrCalculateGn     src/Corr3.cpp:1452-1546
//   multipoleFinish               src/Corr3.cpp:1548-1613
//   Corr3::calculateGn            src/Corr3.cpp:1681-1704
//   MultipoleScratch<N/G>::calculateGn src/Corr3.cpp:1706-1729, 1767-1828
//   Corr3::calculateZeta          src/Corr3.cpp:1861-1974
//   MultipoleHelper<5>::CalculateZeta src/Corr3.cpp:2582-2849 (GGG)
// Scratch layout follows include/MultipoleScratch.h (Windex :80-81, Gindex :334-335).
// Used by tests/ and bench.py only; the product never links it.
#include "tc_oracle.h"

#include <algorithm>
#include <cstring>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace tco {

Tree* build_tree(int D, int C, const double* x, const double* y, const double* z,
                 const double* d1, const double* d2, const double* w, const double* wpos,
                 long n, double minsize, double maxsize, int sm, bool brute,
                 int mintop, int maxtop);

typedef std::complex<double> cd;

struct P3 {
    int nbins, maxn, D;
    double minsep, maxsep, binsize, b, a;
    double logminsep, minsepsq, maxsepsq, bsq, asq;
};

// Per-node sums over the leaves below (BaseCell::calculateSumWSq etc.,
// src/Cell.cpp:143-168): a leaf (size 0 or N==1) contributes its own value,
// otherwise left + right.
struct NodeSums { double wsq; cd wzsq; double abswzsq; };

static void node_sums(const Tree& t, std::vector<NodeSums>& out) {
    out.assign(t.nodes.size(), NodeSums{0, 0, 0});
    // children have larger indices than parents in pre-order: sweep backwards
    for (long i = long(t.nodes.size()) - 1; i >= 0; --i) {
        const Node& nd = t.nodes[i];
        if (nd.size == 0.f || nd.n == 1 || nd.left < 0) {
            double w = nd.w;
            cd wz(nd.gr, nd.gi);
            out[i].wsq = w * w;
            out[i].wzsq = cd(wz.real() * wz.real() - wz.imag() * wz.imag(),
                             wz.real() * wz.imag() + wz.imag() * wz.real());
            out[i].abswzsq = wz.real() * wz.real() + wz.imag() * wz.imag();
        } else {
            const NodeSums& l = out[nd.left];
            const NodeSums& r = out[nd.right];
            out[i].wsq = l.wsq + r.wsq;
            out[i].wzsq = l.wzsq + r.wzsq;
            out[i].abswzsq = l.abswzsq + r.abswzsq;
        }
    }
}

// MultipoleScratch (include/MultipoleScratch.h), GData with buffer 1 or NData.
struct Scratch {
    int nbins, maxn, D, wstride, gstride;
    std::vector<cd> Wn, Gn;
    std::vector<double> npairs, sumw, sumwr, sumwlogr, sumww, sumwwr, sumwwlogr;
    std::vector<cd> gg0, gg1, gg2;
    Scratch(int nb, int mn, int D_) : nbins(nb), maxn(mn), D(D_) {
        wstride = maxn + 1;          // NData buffer 0 for NNN (D1 <= KData -> buffer 0)
        gstride = 2 * (maxn + 1) + 1;
        Wn.resize(size_t(nbins) * wstride);
        if (D == GData) { Gn.resize(size_t(nbins) * gstride); gg0.resize(nbins); gg1.resize(nbins); gg2.resize(nbins); }
        npairs.resize(nbins); sumw.resize(nbins); sumwr.resize(nbins); sumwlogr.resize(nbins);
        sumww.resize(nbins); sumwwr.resize(nbins); sumwwlogr.resize(nbins);
    }
    void clear() {
        std::fill(Wn.begin(), Wn.end(), cd(0));
        std::fill(Gn.begin(), Gn.end(), cd(0));
        for (auto* v : {&npairs, &sumw, &sumwr, &sumwlogr, &sumww, &sumwwr, &sumwwlogr})
            std::fill(v->begin(), v->end(), 0.);
        for (auto* v : {&gg0, &gg1, &gg2}) std::fill(v->begin(), v->end(), cd(0));
    }
    int Windex(int k) const { return k * wstride; }
    int Gindex(int k) const { return (2 * k + 1) * (maxn + 1) + k; }
};

struct Out3 {
    int ntot;
    std::vector<double> z[8];
    std::vector<double> meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri;
    long long gn_calls = 0, zeta_calls = 0, zeta_rows = 0, stages = 0;
    explicit Out3(int n) : ntot(n) {
        for (auto& v : z) v.assign(n, 0.);
        for (auto* v : {&meand2, &meanlogd2, &meand3, &meanlogd3, &weight, &weight_im, &ntri})
            v->assign(n, 0.);
    }
};

// c1 cells come from T1; the cells of a list from T (and, for the second list of the
// three-field form, T3), with their node sums S (S3).
template <int C>
struct MP {
    const P3& P;
    const Tree& T1;
    const Tree& T;
    const std::vector<NodeSums>& S;
    Out3& O;
    const Tree* T3 = nullptr;
    const std::vector<NodeSums>* S3 = nullptr;
    MP(const P3& p, const Tree& t1, const Tree& t, const std::vector<NodeSums>& s, Out3& o)
        : P(p), T1(t1), T(t), S(s), O(o) {}

    static double dist_sq(const Pt& a, const Pt& b) {
        Pt r{a.x - b.x, a.y - b.y, a.z - b.z};
        return normsq(C, r);
    }

    // BinTypeHelper<Log>::singleBin, include/BinType.h:93-151
    bool single_bin(double rsq, double s1ps2, int& ik, double& r, double& logr) const {
        if (s1ps2 == 0.) return true;
        double s1ps2sq = s1ps2 * s1ps2;
        if (s1ps2sq > P.asq * rsq) return false;
        if (s1ps2sq <= P.bsq * rsq) return true;
        if (s1ps2sq > 0.25 * ((P.binsize + P.b) * (P.binsize + P.b)) * rsq) return false;
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
    bool too_small(double rsq, double s1ps2) const {
        return rsq < P.minsepsq && s1ps2 < P.minsep && rsq < (P.minsep - s1ps2) * (P.minsep - s1ps2);
    }
    bool too_large(double rsq, double s1ps2) const {
        return rsq >= P.maxsepsq && rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
    }

    // ProjectHelper<C>::ExpIPhi (Flat :264-266, Sphere :526-535, ThreeD :681-688)
    cd exp_iphi(const Pt& p1, const Pt& p2, double r) const {
        if (C == Flat) return cd((p2.x - p1.x) / r, (p2.y - p1.y) / r);
        Pt a = p1, b = p2;
        if (C == ThreeD) { sphere_normalize_recip(a); sphere_normalize_recip(b); }
        cd z = sphere_direction(b, a);
        double n = std::sqrt(safe_norm(z));
        return cd(z.real() / n, z.imag() / n);
    }
    // _expmsialpha<GData> of the projection in ProjectWithSq (ProjectHelper.h:115-124, 357-368)
    cd proj2(const Pt& p1, const Pt& p2) const {
        if (C == Flat) return expm2ialpha(cd(p2.x - p1.x, p2.y - p1.y));
        Pt a = p1, b = p2;
        if (C == ThreeD) { sphere_normalize_recip(a); sphere_normalize_recip(b); }
        return expm2ialpha(sphere_direction(a, b));
    }

    // Corr3::calculateGn + MultipoleScratch::calculateGn
    void calculate_gn(const Node& c1, const Tree& TL, const std::vector<NodeSums>& S, int i2,
                      double r, double logr, int k, Scratch& mp) const {
        const Node& c2 = TL.nodes[i2];
        O.gn_calls++;
        double n = double(c2.n), w = c2.w;
        mp.npairs[k] += n;
        mp.sumw[k] += w;
        mp.sumwr[k] += w * r;
        mp.sumwlogr[k] += w * logr;
        double wsq = S[i2].wsq;
        mp.sumww[k] += wsq;
        mp.sumwwr[k] += wsq * r;
        mp.sumwwlogr[k] += wsq * logr;
        cd z = exp_iphi(c1.p, c2.p, r);
        int iw = mp.Windex(k);
        mp.Wn[iw] += w;
        cd wz = w;
        for (int m = 1; m <= P.maxn; ++m) { wz = cmul(wz, z); mp.Wn[iw + m] += wz; }
        if (P.D == GData) {
            cd wg(c2.gr, c2.gi);
            cd wgsq = S[i2].wzsq;
            cd e = proj2(c1.p, c2.p);
            wg = cmul(wg, e);
            wgsq = cmul(wgsq, cmul(e, e));
            cd abswgsq = S[i2].abswzsq;
            cd zsq = cmul(z, z);
            cd zsqc = std::conj(zsq);
            mp.gg0[k] += cmul(wgsq, zsqc);
            mp.gg1[k] += cmul(wgsq, zsq);
            mp.gg2[k] += cmul(abswgsq, zsqc);
            int ig = mp.Gindex(k);
            mp.Gn[ig] += wg;
            cd t = wg;
            for (int m = 1; m <= P.maxn + 1; ++m) { t = cmul(t, z); mp.Gn[ig + m] += t; }
            t = wg;
            cd zc = std::conj(z);
            for (int m = 1; m <= P.maxn + 1; ++m) { t = cmul(t, zc); mp.Gn[ig - m] += t; }
        }
    }

    static double cnorm(const cd& z) { return z.real() * z.real() + z.imag() * z.imag(); }

    // Corr3::calculateZeta (src/Corr3.cpp:1861-1974) + MultipoleHelper<0>/<5>
    void calculate_zeta(const Node& c1, Scratch& mp, int kstart, int mink_zeta) const {
        O.zeta_calls++;
        O.zeta_rows += std::max(0, mink_zeta - kstart);
        const int nb = P.nbins, maxn = P.maxn, nn = 2 * maxn + 1;
        double w1 = c1.w, n1 = double(c1.n);
        for (int k2 = kstart; k2 < mink_zeta; ++k2) {
            int i22 = (k2 * nb + k2) * nn + maxn;
            O.ntri[i22] += n1 * mp.npairs[k2] * (mp.npairs[k2] - 1);
            O.meand2[i22] += w1 * (mp.sumw[k2] * mp.sumwr[k2] - mp.sumwwr[k2]);
            O.meanlogd2[i22] += w1 * (mp.sumw[k2] * mp.sumwlogr[k2] - mp.sumwwlogr[k2]);
            O.meand3[i22] += w1 * (mp.sumw[k2] * mp.sumwr[k2] - mp.sumwwr[k2]);
            O.meanlogd3[i22] += w1 * (mp.sumw[k2] * mp.sumwlogr[k2] - mp.sumwwlogr[k2]);
            for (int k3 = k2 + 1; k3 < nb; ++k3) {
                int i23 = (k2 * nb + k3) * nn + maxn;
                int i32 = (k3 * nb + k2) * nn + maxn;
                double nnn = n1 * mp.npairs[k2] * mp.npairs[k3];
                O.ntri[i23] += nnn; O.ntri[i32] += nnn;
                double ww12 = w1 * mp.sumw[k2], ww13 = w1 * mp.sumw[k3];
                double wwwd2 = ww13 * mp.sumwr[k2];
                O.meand2[i23] += wwwd2; O.meand3[i32] += wwwd2;
                double wwwlogd2 = ww13 * mp.sumwlogr[k2];
                O.meanlogd2[i23] += wwwlogd2; O.meanlogd3[i32] += wwwlogd2;
                double wwwd3 = ww12 * mp.sumwr[k3];
                O.meand3[i23] += wwwd3; O.meand2[i32] += wwwd3;
                double wwwlogd3 = ww12 * mp.sumwlogr[k3];
                O.meanlogd3[i23] += wwwlogd3; O.meanlogd2[i32] += wwwlogd3;
            }
        }
        const int step23 = nn, step32 = nb * step23, step22 = step32 + step23;
        for (int k2 = kstart; k2 < mink_zeta; ++k2) {
            int iz22 = maxn + k2 * step22;
            int iw2 = mp.Windex(k2);
            O.weight[iz22] += w1 * (cnorm(mp.Wn[iw2]) - mp.sumww[k2]);
            for (int m = 1; m <= maxn; ++m) {
                double www = w1 * (cnorm(mp.Wn[iw2 + m]) - mp.sumww[k2]);
                O.weight[iz22 + m] += www;
                O.weight[iz22 - m] += www;
            }
            int iz23 = iz22 + step23, iz32 = iz22 + step32;
            for (int k3 = k2 + 1; k3 < nb; ++k3, iz23 += step23, iz32 += step32) {
                int iw3 = mp.Windex(k3);
                for (int m = 0; m <= maxn; ++m) {
                    cd www = cmul(cmul(cd(w1, 0.), mp.Wn[iw2 + m]), std::conj(mp.Wn[iw3 + m]));
                    O.weight[iz23 + m] += www.real(); O.weight_im[iz23 + m] += www.imag();
                    O.weight[iz32 + m] += www.real(); O.weight_im[iz32 + m] -= www.imag();
                    if (m > 0) {
                        O.weight[iz23 - m] += www.real(); O.weight_im[iz23 - m] -= www.imag();
                        O.weight[iz32 - m] += www.real(); O.weight_im[iz32 - m] += www.imag();
                    }
                }
            }
        }
        if (P.D == GData) zeta_ggg(cd(c1.gr, c1.gi), mp, kstart, mink_zeta);
    }

    void add8(int idx, const cd& g0, const cd& g1, const cd& g2, const cd& g3) const {
        O.z[0][idx] += g0.real(); O.z[1][idx] += g0.imag();
        O.z[2][idx] += g1.real(); O.z[3][idx] += g1.imag();
        O.z[4][idx] += g2.real(); O.z[5][idx] += g2.imag();
        O.z[6][idx] += g3.real(); O.z[7][idx] += g3.imag();
    }

    // MultipoleHelper<5>::CalculateZeta, src/Corr3.cpp:2582-2849 (the #else branch)
    void zeta_ggg(const cd& wg1, Scratch& mp, int kstart, int mink_zeta) const {
        const int nb = P.nbins, maxn = P.maxn;
        const int step23 = 2 * maxn + 1, step32 = nb * step23, step22 = step32 + step23;
        const cd wg1c = std::conj(wg1);
        auto G = [&](int ig, int m) { return mp.Gn[ig + m]; };
        for (int k2 = kstart; k2 < mink_zeta; ++k2) {
            int iz22 = maxn + k2 * step22;
            const int ig2 = mp.Gindex(k2);
            const cd g1wwgg0 = cmul(wg1, mp.gg0[k2]);
            const cd g1cwwgg1 = cmul(wg1c, mp.gg1[k2]);
            const cd g1wwgg2 = cmul(wg1, mp.gg2[k2]);
            const cd g1G2m = cmul(wg1, G(ig2, -1));
            const cd g1cG2p = cmul(wg1c, G(ig2, 1));
            const cd gam0 = cmul(g1G2m, G(ig2, -1)) - g1wwgg0;
            const cd gam1 = cmul(g1cG2p, G(ig2, 1)) - g1cwwgg1;
            const cd gam2 = cmul(g1G2m, std::conj(G(ig2, 1))) - g1wwgg2;
            add8(iz22, gam0, gam1, gam2, gam2);
            for (int m = 1; m <= maxn; ++m) {
                const cd G2pp = G(ig2, m + 1), G2pm = G(ig2, m - 1);
                const cd G2mp = G(ig2, -m + 1), G2mm = G(ig2, -m - 1);
                const cd g1G2mm = cmul(wg1, G2mm);
                const cd g1cG2pp = cmul(wg1c, G2pp);
                const cd g0 = cmul(g1G2mm, G2pm) - g1wwgg0;
                const cd g1 = cmul(g1cG2pp, G2mp) - g1cwwgg1;
                const cd g2 = cmul(std::conj(g1cG2pp), G2pm) - g1wwgg2;
                const cd g3 = cmul(g1G2mm, std::conj(G2mp)) - g1wwgg2;
                add8(iz22 + m, g0, g1, g2, g3);
                add8(iz22 - m, g0, g1, g3, g2);
            }
            int iz23 = iz22 + step23, iz32 = iz22 + step32;
            for (int k3 = k2 + 1; k3 < nb; ++k3, iz23 += step23, iz32 += step32) {
                const int ig3 = mp.Gindex(k3);
                const cd G2p = G(ig2, 1), G2m = G(ig2, -1), G3p = G(ig3, 1), G3m = G(ig3, -1);
                const cd g1G2m0 = cmul(wg1, G2m);
                const cd g1cG2p0 = cmul(wg1c, G2p);
                const cd a0 = cmul(g1G2m0, G3m), a1 = cmul(g1cG2p0, G3p);
                const cd a2 = cmul(g1G2m0, std::conj(G3p)), a3 = cmul(std::conj(g1cG2p0), G3m);
                add8(iz23, a0, a1, a2, a3);
                add8(iz32, a0, a1, a3, a2);
                for (int m = 1; m <= maxn; ++m) {
                    const cd G2pp = G(ig2, m + 1), G2pm = G(ig2, m - 1);
                    const cd G2mp = G(ig2, -m + 1), G2mm = G(ig2, -m - 1);
                    const cd G3pp = G(ig3, m + 1), G3pm = G(ig3, m - 1);
                    const cd G3mp = G(ig3, -m + 1), G3mm = G(ig3, -m - 1);
                    const cd g1G2pm = cmul(wg1, G2pm);
                    const cd g1cG2pp = cmul(wg1c, G2pp);
                    const cd g1G2mpc = cmul(wg1, std::conj(G2mp));
                    const cd g1G2mm = cmul(wg1, G2mm);
                    const cd p0 = cmul(g1G2pm, G3mm), p1 = cmul(g1cG2pp, G3mp);
                    const cd p2 = cmul(g1G2pm, std::conj(G3pp)), p3 = cmul(g1G2mpc, G3mm);
                    add8(iz23 + m, p0, p1, p2, p3);
                    add8(iz32 - m, p0, p1, p3, p2);
                    const cd q0 = cmul(g1G2mm, G3pm), q1 = cmul(std::conj(g1G2mpc), G3pp);
                    const cd q2 = cmul(std::conj(g1cG2pp), G3pm), q3 = cmul(g1G2mm, std::conj(G3mp));
                    add8(iz32 + m, q0, q1, q2, q3);
                    add8(iz23 - m, q0, q1, q3, q2);
                }
            }
        }
    }

    static void calc_split_sq(bool& split1, bool& split2, double s1, double s2, double bsq) {
        const double splitfactorsq = 0.3422;
        if (split1 && split2) return;
        if (s2 > s1) calc_split_sq(split2, split1, s2, s1, bsq);
        else if (s1 > 2. * s2) split1 = true;
        else { split1 = true; split2 = s2 * s2 > splitfactorsq * bsq; }
    }

    // splitC2Cells, src/Corr3.cpp:1323-1372
    void split_c2(const Node& c1, const Tree& TL, const std::vector<int>& in,
                  std::vector<int>& out) const {
        double s1 = c1.size;
        for (int i2 : in) {
            const Node& c2 = TL.nodes[i2];
            double s2 = c2.size;
            double rsq = dist_sq(c1.p, c2.p);
            double s1ps2 = s1 + s2;
            if (too_small(rsq, s1ps2) || too_large(rsq, s1ps2)) continue;
            int k = -1; double r = 0, logr = 0;
            bool split = !single_bin(rsq, s1ps2, k, r, logr);
            if (split && s2 > s1) { out.push_back(c2.left); out.push_back(c2.right); }
            else out.push_back(i2);
        }
    }

    // splitC2CellsOrCalculateGn, src/Corr3.cpp:1452-1546
    double split_or_gn(const Node& c1, const Tree& TL, const std::vector<NodeSums>& SL,
                       const std::vector<int>& in, std::vector<int>& out,
                       bool& anysplit1, Scratch& mp, double prev_max) const {
        double s1 = c1.size;
        double max_remaining_r = 0.;
        double thresh1 = (1. - P.binsize) * prev_max;   // oneBinLessThan, BinType.h:153-154
        for (int i2 : in) {
            const Node& c2 = TL.nodes[i2];
            double s2 = c2.size;
            double rsq = dist_sq(c1.p, c2.p);
            double s1ps2 = s1 + s2;
            if (too_small(rsq, s1ps2) || too_large(rsq, s1ps2)) continue;
            int k = -1;
            double r = std::sqrt(rsq);
            double logr = 0;
            if (single_bin(rsq, s1ps2, k, r, logr)) {
                if (rsq >= P.minsepsq && rsq < P.maxsepsq) {
                    if (k < 0) {
                        logr = std::log(r);
                        k = int((logr - P.logminsep) / P.binsize);
                    }
                    calculate_gn(c1, TL, SL, i2, r, logr, k, mp);
                }
                continue;
            }
            double rtot = r + s1ps2;
            max_remaining_r = std::max(rtot, max_remaining_r);
            bool split1 = false, split2 = false;
            double bsq_eff = rsq * std::min(P.bsq, P.asq);
            calc_split_sq(split1, split2, s1, s2, bsq_eff);
            if (split1 && r + s1ps2 > thresh1) anysplit1 = true;
            if (split2) { out.push_back(c2.left); out.push_back(c2.right); }
            else out.push_back(i2);
        }
        return max_remaining_r;
    }

    // multipoleFinish, src/Corr3.cpp:1548-1613
    void finish(int i1, const std::vector<int>& c2list, Scratch& mp, int mink_zeta, double maxr) const {
        O.stages++;
        const Node& c1 = T1.nodes[i1];
        bool anysplit1 = false;
        std::vector<int> newlist;
        maxr = split_or_gn(c1, T, S, c2list, newlist, anysplit1, mp, maxr);
        if (!newlist.empty()) {
            if (maxr < P.maxsep) {
                int k = maxr < P.minsep ? 0
                        : int((std::log(maxr) - P.logminsep) / P.binsize) + 1;
                if (k < mink_zeta) {
                    calculate_zeta(c1, mp, k, mink_zeta);
                    mink_zeta = k;
                }
            }
            if (anysplit1) {
                Scratch copy = mp;
                finish(c1.left, newlist, mp, mink_zeta, maxr);
                finish(c1.right, newlist, copy, mink_zeta, maxr);
            } else {
                finish(i1, newlist, mp, mink_zeta, maxr);
            }
        } else {
            calculate_zeta(c1, mp, 0, mink_zeta);
        }
    }

    // multipoleSplit1, src/Corr3.cpp:1374-1408
    void split1(int i1, const std::vector<int>& c2list, Scratch& mp) const {
        const Node& c1 = T1.nodes[i1];
        double s1 = c1.size;
        std::vector<int> newlist;
        split_c2(c1, T, c2list, newlist);
        double maxbsq_eff = P.maxsepsq * std::min(P.bsq, P.asq);
        if (s1 * s1 > maxbsq_eff) {
            split1(c1.left, newlist, mp);
            split1(c1.right, newlist, mp);
        } else {
            mp.clear();
            finish(i1, newlist, mp, P.nbins, 0.);
        }
    }

    // ---- three fields: c1 from T1, the first list from T, the second from T3 ----

    // Corr3::calculateZeta with mp2, mp3 (src/Corr3.cpp:1976-2085) and
    // MultipoleHelper<5>::CalculateZeta(wg1, ordered, mp2, mp3, ...) (:2851-3003).
    // ordered 4: vertex 2 from the first list, 3 from the second; ordered 1: both ways.
    void calculate_zeta3(const Node& c1, int ordered, Scratch& mp2, Scratch& mp3, int kstart,
                         int mink_zeta) const {
        O.zeta_calls++;
        O.zeta_rows += std::max(0, P.nbins - kstart);
        const int nb = P.nbins, maxn = P.maxn, nn = 2 * maxn + 1;
        const double w1 = c1.w, n1 = double(c1.n);
        const bool swap23 = ordered == 1;
        for (int k2 = kstart; k2 < nb; ++k2) {
            const int k3end = k2 < mink_zeta ? nb : mink_zeta;
            for (int k3 = kstart; k3 < k3end; ++k3) {
                const int i = (k2 * nb + k3) * nn + maxn;
                if (!swap23) {
                    // mp2 holds the d3 sums (cells across from p2), mp3 the d2 sums
                    O.ntri[i] += n1 * mp3.npairs[k2] * mp2.npairs[k3];
                    const double ww12 = w1 * mp3.sumw[k2], ww13 = w1 * mp2.sumw[k3];
                    O.meand2[i] += ww13 * mp3.sumwr[k2];
                    O.meanlogd2[i] += ww13 * mp3.sumwlogr[k2];
                    O.meand3[i] += ww12 * mp2.sumwr[k3];
                    O.meanlogd3[i] += ww12 * mp2.sumwlogr[k3];
                } else {
                    O.ntri[i] += n1 * (mp2.npairs[k2] * mp3.npairs[k3] +
                                       mp3.npairs[k2] * mp2.npairs[k3]);
                    O.meand2[i] += w1 * (mp2.sumw[k3] * mp3.sumwr[k2] + mp3.sumw[k3] * mp2.sumwr[k2]);
                    O.meanlogd2[i] += w1 * (mp2.sumw[k3] * mp3.sumwlogr[k2] +
                                            mp3.sumw[k3] * mp2.sumwlogr[k2]);
                    O.meand3[i] += w1 * (mp2.sumw[k2] * mp3.sumwr[k3] + mp3.sumw[k2] * mp2.sumwr[k3]);
                    // the second term pairs sumw and sumwlogr of bin k3, as the reference does
                    O.meanlogd3[i] += w1 * (mp2.sumw[k2] * mp3.sumwlogr[k3] +
                                            mp3.sumw[k3] * mp2.sumwlogr[k3]);
                }
                const int iw2 = mp3.Windex(k2), iw3 = mp2.Windex(k3);
                for (int m = 0; m <= maxn; ++m) {
                    cd www = cmul(cmul(cd(w1, 0.), mp3.Wn[iw2 + m]), std::conj(mp2.Wn[iw3 + m]));
                    if (swap23)
                        www += cmul(cmul(cd(w1, 0.), mp2.Wn[iw2 + m]), std::conj(mp3.Wn[iw3 + m]));
                    O.weight[i + m] += www.real();
                    O.weight_im[i + m] += www.imag();
                    if (m > 0) {
                        O.weight[i - m] += www.real();
                        O.weight_im[i - m] -= www.imag();
                    }
                }
                if (P.D != GData) continue;
                // gam_mu(n) from G_a at bin k2 and G_b at bin k3 (the general form of
                // :2883-2897): a = mp3, b = mp2, and with swap23 also a = mp2, b = mp3
                const cd wg1(c1.gr, c1.gi), wg1c = std::conj(wg1);
                auto gam = [&](const Scratch& A, const Scratch& B, int n, cd* g) {
                    const int ia = A.Gindex(k2), ib = B.Gindex(k3);
                    g[0] += cmul(cmul(wg1, A.Gn[ia + n - 1]), B.Gn[ib - n - 1]);
                    g[1] += cmul(cmul(wg1c, A.Gn[ia + n + 1]), B.Gn[ib - n + 1]);
                    g[2] += cmul(cmul(wg1, A.Gn[ia + n - 1]), std::conj(B.Gn[ib + n + 1]));
                    g[3] += cmul(cmul(wg1, std::conj(A.Gn[ia - n + 1])), B.Gn[ib - n - 1]);
                };
                for (int n = -maxn; n <= maxn; ++n) {
                    cd g[4] = {cd(0), cd(0), cd(0), cd(0)};
                    gam(mp3, mp2, n, g);
                    if (swap23) gam(mp2, mp3, n, g);
                    add8(i + n, g[0], g[1], g[2], g[3]);
                }
            }
        }
    }

    // multipoleFinish with two lists, src/Corr3.cpp:1615-1679
    void finish3(int i1, const std::vector<int>& l2, const std::vector<int>& l3, int ordered,
                 Scratch& mp2, Scratch& mp3, int mink_zeta, double maxr2, double maxr3) const {
        O.stages++;
        const Node& c1 = T1.nodes[i1];
        bool anysplit1 = false;
        std::vector<int> n2, n3;
        maxr2 = split_or_gn(c1, T, S, l2, n2, anysplit1, mp2, maxr2);
        maxr3 = split_or_gn(c1, *T3, *S3, l3, n3, anysplit1, mp3, maxr3);
        const double maxr = std::max(maxr2, maxr3);
        if (!n2.empty() || !n3.empty()) {
            if (maxr < P.maxsep) {
                int k = maxr < P.minsep ? 0
                        : int((std::log(maxr) - P.logminsep) / P.binsize) + 1;
                if (k < mink_zeta) {
                    calculate_zeta3(c1, ordered, mp2, mp3, k, mink_zeta);
                    mink_zeta = k;
                }
            }
            if (anysplit1) {
                Scratch c2 = mp2, c3 = mp3;
                finish3(c1.left, n2, n3, ordered, mp2, mp3, mink_zeta, maxr2, maxr3);
                finish3(c1.right, n2, n3, ordered, c2, c3, mink_zeta, maxr2, maxr3);
            } else {
                finish3(i1, n2, n3, ordered, mp2, mp3, mink_zeta, maxr2, maxr3);
            }
        } else {
            calculate_zeta3(c1, ordered, mp2, mp3, 0, mink_zeta);
        }
    }

    // multipoleSplit1 with two lists, src/Corr3.cpp:1410-1450
    void split13(int i1, const std::vector<int>& l2, const std::vector<int>& l3, int ordered,
                 Scratch& mp2, Scratch& mp3) const {
        const Node& c1 = T1.nodes[i1];
        const double s1 = c1.size;
        std::vector<int> n2, n3;
        split_c2(c1, T, l2, n2);
        split_c2(c1, *T3, l3, n3);
        const double maxbsq_eff = P.maxsepsq * std::min(P.bsq, P.asq);
        if (s1 * s1 > maxbsq_eff) {
            split13(c1.left, n2, n3, ordered, mp2, mp3);
            split13(c1.right, n2, n3, ordered, mp2, mp3);
        } else {
            mp2.clear();
            mp3.clear();
            finish3(i1, n2, n3, ordered, mp2, mp3, P.nbins, 0., 0.);
        }
    }
};

static void merge_outs(std::vector<std::unique_ptr<Out3>>& outs, Out3& total) {
    for (auto& o : outs) {
        for (int q = 0; q < 8; ++q) for (int j = 0; j < total.ntot; ++j) total.z[q][j] += o->z[q][j];
        for (int j = 0; j < total.ntot; ++j) {
            total.meand2[j] += o->meand2[j]; total.meanlogd2[j] += o->meanlogd2[j];
            total.meand3[j] += o->meand3[j]; total.meanlogd3[j] += o->meanlogd3[j];
            total.weight[j] += o->weight[j]; total.weight_im[j] += o->weight_im[j];
            total.ntri[j] += o->ntri[j];
        }
        total.gn_calls += o->gn_calls; total.zeta_calls += o->zeta_calls;
        total.zeta_rows += o->zeta_rows; total.stages += o->stages;
    }
}

static int thread_count(int nthreads) {
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    (void)nthreads;
    return nt;
}

// BaseCorr3::multipole(field1, field2) (src/Corr3.cpp:1124-1175 for one field,
// :1177-1236 for two): c1 runs over the top cells of t1, the c2 list starts as the top
// cells of t2.
template <int C>
static void run_multipole(const P3& P, const Tree& t1, const Tree& t2, int nthreads, Out3& total) {
    std::vector<NodeSums> S;
    node_sums(t2, S);
    const int nt = thread_count(nthreads);
    std::vector<std::unique_ptr<Out3>> outs;
    for (int i = 0; i < nt; ++i) outs.emplace_back(new Out3(total.ntot));
    const long n1 = long(t1.top.size());
#pragma omp parallel for schedule(dynamic) num_threads(nt)
    for (long i = 0; i < n1; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        MP<C> mp(P, t1, t2, S, *outs[tid]);
        Scratch sc(P.nbins, P.maxn, P.D);
        mp.split1(t1.top[i], t2.top, sc);
    }
    merge_outs(outs, total);
}

// BaseCorr3::multipole(field1, field2, field3, ordered) (src/Corr3.cpp:1238-1305)
template <int C>
static void run_multipole3(const P3& P, const Tree& t1, const Tree& t2, const Tree& t3,
                           int ordered, int nthreads, Out3& total) {
    std::vector<NodeSums> S2, S3;
    node_sums(t2, S2);
    node_sums(t3, S3);
    const int nt = thread_count(nthreads);
    std::vector<std::unique_ptr<Out3>> outs;
    for (int i = 0; i < nt; ++i) outs.emplace_back(new Out3(total.ntot));
    const long n1 = long(t1.top.size());
#pragma omp parallel for schedule(dynamic) num_threads(nt)
    for (long i = 0; i < n1; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        MP<C> mp(P, t1, t2, S2, *outs[tid]);
        mp.T3 = &t3;
        mp.S3 = &S3;
        Scratch sc2(P.nbins, P.maxn, P.D), sc3(P.nbins, P.maxn, P.D);
        mp.split13(t1.top[i], t2.top, t3.top, ordered, sc2, sc3);
    }
    merge_outs(outs, total);
}

}  // namespace tco

using namespace tco;

extern "C" {

static int multipole_common(void* f1, void* f2, void* f3, int ordered, int d, int coords,
                            int nbins, int maxn, double minsep, double maxsep, double binsize,
                            double b, double a, int nthreads, double* const* zs,
                            double* const* ps, long long* counters) {
    if (d != NData && d != GData) return 1;
    P3 P;
    P.nbins = nbins; P.maxn = maxn; P.D = d;
    P.minsep = minsep; P.maxsep = maxsep; P.binsize = binsize; P.b = b; P.a = a;
    P.logminsep = std::log(minsep); P.minsepsq = minsep * minsep; P.maxsepsq = maxsep * maxsep;
    P.bsq = b * b; P.asq = a * a;
    const int ntot = nbins * nbins * (2 * maxn + 1);
    Out3 O(ntot);
    const Tree& t1 = *static_cast<Tree*>(f1);
    const Tree& t2 = *static_cast<Tree*>(f2);
    if (f3) {
        if (ordered != 1 && ordered != 4) return 1;
        const Tree& t3 = *static_cast<Tree*>(f3);
        if (coords == Flat) run_multipole3<Flat>(P, t1, t2, t3, ordered, nthreads, O);
        else if (coords == ThreeD) run_multipole3<ThreeD>(P, t1, t2, t3, ordered, nthreads, O);
        else run_multipole3<Sphere>(P, t1, t2, t3, ordered, nthreads, O);
    } else {
        if (coords == Flat) run_multipole<Flat>(P, t1, t2, nthreads, O);
        else if (coords == ThreeD) run_multipole<ThreeD>(P, t1, t2, nthreads, O);
        else run_multipole<Sphere>(P, t1, t2, nthreads, O);
    }
    for (int q = 0; q < 8; ++q) if (zs[q]) for (int j = 0; j < ntot; ++j) zs[q][j] += O.z[q][j];
    std::vector<double>* vs[7] = {&O.meand2, &O.meanlogd2, &O.meand3, &O.meanlogd3, &O.weight,
                                  &O.weight_im, &O.ntri};
    for (int q = 0; q < 7; ++q) if (ps[q]) for (int j = 0; j < ntot; ++j) ps[q][j] += (*vs[q])[j];
    if (counters) {
        counters[0] += O.gn_calls; counters[1] += O.zeta_calls;
        counters[2] += O.zeta_rows; counters[3] += O.stages;
    }
    return 0;
}

// Multipole auto-correlation of one oracle tree.  D in {0 (NNN), 4 (GGG)}.
// Arrays (ntot = nbins*nbins*(2*maxn+1) each, NULL allowed): zeta0..7, meand2,
// meanlogd2, meand3, meanlogd3, weight, weight_im, ntri, and 4 unused slots
// (meand1, meanlogd1, meanu, meanv).  Values are ADDED.
// counters[4]: Gn calls, zeta calls, zeta rows, finish stages.
int tco_corr3_multipole(void* f, int d, int coords, int nbins, int maxn, double minsep,
                        double maxsep, double binsize, double b, double a, int quick,
                        int exact, int nthreads, double* z0, double* z1, double* z2, double* z3,
                        double* z4, double* z5, double* z6, double* z7, double* meand2,
                        double* meanlogd2, double* meand3, double* meanlogd3, double* weight,
                        double* weight_im, double* ntri, double* unused1, double* unused2,
                        long long* counters) {
    (void)quick; (void)exact; (void)unused1; (void)unused2;
    double* zs[8] = {z0, z1, z2, z3, z4, z5, z6, z7};
    double* ps[7] = {meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri};
    return multipole_common(f, f, nullptr, 0, d, coords, nbins, maxn, minsep, maxsep, binsize, b,
                            a, nthreads, zs, ps, counters);
}

// One multipole pass over several trees, as BaseCorr3::multipole is called by the
// cross-correlation drivers (src/Corr3.cpp:3731-3768, 3831-3870, 3938-3985):
// f3 NULL: c1 from f1, the c2 list from f2 (the one-list form, :1177-1236);
// otherwise c1 from f1, lists from f2 and f3, ordered 1 or 4 (:1238-1305).
// The same arrays as tco_corr3_multipole; values are ADDED.
int tco_corr3_multipole_x(void* f1, void* f2, void* f3, int ordered, int d, int coords,
                          int nbins, int maxn, double minsep, double maxsep, double binsize,
                          double b, double a, int nthreads, double* z0, double* z1, double* z2,
                          double* z3, double* z4, double* z5, double* z6, double* z7,
                          double* meand2, double* meanlogd2, double* meand3, double* meanlogd3,
                          double* weight, double* weight_im, double* ntri, long long* counters) {
    double* zs[8] = {z0, z1, z2, z3, z4, z5, z6, z7};
    double* ps[7] = {meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri};
    return multipole_common(f1, f2, f3, ordered, d, coords, nbins, maxn, minsep, maxsep, binsize,
                            b, a, nthreads, zs, ps, counters);
}

}  // extern "C"
