// treecorr_amd: the multipole's zeta contraction on the matrix cores (gfx950 MFMA),
// FP32 or split BF16, for the cells of a level's zeta jobs (see tc_corr3.hip for the
// jobs and the compressed output layout).
#include "tc_mp.h"

namespace tc {

// ---------------------------------------------------------------------------
// Zeta on the matrix cores (off-diagonal cells k3 > k2), FP32 or split BF16
// ---------------------------------------------------------------------------
// For one multipole m and one output, the contraction over the jobs j of a chunk is a
// product of an (rows x jobs) and a (jobs x columns) complex matrix:
//   Z[k2, k3] = sum_j A_j[k2] B_j[k3],
// A_j[k2] zero outside the job's rows [ks, mk).  Per job and bin the operands are
//   A: a_w = w1 W(k2,m), a0 = g1 G(k2,m-1), a1 = conj(g1) G(k2,m+1),
//      a3 = g1 conj(G(k2,-m+1)), aq = g1 G(k2,-m-1)
//   B: b_w = conj(W(k3,m)), b_mm = G(k3,-m-1), b_mp = G(k3,-m+1), b_pp = G(k3,m+1),
//      b_pm = G(k3,m-1)
// and the outputs (MultipoleHelper<5>::CalculateZeta, src/Corr3.cpp:2582-2849) are
//   weight (a_w, b_w); P0 (a0, b_mm); P1 (a1, b_mp); P2 (a0, conj b_pp); P3 (a3, b_mm);
//   Q0 (aq, b_pm); Q1 (conj a3, b_pp); Q2 (conj a1, b_pm); Q3 (aq, conj b_mp).
// A complex product is four real ones: re = Ar Br - Ai Bi, im = Ar Bi + Ai Br.  One
// workgroup per (chunk, m); one wavefront per output; the operands of 32 jobs at a time
// are built in FP64 and stored in LDS as FP32; each block of 32 jobs is accumulated on
// the matrix cores in FP32 (one 16x16x4 FP32 MFMA per 4 jobs, or each FP32 value split
// into three BF16 parts hi + mid + lo and the six leading products on 16x16x32 BF16
// MFMAs), and the block sums are added in FP64.
typedef float zf4 __attribute__((ext_vector_type(4)));
typedef __bf16 zbf8 __attribute__((ext_vector_type(8)));
constexpr int ZM_KB = 32;             // jobs per block
constexpr int ZM_NA = 5, ZM_NB = 5;   // operand kinds per job and bin (GGG; NNN uses 1 + 1)

struct ZOp { int ia, ib; bool ca, cb; };
__device__ __forceinline__ ZOp zeta_op(int t) {
    switch (t) {
        case 0: return ZOp{0, 0, false, false};   // weight
        case 1: return ZOp{1, 1, false, false};   // P0
        case 2: return ZOp{2, 2, false, false};   // P1
        case 3: return ZOp{1, 3, false, true};    // P2
        case 4: return ZOp{3, 1, false, false};   // P3
        case 5: return ZOp{4, 4, false, false};   // Q0
        case 6: return ZOp{3, 3, true, false};    // Q1
        case 7: return ZOp{2, 4, true, false};    // Q2
        default: return ZOp{4, 2, false, true};   // Q3
    }
}

__device__ __forceinline__ void bf16_split3(float x, __bf16& h, __bf16& m, __bf16& l) {
    h = (__bf16)x;
    const float r1 = x - (float)h;
    m = (__bf16)r1;
    const float r2 = r1 - (float)m;
    l = (__bf16)r2;
}

template <int D, int MODE>
__global__ __launch_bounds__(D == GData ? 576 : 64) void k_mp_zeta_mfma(
    const MPJob* __restrict__ jobs, int njobs, const DNode* __restrict__ nodes,
    double* const* __restrict__ slots, C3Params P, int jpc, double* __restrict__ partial,
    int nq) {
    HIP_DYNAMIC_SHARED(float, tab)     // [ZM_KB][nb][NA + NB + NC] complex (re, im)
    __shared__ int s_kmin[9], s_kmax[9];
    constexpr int NV = zeta_nv(D);
    constexpr int NA = D == GData ? ZM_NA : 1, NB = D == GData ? ZM_NB : 1;
    constexpr int NC = D == GData ? 4 : 1, NT = NA + NB + NC;
    const int nb = P.nbins, M = P.maxn;
    int q, m;
    if (!xcd_chunk(nq, M + 1, q, m)) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    ScratchLayout Lo;
    Lo.init(nb, M, D);
    const int j0 = q * jpc;
    const int j1 = njobs < j0 + jpc ? njobs : j0 + jpc;
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
    if (lane == 0) { s_kmin[wv] = kmin; s_kmax[wv] = kmax; }
    __syncthreads();
    kmin = nb; kmax = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        kmin = s_kmin[w] < kmin ? s_kmin[w] : kmin;
        kmax = s_kmax[w] > kmax ? s_kmax[w] : kmax;
    }
    if (kmin >= kmax) return;                                   // uniform
    const int nrt = (kmax - kmin + 15) / 16, nct = (nb - kmin + 15) / 16;   // <= 2 each
    const int t = wv;                                           // this wave's output
    const bool active = !(t >= 5 && m == 0);                    // Q is unused at m = 0
    const ZOp op = zeta_op(t);
    // the diagonal's correction (the c2 = c3 terms): weight w1 sumww, P0 g1 sumwwgg0,
    // P1 conj(g1) sumwwgg1, P2 and P3 g1 sumwwgg2 (src/Corr3.cpp:1921-1925, 2595-2624)
    const int cidx = t == 0 ? 0 : t == 1 ? 1 : t == 2 ? 2 : t <= 4 ? 3 : -1;
    double acc[2][2][2][4];                                     // [rt][ct][re/im][r]
    double dcr[2][2], dci[2][2];                                // diagonal corrections
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            dcr[a][b] = 0.; dci[a][b] = 0.;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][b][c][r] = 0.;
        }
    for (int jb0 = j0; jb0 < j1; jb0 += ZM_KB) {
        __syncthreads();                 // the previous block's operands are no longer read
        for (int e = threadIdx.x; e < ZM_KB * nb; e += blockDim.x) {
            const int sj = e / nb, k = e - sj * nb;
            const int j = jb0 + sj;
            float* o = tab + (size_t)(sj * nb + k) * NT * 2;
            double vr[NT], vi[NT];
#pragma unroll
            for (int u = 0; u < NT; ++u) { vr[u] = 0.; vi[u] = 0.; }
            if (j < j1) {
                const MPJob jb = jobs[j];
                if (jb.c1 >= 0 && jb.ks < jb.mk) {
                    const double* sc = slot_at(slots, jb.slot, Lo.total);
                    const DNode c1 = nodes[jb.c1];
                    const bool row = k >= jb.ks && k < jb.mk;
                    const double wr = sc[Lo.W(k, m)], wi = sc[Lo.W(k, m) + 1];
                    if (row) {
                        vr[0] = c1.w * wr; vi[0] = c1.w * wi;
                        vr[NA + NB] = c1.w * sc[Lo.S(S_SUMWW, k)];
                    }
                    vr[NA] = wr; vi[NA] = -wi;                           // conj W(k3, m)
                    if (D == GData) {
                        const double g1r = c1.a, g1i = c1.b;
                        const double gmr = sc[Lo.G(k, m - 1)], gmi = sc[Lo.G(k, m - 1) + 1];
                        const double gpr = sc[Lo.G(k, m + 1)], gpi = sc[Lo.G(k, m + 1) + 1];
                        const double hpr = sc[Lo.G(k, -m + 1)], hpi = sc[Lo.G(k, -m + 1) + 1];
                        const double hmr = sc[Lo.G(k, -m - 1)], hmi = sc[Lo.G(k, -m - 1) + 1];
                        if (row) {
                            cmul3(g1r, g1i, gmr, gmi, vr[1], vi[1]);          // a0
                            cmul3(g1r, -g1i, gpr, gpi, vr[2], vi[2]);         // a1
                            cmul3(g1r, g1i, hpr, -hpi, vr[3], vi[3]);         // a3
                            cmul3(g1r, g1i, hmr, hmi, vr[4], vi[4]);          // aq
                            cmul3(g1r, g1i, sc[Lo.GG(0, k)], sc[Lo.GG(0, k) + 1],
                                  vr[NA + NB + 1], vi[NA + NB + 1]);
                            cmul3(g1r, -g1i, sc[Lo.GG(1, k)], sc[Lo.GG(1, k) + 1],
                                  vr[NA + NB + 2], vi[NA + NB + 2]);
                            cmul3(g1r, g1i, sc[Lo.GG(2, k)], sc[Lo.GG(2, k) + 1],
                                  vr[NA + NB + 3], vi[NA + NB + 3]);
                        }
                        vr[NA + 1] = hmr; vi[NA + 1] = hmi;                   // b_mm
                        vr[NA + 2] = hpr; vi[NA + 2] = hpi;                   // b_mp
                        vr[NA + 3] = gpr; vi[NA + 3] = gpi;                   // b_pp
                        vr[NA + 4] = gmr; vi[NA + 4] = gmi;                   // b_pm
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < NT; ++u) { o[2 * u] = (float)vr[u]; o[2 * u + 1] = (float)vi[u]; }
        }
        __syncthreads();
        if (!active) continue;
        const int nj = j1 - jb0 < ZM_KB ? j1 - jb0 : ZM_KB;
        const int ri = lane & 15, kg = lane >> 4;
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                if (rt >= nrt || ct >= nct) continue;              // uniform
                const int k2 = kmin + 16 * rt + ri, k3 = kmin + 16 * ct + ri;
                zf4 cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
                if (MODE == 1) {
                    for (int s0 = 0; s0 < nj; s0 += 4) {
                        const int sj = s0 + kg;
                        float ar = 0.f, ai = 0.f, br = 0.f, bi = 0.f;
                        if (sj < nj) {
                            if (k2 < kmax) {
                                const float* a = tab + (size_t)(sj * nb + k2) * NT * 2 + 2 * op.ia;
                                ar = a[0]; ai = op.ca ? -a[1] : a[1];
                            }
                            if (k3 < nb) {
                                const float* b = tab + (size_t)(sj * nb + k3) * NT * 2 + 2 * (NA + op.ib);
                                br = b[0]; bi = op.cb ? -b[1] : b[1];
                            }
                        }
                        cre = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, cre, 0, 0, 0);
                        cre = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, cre, 0, 0, 0);
                        cim = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, cim, 0, 0, 0);
                        cim = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, cim, 0, 0, 0);
                    }
                } else {
                    zbf8 arh, arm, arl, aih, aim, ail, nih, nim, nil;
                    zbf8 brh, brm, brl, bih, bim, bil;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int sj = 8 * kg + e;
                        float ar = 0.f, ai = 0.f, br = 0.f, bi = 0.f;
                        if (sj < nj) {
                            if (k2 < kmax) {
                                const float* a = tab + (size_t)(sj * nb + k2) * NT * 2 + 2 * op.ia;
                                ar = a[0]; ai = op.ca ? -a[1] : a[1];
                            }
                            if (k3 < nb) {
                                const float* b = tab + (size_t)(sj * nb + k3) * NT * 2 + 2 * (NA + op.ib);
                                br = b[0]; bi = op.cb ? -b[1] : b[1];
                            }
                        }
                        __bf16 h, mi, lo;
                        bf16_split3(ar, h, mi, lo); arh[e] = h; arm[e] = mi; arl[e] = lo;
                        bf16_split3(ai, h, mi, lo); aih[e] = h; aim[e] = mi; ail[e] = lo;
                        bf16_split3(-ai, h, mi, lo); nih[e] = h; nim[e] = mi; nil[e] = lo;
                        bf16_split3(br, h, mi, lo); brh[e] = h; brm[e] = mi; brl[e] = lo;
                        bf16_split3(bi, h, mi, lo); bih[e] = h; bim[e] = mi; bil[e] = lo;
                    }
                    // x y ~ hh + hm + mh + hl + mm + lh (the terms down to 2^-24 relative)
#define TC_MF(C, X, Y) C = __builtin_amdgcn_mfma_f32_16x16x32_bf16(X, Y, C, 0, 0, 0)
                    TC_MF(cre, arl, brh); TC_MF(cre, arm, brm); TC_MF(cre, arh, brl);
                    TC_MF(cre, arm, brh); TC_MF(cre, arh, brm); TC_MF(cre, arh, brh);
                    TC_MF(cre, nil, bih); TC_MF(cre, nim, bim); TC_MF(cre, nih, bil);
                    TC_MF(cre, nim, bih); TC_MF(cre, nih, bim); TC_MF(cre, nih, bih);
                    TC_MF(cim, arl, bih); TC_MF(cim, arm, bim); TC_MF(cim, arh, bil);
                    TC_MF(cim, arm, bih); TC_MF(cim, arh, bim); TC_MF(cim, arh, bih);
                    TC_MF(cim, ail, brh); TC_MF(cim, aim, brm); TC_MF(cim, aih, brl);
                    TC_MF(cim, aim, brh); TC_MF(cim, aih, brm); TC_MF(cim, aih, brh);
#undef TC_MF
                }
                // this block's FP32 sums, added in FP64
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    acc[rt][ct][0][r] += (double)cre[r];
                    acc[rt][ct][1][r] += (double)cim[r];
                }
                // the lane's diagonal cell of this tile, if any: its corrections
                if (cidx >= 0 && rt == ct) {
                    const int r0 = ri - 4 * kg;         // row offset of k3 within the lane's rows
                    if (r0 >= 0 && r0 < 4 && k3 < kmax) {
                        double sr = 0., si = 0.;
                        for (int sj = 0; sj < nj; ++sj) {
                            const float* c = tab + (size_t)(sj * nb + k3) * NT * 2 + 2 * (NA + NB + cidx);
                            sr += c[0];
                            si += c[1];
                        }
                        dcr[rt][ct] += sr;
                        dci[rt][ct] += si;
                    }
                }
            }
        }
    }
    // cells (k2, k3 >= k2, m): this wave's output (the diagonal with its correction); the
    // weight wave also clears the n = 0 sums of cells with m > 0 (the FP64 pass writes
    // those of m = 0)
    const size_t ncomp_chunk = (size_t)nb * nb * (M + 1) * NV;
    const int vre = t == 0 ? 0 : t <= 4 ? 2 + 2 * (t - 1) : 10 + 2 * (t - 5);
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (rt >= nrt || ct >= nct) continue;
                const int k2 = kmin + 16 * rt + 4 * (lane >> 4) + r;
                const int k3 = kmin + 16 * ct + (lane & 15);
                if (k2 >= kmax || k3 >= nb || k3 < k2) continue;
                double* dst = partial + (size_t)q * ncomp_chunk +
                              (((size_t)k2 * nb + k3) * (M + 1) + m) * NV;
                double vr = active ? acc[rt][ct][0][r] : 0.;
                double vi = active ? acc[rt][ct][1][r] : 0.;
                if (k3 == k2) {
                    if (cidx >= 0) { vr -= dcr[rt][ct]; vi -= dci[rt][ct]; }
                    else { vr = 0.; vi = 0.; }
                    if (t == 0) vi = 0.;
                }
                dst[vre] = vr;
                dst[vre + 1] = vi;
                if (t == 0 && m > 0)
                    for (int v = NV - 5; v < NV; ++v) dst[v] = 0.;
            }
}

ZetaMfmaFn pick_zeta_mfma(int D, int mode, int nbins, int& threads, size_t& lds_bytes) {
    threads = D == GData ? 576 : 64;
    lds_bytes = sizeof(float) * 2 * (size_t)ZM_KB * nbins * (D == GData ? ZM_NA + ZM_NB + 4 : 3);
    if (mode == 1) return D == GData ? &k_mp_zeta_mfma<GData, 1> : &k_mp_zeta_mfma<NData, 1>;
    if (mode == 2) return D == GData ? &k_mp_zeta_mfma<GData, 2> : &k_mp_zeta_mfma<NData, 2>;
    return nullptr;
}

}  // namespace tc
