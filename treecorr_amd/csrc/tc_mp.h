// treecorr_amd: definitions shared by the multipole kernels (tc_corr3.hip) and the
// zeta contraction on the matrix cores (tc_zeta_mfma.hip).
#pragma once
#include "tc_engine.h"

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

__device__ __forceinline__ void cmul3(double ar, double ai, double br, double bi, double& cr,
                                      double& ci) {
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

// calculateZeta(c1, mp, ks, mk) with mp in slot `slot` (c1 < 0: no job)
struct MPJob { int c1, slot, ks, mk; };
// rows [lo, hi) of a chunk of zeta jobs
struct RowRange { int lo, hi; };

// Error bits reported through `err` (all of them end the kernel cleanly).
enum { E_POOL = 1, E_INDEX = 2, E_SLOT = 4 };

// Scratch slots in chunks of 2^SLOT_SHIFT; the chunk table is a device array.
constexpr int SLOT_SHIFT = 12;
constexpr int SLOT_MASK = (1 << SLOT_SHIFT) - 1;
__device__ __forceinline__ double* slot_at(double* const* chunks, int s, int total) {
    return chunks[s >> SLOT_SHIFT] + (size_t)(s & SLOT_MASK) * (size_t)total;
}

// Compressed zeta outputs per (k2, k3 >= k2, m >= 0) cell, NV doubles: the weight (2),
// for GGG P0..P3 and Q0..Q3 (16), then ntri, meand2, meanlogd2, meand3, meanlogd3 (5).
__host__ __device__ constexpr int zeta_nv(int D) { return D == GData ? 23 : 7; }

// Blocks are dealt to the chip's 8 XCDs round robin, and every XCD has its own L2.  A
// grid of 8 * ceil(nq / 8) * nsub blocks is read as (chunk q, sub-range s) so that all
// sub-ranges of one chunk run on the same XCD and share the chunk's scratch slots in
// its L2.  Returns false for the padding blocks (q >= nq).
__device__ __forceinline__ bool xcd_chunk(int nq, int nsub, int& q, int& sub) {
    const int b = (int)blockIdx.x;
    const int r = b >> 3;
    q = (b & 7) + 8 * (r / nsub);
    sub = r - (r / nsub) * nsub;
    return q < nq;
}

// The zeta contraction of the cells of a level's jobs on the matrix cores: mode 1 FP32
// MFMA, mode 2 split-BF16 MFMA (tc_zeta_mfma.hip).  Returns the kernel, its threads per
// workgroup and LDS bytes, or nullptr for mode 0.
using ZetaMfmaFn = void (*)(const MPJob*, int, const DNode*, double* const*, C3Params, int,
                            double*, int);
ZetaMfmaFn pick_zeta_mfma(int D, int mode, int nbins, int& threads, size_t& lds_bytes);

}  // namespace tc
