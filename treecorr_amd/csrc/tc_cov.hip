// treecorr_amd: the patch-based covariance estimates' linear algebra on the GPU.
//
// Every row of a jackknife, sample or bootstrap design matrix (treecorr/corr2base.py:
// 2321-2368) is a weighted sum of patch-pair sums: with the pair sums stacked as the rows
// of S (npair x nq) and each row's pair weights as the rows of W (nrow x npair), all rows
// are R = W S, one FP64 GEMM.  The covariance of the finished rows V (nrow x ncol) is
// C = sum_r w_r (v_r - mean)^T (v_r - mean), a second one (the reference's
// v.conj().T.dot(v) after centring, :2432-2475, 2494-2538).  Both are 64 x 64 output tiles
// per workgroup, each thread a 4 x 4 block, k in steps of 16 through LDS; every output
// element sums its k terms in order, so results do not depend on the launch.
#include "tc_engine.h"

#include <algorithm>
#include <vector>

namespace tc {

namespace {

constexpr int GT = 64, GK = 16;

// C (m x n) = A (m x k) B (k x n), all row major, FP64.  aT: A is given transposed
// (k x m) -- the covariance's (V - mean)^T; then each of A's k rows is scaled by wa[k]
// and centred by mean (length m) before use.
__global__ __launch_bounds__(256) void k_gemm_f64(int m, int n, int k, const double* __restrict__ A,
                                                  const double* __restrict__ B,
                                                  double* __restrict__ C, int aT,
                                                  const double* __restrict__ mean,
                                                  const double* __restrict__ wa) {
    __shared__ double sa[GK][GT + 1];
    __shared__ double sb[GK][GT + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int r0 = blockIdx.y * GT, c0 = blockIdx.x * GT;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < k; k0 += GK) {
        for (int t = threadIdx.x; t < GK * GT; t += 256) {
            const int kk = t / GT, ii = t % GT;
            const int r = r0 + ii, c = c0 + ii, kg = k0 + kk;
            double a = 0.;
            if (r < m && kg < k) {
                if (aT) {
                    a = A[(size_t)kg * m + r];
                    if (mean) a -= mean[r];
                    if (wa) a *= wa[kg];
                } else {
                    a = A[(size_t)r * k + kg];
                }
            }
            sa[kk][ii] = a;
            double b = 0.;
            if (c < n && kg < k) {
                b = B[(size_t)kg * n + c];
                if (aT && mean) b -= mean[c];    // the Gram product centres both sides
            }
            sb[kk][ii] = b;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = sa[kk][ty * 4 + q];
                bv[q] = sb[kk][tx * 4 + q];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int p = 0; p < 4; ++p) acc[q][p] = fma(av[q], bv[p], acc[q][p]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int r = r0 + ty * 4 + q, c = c0 + tx * 4 + p;
            if (r < m && c < n) C[(size_t)r * n + c] = acc[q][p];
        }
}

// mean over the rows of V (nrow x ncol), one thread per column, rows in order
__global__ void k_col_mean(int nrow, int ncol, const double* __restrict__ V,
                           double* __restrict__ mean) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    double s = 0.;
    for (int r = 0; r < nrow; ++r) s += V[(size_t)r * ncol + c];
    mean[c] = s / nrow;
}

struct Dev {
    void* p = nullptr;
    size_t b = 0;
    explicit Dev(size_t bytes) : b(bytes ? bytes : 8) { p = dev_alloc(b); }
    ~Dev() { dev_free(p, b); }
    double* d() const { return static_cast<double*>(p); }
};

int gemm(int m, int n, int k, const double* dA, const double* dB, double* dC, bool aT,
         const double* dmean, const double* dwa) {
    hipStream_t st = stream();
    dim3 grid((n + GT - 1) / GT, (m + GT - 1) / GT);
    hipLaunchKernelGGL(k_gemm_f64, grid, dim3(256), 0, st, m, n, k, dA, dB, dC, aT ? 1 : 0,
                       dmean, dwa);
    TC_HIP(hipGetLastError());
    return 0;
}

}  // namespace

int cov_rowsums(int nrow, int npair, int nq, const double* W, const double* S, double* R) {
    hipStream_t st = stream();
    if (nrow <= 0 || nq <= 0) return 0;
    if (npair <= 0) {
        std::fill(R, R + (size_t)nrow * nq, 0.);
        return 0;
    }
    Dev dW(sizeof(double) * (size_t)nrow * npair), dS(sizeof(double) * (size_t)npair * nq),
        dR(sizeof(double) * (size_t)nrow * nq);
    if (!dW.p || !dS.p || !dR.p) return fail("out of device memory for the design matrix");
    TC_HIP(hipMemcpyAsync(dW.p, W, sizeof(double) * (size_t)nrow * npair, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(dS.p, S, sizeof(double) * (size_t)npair * nq, hipMemcpyHostToDevice, st));
    if (int rc = gemm(nrow, nq, npair, dW.d(), dS.d(), dR.d(), false, nullptr, nullptr)) return rc;
    TC_HIP(hipMemcpyAsync(R, dR.p, sizeof(double) * (size_t)nrow * nq, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    return 0;
}

int cov_gram(int nrow, int ncol, const double* V, const double* w, double* C) {
    hipStream_t st = stream();
    if (nrow <= 0 || ncol <= 0) return fail("the design matrix is empty");
    Dev dV(sizeof(double) * (size_t)nrow * ncol), dM(sizeof(double) * ncol),
        dW(sizeof(double) * nrow), dC(sizeof(double) * (size_t)ncol * ncol);
    if (!dV.p || !dM.p || !dW.p || !dC.p) return fail("out of device memory for the covariance");
    TC_HIP(hipMemcpyAsync(dV.p, V, sizeof(double) * (size_t)nrow * ncol, hipMemcpyHostToDevice, st));
    if (w) TC_HIP(hipMemcpyAsync(dW.p, w, sizeof(double) * nrow, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_col_mean, dim3((ncol + 255) / 256), dim3(256), 0, st, nrow, ncol, dV.d(),
                       dM.d());
    TC_HIP(hipGetLastError());
    // C = (V - mean)^T diag(w) (V - mean): A = V read transposed (ncol x nrow)
    if (int rc = gemm(ncol, ncol, nrow, dV.d(), dV.d(), dC.d(), true, dM.d(), w ? dW.d() : nullptr))
        return rc;
    TC_HIP(hipMemcpyAsync(C, dC.p, sizeof(double) * (size_t)ncol * ncol, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // namespace tc
