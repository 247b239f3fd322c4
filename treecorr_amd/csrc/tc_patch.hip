// Patch assignment on the device: QuickAssign (src/KMeans.cpp:890-942), the
// nearest-centre labelling TreeCorr applies to a catalog given patch_centers
// (treecorr/catalog.py:1556-1566).
#include "tc_engine.h"

#include <algorithm>
#include <vector>

namespace tc {

constexpr int QA_THREADS = 256;
constexpr int QA_CHUNK = 1024;     // centres staged in LDS at a time

// One thread per object.  Every thread of a workgroup reads the same staged centre,
// so the LDS reads are broadcasts.  The squared distance is (dx^2 + dy^2) + dz^2 and
// the first centre wins a tie, as in the reference's loop (kmin starts at centre 0
// and moves only on a strictly smaller distance).
template <int DIM>
__global__ __launch_bounds__(QA_THREADS) void k_quick_assign(const double* __restrict__ centers,
                                                             int npatch,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             const double* __restrict__ z,
                                                             long long n, int* __restrict__ out) {
    __shared__ double sc[QA_CHUNK * DIM];
    const long long i = (long long)blockIdx.x * QA_THREADS + threadIdx.x;
    const bool live = i < n;
    const double px = live ? x[i] : 0., py = live ? y[i] : 0.;
    const double pz = (DIM == 3 && live) ? z[i] : 0.;
    int kmin = 0;
    double min_rsq = 0.;
    for (int c0 = 0; c0 < npatch; c0 += QA_CHUNK) {
        const int nc = npatch - c0 < QA_CHUNK ? npatch - c0 : QA_CHUNK;
        __syncthreads();
        for (int t = threadIdx.x; t < nc * DIM; t += QA_THREADS) sc[t] = centers[(size_t)c0 * DIM + t];
        __syncthreads();
        for (int k = 0; k < nc; ++k) {
            const double dx = px - sc[k * DIM], dy = py - sc[k * DIM + 1];
            double rsq = dx * dx + dy * dy;
            if (DIM == 3) {
                const double dz = pz - sc[k * DIM + 2];
                rsq = rsq + dz * dz;
            }
            if (c0 + k == 0) min_rsq = rsq;
            else if (rsq < min_rsq) { kmin = c0 + k; min_rsq = rsq; }
        }
    }
    if (live) out[i] = kmin;
}

int quick_assign(const double* centers, int npatch, const double* x, const double* y,
                 const double* z, long long n, int* patch) {
    hipStream_t st = stream();
    const int dim = z ? 3 : 2;
    const size_t cb = sizeof(double) * (size_t)npatch * dim, pb = sizeof(double) * (size_t)n;
    std::vector<std::pair<void*, size_t>> bufs;
    auto alloc = [&](size_t bytes) {
        void* p = dev_alloc(bytes);
        if (p) bufs.push_back({p, bytes});
        return p;
    };
    auto release = [&]() { for (auto& b : bufs) dev_free(b.first, b.second); };
    double* dc = static_cast<double*>(alloc(cb));
    double* dx = static_cast<double*>(alloc(pb));
    double* dy = static_cast<double*>(alloc(pb));
    double* dz = z ? static_cast<double*>(alloc(pb)) : nullptr;
    int* dp = static_cast<int*>(alloc(sizeof(int) * (size_t)n));
    if (!dc || !dx || !dy || (z && !dz) || !dp) {
        release();
        return fail("out of device memory assigning patches");
    }
    hipError_t e = hipMemcpyAsync(dc, centers, cb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, pb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, pb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && z) e = hipMemcpyAsync(dz, z, pb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const dim3 grid((unsigned)((n + QA_THREADS - 1) / QA_THREADS));
        if (z) hipLaunchKernelGGL(k_quick_assign<3>, grid, dim3(QA_THREADS), 0, st, dc, npatch, dx, dy, dz, n, dp);
        else hipLaunchKernelGGL(k_quick_assign<2>, grid, dim3(QA_THREADS), 0, st, dc, npatch, dx, dy, dz, n, dp);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(patch, dp, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    release();
    if (e != hipSuccess) return hip_fail(e, "assigning patches");
    return 0;
}

}  // namespace tc
