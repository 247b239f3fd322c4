// Patch assignment on the device: QuickAssign (src/KMeans.cpp:890-942), the
// nearest-centre labelling TreeCorr applies to a catalog given patch_centers
// (treecorr/catalog.py:1556-1566).
#include "tc_engine.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
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

// SelectPatch (src/KMeans.cpp:944-1011): use[i] = 1 unless another centre is strictly
// nearer object i than centre `patch` (a tie keeps the object).  The squared distances
// are added as the compiled reference adds them: dx^2 + dy^2 in 2-d; in 3-d
// (dx^2 + dy^2) + dz^2 for the selected centre and dx^2 + (dy^2 + dz^2) for the others
// (its -ffast-math build; this decides objects equidistant from two equal centres).
template <int DIM>
__global__ __launch_bounds__(QA_THREADS) void k_select_patch(const double* __restrict__ centers,
                                                             int npatch, int patch,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             const double* __restrict__ z,
                                                             long long n, int* __restrict__ use) {
    __shared__ double sc[QA_CHUNK * DIM];
    const long long i = (long long)blockIdx.x * QA_THREADS + threadIdx.x;
    const bool live = i < n;
    const double px = live ? x[i] : 0., py = live ? y[i] : 0.;
    const double pz = (DIM == 3 && live) ? z[i] : 0.;
    double p_dsq;
    {
        const double dx = px - centers[(size_t)patch * DIM], dy = py - centers[(size_t)patch * DIM + 1];
        p_dsq = dx * dx + dy * dy;
        if (DIM == 3) {
            const double dz = pz - centers[(size_t)patch * DIM + 2];
            p_dsq = p_dsq + dz * dz;
        }
    }
    int keep = 1;
    for (int c0 = 0; c0 < npatch; c0 += QA_CHUNK) {
        const int nc = npatch - c0 < QA_CHUNK ? npatch - c0 : QA_CHUNK;
        __syncthreads();
        for (int t = threadIdx.x; t < nc * DIM; t += QA_THREADS) sc[t] = centers[(size_t)c0 * DIM + t];
        __syncthreads();
        for (int k = 0; k < nc; ++k) {
            if (c0 + k == patch) continue;
            const double dx = px - sc[k * DIM], dy = py - sc[k * DIM + 1];
            double q_dsq;
            if (DIM == 3) {
                const double dz = pz - sc[k * DIM + 2];
                q_dsq = dx * dx + (dy * dy + dz * dz);
            } else {
                q_dsq = dx * dx + dy * dy;
            }
            if (q_dsq < p_dsq) keep = 0;
        }
    }
    if (live) use[i] = keep;
}

static int assign_or_select(const double* centers, int npatch, int select, const double* x,
                            const double* y, const double* z, long long n, int* out);

int quick_assign(const double* centers, int npatch, const double* x, const double* y,
                 const double* z, long long n, int* patch) {
    return assign_or_select(centers, npatch, -1, x, y, z, n, patch);
}

int select_patch(int patch, const double* centers, int npatch, const double* x, const double* y,
                 const double* z, long long n, int* use) {
    if (patch < 0 || patch >= npatch) return fail("patch must be in [0, npatch)");
    return assign_or_select(centers, npatch, patch, x, y, z, n, use);
}

// QuickAssign (select < 0) or SelectPatch (select = the patch), host arrays in and out.
static int assign_or_select(const double* centers, int npatch, int select, const double* x,
                            const double* y, const double* z, long long n, int* patch) {
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
        if (select >= 0) {
            if (z) hipLaunchKernelGGL(k_select_patch<3>, grid, dim3(QA_THREADS), 0, st, dc, npatch, select, dx, dy, dz, n, dp);
            else hipLaunchKernelGGL(k_select_patch<2>, grid, dim3(QA_THREADS), 0, st, dc, npatch, select, dx, dy, dz, n, dp);
        } else if (z) {
            hipLaunchKernelGGL(k_quick_assign<3>, grid, dim3(QA_THREADS), 0, st, dc, npatch, dx, dy, dz, n, dp);
        } else {
            hipLaunchKernelGGL(k_quick_assign<2>, grid, dim3(QA_THREADS), 0, st, dc, npatch, dx, dy, dz, n, dp);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(patch, dp, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    release();
    if (e != hipSuccess) return hip_fail(e, "assigning patches");
    return 0;
}

// One Lloyd step of k-means over the objects (KMeansRun, src/KMeans.cpp:773-833, with
// the per-object assignment its FindCellsInPatches ends in): each object joins its
// nearest centre (the lowest index on ties) and adds w*x, w*y, w*z and w to that
// centre's sums.  Workgroups sum into LDS first when npatch <= KM_LDS_MAX, then add
// to sums.
constexpr int KM_LDS_MAX = 768;
constexpr int KM_NS = 4;

// pen (may be NULL): a per-centre penalty added to the squared distance, the inertia
// term of the alternate k-means (FindCellsInPatches with inertia, src/KMeans.cpp:563-585:
// a cell goes to the centre of least dsq + I_i; the candidate pruning there never drops
// the least one for any point of a cell, so every object goes to its own least
// |p - c_i|^2 + I_i, which is what this step computes).
template <int DIM>
__global__ __launch_bounds__(QA_THREADS) void k_kmeans_step(const double* __restrict__ centers,
                                                            int npatch,
                                                            const double* __restrict__ pen,
                                                            const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ z,
                                                            const double* __restrict__ w,
                                                            long long n, double* __restrict__ sums) {
    __shared__ double sc[QA_CHUNK * DIM];
    __shared__ double sp[QA_CHUNK];
    __shared__ double acc[KM_LDS_MAX * KM_NS];
    const bool lds = npatch <= KM_LDS_MAX;
    if (lds)
        for (int t = threadIdx.x; t < npatch * KM_NS; t += QA_THREADS) acc[t] = 0.;
    const long long i = (long long)blockIdx.x * QA_THREADS + threadIdx.x;
    const bool live = i < n;
    const double px = live ? x[i] : 0., py = live ? y[i] : 0.;
    const double pz = (DIM == 3 && live) ? z[i] : 0.;
    int kmin = 0;
    double best = 0.;
    for (int c0 = 0; c0 < npatch; c0 += QA_CHUNK) {
        const int nc = npatch - c0 < QA_CHUNK ? npatch - c0 : QA_CHUNK;
        __syncthreads();
        for (int t = threadIdx.x; t < nc * DIM; t += QA_THREADS) sc[t] = centers[(size_t)c0 * DIM + t];
        for (int t = threadIdx.x; t < nc; t += QA_THREADS) sp[t] = pen ? pen[c0 + t] : 0.;
        __syncthreads();
        for (int k = 0; k < nc; ++k) {
            const double dx = px - sc[k * DIM], dy = py - sc[k * DIM + 1];
            double dsq = dx * dx + dy * dy;
            if (DIM == 3) {
                const double dz = pz - sc[k * DIM + 2];
                dsq = dsq + dz * dz;
            }
            if (pen) dsq += sp[k];
            if (c0 + k == 0 || dsq < best) { kmin = c0 + k; best = dsq; }
        }
    }
    if (live) {
        const double wi = w[i];
        const double v[KM_NS] = {wi * px, wi * py, wi * pz, wi};
        double* dst = lds ? acc : sums;
#pragma unroll
        for (int q = 0; q < KM_NS; ++q) atomicAdd(&dst[(size_t)kmin * KM_NS + q], v[q]);
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < npatch * KM_NS; t += QA_THREADS)
            if (acc[t] != 0.) atomicAdd(&sums[t], acc[t]);
    }
}

// ---------------------------------------------------------------------------
// The alternate k-means' patch inertia (CalculateInertia, src/KMeans.cpp:388-447, run by
// FindCellsInPatches without inertia, :544-643): the tree is walked from the top cells
// with the list of candidate centres; at each cell the nearest candidate (of the cell's
// centroid) is found, the candidates farther than (min_d + 2 s)^2 are dropped, and the
// walk stops where one candidate is left or the cell has size 0; there the cell adds
// |pos - c|^2 w + 0.75 s^2 w (s > 0) to its centre's inertia.  Breadth first: one wave per
// (cell, candidate list) of a level; the kept list (in list order) is written once and
// shared by the two children.
// ---------------------------------------------------------------------------
struct KItem { int node, off, len, pad; };
constexpr int KI_WAVES = 4;

template <int DIM>
__global__ __launch_bounds__(64 * KI_WAVES) void k_kinertia_level(
    const DNode* __restrict__ nodes, int nnodes, const KItem* __restrict__ items, int nitems,
    const int* __restrict__ pool_in, const double* __restrict__ centers,
    KItem* __restrict__ items_out, int items_cap, int* __restrict__ nout,
    int* __restrict__ pool_out, int pool_cap, int* __restrict__ pool_top,
    double* __restrict__ inertia, double* __restrict__ sumw, int* __restrict__ err) {
    const int lane = threadIdx.x & 63;
    const int it = blockIdx.x * KI_WAVES + (threadIdx.x >> 6);
    if (it >= nitems) return;                       // a whole wave
    const KItem item = items[it];
    if ((unsigned)item.node >= (unsigned)nnodes || item.len <= 0) {
        if (lane == 0) atomicOr(err, 1);
        return;
    }
    const DNode c = nodes[item.node];
    const double s = c.size;
    const int* list = pool_in + item.off;
    auto dsq_of = [&](int k) {
        const double dx = c.x - centers[(size_t)k * DIM], dy = c.y - centers[(size_t)k * DIM + 1];
        double d = dx * dx + dy * dy;
        if (DIM == 3) {
            const double dz = c.z - centers[(size_t)k * DIM + 2];
            d = d + dz * dz;
        }
        return d;
    };
    // the nearest candidate (the smallest index among equal distances)
    double best = 0.;
    int bk = -1;
    for (int j = lane; j < item.len; j += 64) {
        const int k = list[j];
        const double d = dsq_of(k);
        if (bk < 0 || d < best || (d == best && k < bk)) { best = d; bk = k; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int ok = __shfl_xor(bk, o, 64);
        if (ok >= 0 && (bk < 0 || ob < best || (ob == best && ok < bk))) { best = ob; bk = ok; }
    }
    const double md = sqrt(best) + 2. * s;
    const double thresh = md * md;
    // candidates kept: dsq <= thresh (the nearest one always)
    int ncand = 0;
    for (int j0 = 0; j0 < item.len; j0 += 64) {
        const int j = j0 + lane;
        const bool keep = j < item.len && dsq_of(list[j]) <= thresh;
        ncand += __popcll(__ballot(keep));
    }
    if (ncand <= 1 || s == 0.) {
        if (lane == 0) {
            const double w = c.w;
            double add = best * w;
            if (s * s > 0.) add += 0.75 * (s * s) * w;
            atomicAdd(&inertia[bk], add);
            atomicAdd(sumw, w);
        }
        return;
    }
    if (c.child < 0 || c.child >= nnodes - 1) {
        if (lane == 0) atomicOr(err, 2);
        return;
    }
    int slot = 0, poff = 0;
    if (lane == 0) {
        slot = atomicAdd(nout, 2);
        poff = atomicAdd(pool_top, ncand);
    }
    slot = __shfl(slot, 0, 64);
    poff = __shfl(poff, 0, 64);
    if (slot + 2 > items_cap || poff + ncand > pool_cap) {
        if (lane == 0) atomicOr(err, 4);
        return;
    }
    int pos = 0;
    for (int j0 = 0; j0 < item.len; j0 += 64) {
        const int j = j0 + lane;
        const int k = j < item.len ? list[j] : -1;
        const bool keep = k >= 0 && dsq_of(k) <= thresh;
        const unsigned long long m = __ballot(keep);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (keep) pool_out[poff + pos + before] = k;
        pos += __popcll(m);
    }
    if (lane == 0) {
        items_out[slot] = KItem{c.child, poff, ncand, 0};
        items_out[slot + 1] = KItem{c.child + 1, poff, ncand, 0};
    }
}

// The patch inertia of the current centres dc (device, npatch x dim) on the tree, divided
// by sumw / npatch / 3 as CalculateInertia::finalize does (src/KMeans.cpp:431-447), into
// dpen (device).
static int kmeans_inertia(const DNode* nodes, int nnodes, const int* dtop, int ntop,
                          const double* dc, int npatch, int dim, double* dpen) {
    hipStream_t st = stream();
    struct DevBuf {
        void* p = nullptr; size_t b = 0;
        DevBuf() = default;
        DevBuf(const DevBuf&) = delete;
        DevBuf& operator=(const DevBuf&) = delete;
        ~DevBuf() { if (p) { (void)hipStreamSynchronize(stream()); dev_free(p, b); } }
        bool ensure(size_t bytes) {
            if (bytes <= b) return true;
            if (p) { (void)hipStreamSynchronize(stream()); dev_free(p, b); }
            b = std::max(bytes, b + b / 2);
            p = dev_alloc(b);
            if (!p) b = 0;
            return p != nullptr;
        }
    } itA, itB, plA, plB, acc, ctl;
    // level 0: every top cell with every centre
    std::vector<KItem> h0(ntop);
    std::vector<int> top(ntop), p0(npatch);
    TC_HIP(hipMemcpyAsync(top.data(), dtop, sizeof(int) * ntop, hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < ntop; ++i) h0[i] = KItem{top[i], 0, npatch, 0};
    for (int k = 0; k < npatch; ++k) p0[k] = k;
    if (!itA.ensure(sizeof(KItem) * ntop) || !plA.ensure(sizeof(int) * npatch) ||
        !acc.ensure(sizeof(double) * (npatch + 1)) || !ctl.ensure(sizeof(int) * 3))
        return fail("out of device memory for the k-means inertia");
    TC_HIP(hipMemcpyAsync(itA.p, h0.data(), sizeof(KItem) * ntop, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemcpyAsync(plA.p, p0.data(), sizeof(int) * npatch, hipMemcpyHostToDevice, st));
    TC_HIP(hipMemsetAsync(acc.p, 0, sizeof(double) * (npatch + 1), st));
    double* dI = static_cast<double*>(acc.p);
    long long nitems = ntop, sumlen = (long long)ntop * npatch;
    int* c = static_cast<int*>(ctl.p);      // [0] items out, [1] pool top, [2] error
    for (int level = 0; nitems > 0; ++level) {
        if (level > 200) return fail("k-means inertia: the tree is deeper than 200 levels");
        if (2 * nitems > 0x7fffffffll || sumlen > 0x7fffffffll)
            return fail("k-means inertia: candidate lists beyond 2^31 entries");
        if (!itB.ensure(sizeof(KItem) * 2 * nitems) || !plB.ensure(sizeof(int) * std::max<long long>(sumlen, 1)))
            return fail("out of device memory for the k-means inertia");
        TC_HIP(hipMemsetAsync(c, 0, sizeof(int) * 3, st));
        const unsigned grid = (unsigned)((nitems + KI_WAVES - 1) / KI_WAVES);
        if (dim == 3)
            hipLaunchKernelGGL(k_kinertia_level<3>, dim3(grid), dim3(64 * KI_WAVES), 0, st, nodes,
                               nnodes, (const KItem*)itA.p, (int)nitems, (const int*)plA.p, dc,
                               (KItem*)itB.p, (int)(2 * nitems), c, (int*)plB.p, (int)sumlen,
                               c + 1, dI, dI + npatch, c + 2);
        else
            hipLaunchKernelGGL(k_kinertia_level<2>, dim3(grid), dim3(64 * KI_WAVES), 0, st, nodes,
                               nnodes, (const KItem*)itA.p, (int)nitems, (const int*)plA.p, dc,
                               (KItem*)itB.p, (int)(2 * nitems), c, (int*)plB.p, (int)sumlen,
                               c + 1, dI, dI + npatch, c + 2);
        TC_HIP(hipGetLastError());
        int h[3];
        TC_HIP(hipMemcpyAsync(h, c, sizeof(h), hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (h[2]) return fail("k-means inertia: inconsistent tree or capacity (code " +
                              std::to_string(h[2]) + ")");
        nitems = h[0];
        sumlen = 2ll * h[1];
        std::swap(itA.p, itB.p); std::swap(itA.b, itB.b);
        std::swap(plA.p, plB.p); std::swap(plA.b, plB.b);
    }
    std::vector<double> hI(npatch + 1);
    TC_HIP(hipMemcpyAsync(hI.data(), dI, sizeof(double) * (npatch + 1), hipMemcpyDeviceToHost, st));
    TC_HIP(hipStreamSynchronize(st));
    const double meanw = hI[npatch] / npatch / 3.;
    for (int k = 0; k < npatch; ++k) hI[k] /= meanw;
    TC_HIP(hipMemcpyAsync(dpen, hI.data(), sizeof(double) * npatch, hipMemcpyHostToDevice, st));
    TC_HIP(hipStreamSynchronize(st));
    return 0;
}

int kmeans_run(double* centers, int npatch, int dim, const double* x, const double* y,
               const double* z, const double* w, long long n, int max_iter, double tol,
               double size, int sphere, int* niter) {
    return kmeans_run_tree(centers, npatch, dim, x, y, z, w, n, max_iter, tol, size, sphere,
                           nullptr, 0, nullptr, 0, niter);
}

int kmeans_run_tree(double* centers, int npatch, int dim, const double* x, const double* y,
                    const double* z, const double* w, long long n, int max_iter, double tol,
                    double size, int sphere, const DNode* tnodes, int tnn, const int* ttop,
                    int tntop, int* niter) {
    hipStream_t st = stream();
    const size_t pb = sizeof(double) * (size_t)n;
    const size_t cb = sizeof(double) * (size_t)npatch * dim;
    const size_t sb = sizeof(double) * (size_t)npatch * KM_NS;
    std::vector<std::pair<void*, size_t>> bufs;
    auto alloc = [&](size_t bytes) {
        void* p = dev_alloc(bytes);
        if (p) bufs.push_back({p, bytes});
        return p;
    };
    auto release = [&]() { for (auto& b : bufs) dev_free(b.first, b.second); };
    double* dx = static_cast<double*>(alloc(pb));
    double* dy = static_cast<double*>(alloc(pb));
    double* dz = dim == 3 ? static_cast<double*>(alloc(pb)) : nullptr;
    double* dw = static_cast<double*>(alloc(pb));
    double* dc = static_cast<double*>(alloc(cb));
    double* ds = static_cast<double*>(alloc(sb));
    double* dpen = tnodes ? static_cast<double*>(alloc(sizeof(double) * npatch)) : nullptr;
    if (!dx || !dy || (dim == 3 && !dz) || !dw || !dc || !ds || (tnodes && !dpen)) {
        release();
        return fail("out of device memory for k-means");
    }
    std::vector<double> hs((size_t)npatch * KM_NS), hc(centers, centers + (size_t)npatch * dim);
    std::vector<double> nc((size_t)npatch * dim);
    const dim3 grid((unsigned)((n + QA_THREADS - 1) / QA_THREADS));
    // one step: the sums of the assignment under the current centres
    int alt_rc = 0;
    auto step = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(dc, hc.data(), cb, hipMemcpyHostToDevice, st);
        // the alternate algorithm: the inertia of the current patches first
        // (KMeansRun1, src/KMeans.cpp:802-808)
        if (e == hipSuccess && tnodes) {
            alt_rc = kmeans_inertia(tnodes, tnn, ttop, tntop, dc, npatch, dim, dpen);
            if (alt_rc) return hipErrorInvalidValue;   // the error is alt_rc's
        }
        if (e == hipSuccess) e = hipMemsetAsync(ds, 0, sb, st);
        if (e == hipSuccess) {
            if (dim == 3) hipLaunchKernelGGL(k_kmeans_step<3>, grid, dim3(QA_THREADS), 0, st, dc, npatch, dpen, dx, dy, dz, dw, n, ds);
            else hipLaunchKernelGGL(k_kmeans_step<2>, grid, dim3(QA_THREADS), 0, st, dc, npatch, dpen, dx, dy, dz, dw, n, ds);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(hs.data(), ds, sb, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        return e;
    };
    hipError_t e = hipMemcpyAsync(dx, x, pb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, pb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && dim == 3) e = hipMemcpyAsync(dz, z, pb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dw, w, pb, hipMemcpyHostToDevice, st);
    // rms shift / size < tol  <=>  total shift^2 < tol^2 size^2 npatch
    const double tolsq = (tol * size) * (tol * size) * npatch;
    int it = 0;
    for (; e == hipSuccess && it < max_iter; ++it) {
        e = step();
        if (e != hipSuccess) break;
        // UpdateCenters::finalize (src/KMeans.cpp:362-372): an empty patch's centre
        // becomes the origin, as the reference's reset value
        double shiftsq = 0.;
        for (int k = 0; k < npatch; ++k) {
            const double* sk = &hs[(size_t)k * KM_NS];
            double c[3] = {0., 0., 0.};
            if (sk[3] > 0.) {
                for (int q = 0; q < dim; ++q) c[q] = sk[q] / sk[3];
                if (sphere) {
                    const double nrm = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                    if (nrm == 0.) { c[0] = 1.; c[1] = c[2] = 0.; }
                    else for (int q = 0; q < 3; ++q) c[q] /= nrm;
                }
            }
            double d = 0.;
            for (int q = 0; q < dim; ++q) {
                const double t = hc[(size_t)k * dim + q] - c[q];
                d += t * t;
                nc[(size_t)k * dim + q] = c[q];
            }
            shiftsq += d;
        }
        hc = nc;
        if (shiftsq < tolsq) { ++it; break; }
    }
    release();
    if (alt_rc) return alt_rc;
    if (e != hipSuccess) return hip_fail(e, "running k-means");
    std::copy(hc.begin(), hc.end(), centers);
    if (niter) *niter = it;
    return 0;
}

// ---------------------------------------------------------------------------
// k-means without allocation: the same Lloyd iterations on device-resident columns,
// for callers that cannot allocate (HIP graph capture, serving loops).  The per-centre
// sums and the centre update live in a fixed static device array, `scratch`.
// ---------------------------------------------------------------------------
constexpr int KM_NOALLOC_MAX = 8192;                       // centres at most
// per-centre sums (KM_NS each), then per-centre squared shifts, then the total shift
constexpr size_t KM_SCRATCH_DOUBLES = (size_t)KM_NOALLOC_MAX * (KM_NS + 1) + 8;
static __device__ __align__(256) double scratch[KM_SCRATCH_DOUBLES];

__global__ void k_kmeans_zero(double* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.;
}

// UpdateCenters::finalize (src/KMeans.cpp:362-372), as kmeans_run does it on the host:
// each centre moves to its objects' w-weighted mean (normalised for the sphere; an empty
// patch's centre becomes the origin), then one thread adds the squared shifts in centre
// order.  The sums are cleared for the next step.
__global__ void k_kmeans_update(double* __restrict__ centers, int npatch, int dim, int sphere,
                                double* __restrict__ sums, double* __restrict__ dsq,
                                double* __restrict__ shiftsq) {
    for (int k = threadIdx.x; k < npatch; k += blockDim.x) {
        double* sk = &sums[(size_t)k * KM_NS];
        double c[3] = {0., 0., 0.};
        if (sk[3] > 0.) {
            for (int q = 0; q < dim; ++q) c[q] = sk[q] / sk[3];
            if (sphere) {
                const double nrm = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                if (nrm == 0.) { c[0] = 1.; c[1] = c[2] = 0.; }
                else for (int q = 0; q < 3; ++q) c[q] /= nrm;
            }
        }
        double d = 0.;
        for (int q = 0; q < dim; ++q) {
            const double t = centers[(size_t)k * dim + q] - c[q];
            d += t * t;
            centers[(size_t)k * dim + q] = c[q];
        }
        dsq[k] = d;
        for (int q = 0; q < KM_NS; ++q) sk[q] = 0.;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.;
        for (int k = 0; k < npatch; ++k) s += dsq[k];
        *shiftsq = s;
    }
}

namespace {
std::mutex g_km_m;
double* g_km_scratch[64] = {nullptr};     // the array's address on each device
}  // namespace

int kmeans_run_noalloc(double* centers, int npatch, int dim, const double* x, const double* y,
                       const double* z, const double* w, long long n, int max_iter, double tol,
                       double size, int sphere, int* niter) {
    if (npatch > KM_NOALLOC_MAX)
        return fail("k-means without allocation: at most 8192 centres (npatch = " +
                    std::to_string(npatch) + ")");
    int dev = 0;
    TC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail("k-means without allocation: device index beyond 63");
    double* ptr = nullptr;
    {
        std::lock_guard<std::mutex> g(g_km_m);
        if (!g_km_scratch[dev]) {
            void* p = nullptr;
            TC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_km_scratch[dev] = static_cast<double*>(p);
        }
        ptr = g_km_scratch[dev];
    }
    double* sums = ptr;
    double* dsq = ptr + (size_t)KM_NOALLOC_MAX * KM_NS;
    double* shift = dsq + KM_NOALLOC_MAX;
    hipStream_t st = stream();
    const dim3 grid((unsigned)((n + QA_THREADS - 1) / QA_THREADS));
    const int nz = npatch * KM_NS;
    hipLaunchKernelGGL(k_kmeans_zero, dim3((nz + 255) / 256), dim3(256), 0, st, sums, nz);
    TC_HIP(hipGetLastError());
    const double tolsq = (tol * size) * (tol * size) * npatch;
    int it = 0;
    for (; it < max_iter; ++it) {
        if (dim == 3) hipLaunchKernelGGL(k_kmeans_step<3>, grid, dim3(QA_THREADS), 0, st, centers, npatch, nullptr, x, y, z, w, n, sums);
        else hipLaunchKernelGGL(k_kmeans_step<2>, grid, dim3(QA_THREADS), 0, st, centers, npatch, nullptr, x, y, z, w, n, sums);
        TC_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_kmeans_update, dim3(1), dim3(256), 0, st, centers, npatch, dim, sphere,
                           sums, dsq, shift);
        TC_HIP(hipGetLastError());
        double shiftsq = 0.;
        TC_HIP(hipMemcpyAsync(&shiftsq, shift, sizeof(double), hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        if (shiftsq < tolsq) { ++it; break; }
    }
    if (niter) *niter = it;
    return 0;
}

}  // namespace tc
