// Runs the library's device tree build (treecorr_amd/csrc/tc_field.hip) and the k-means
// driver with the alternate algorithm's inertia (tc_patch.hip) on the CPU, through the
// HIP emulation in hip_emu/.
//
//   emu_kmeans <input> <output>
// input:  int64 header [C, n, npatch, max_iter, max_top, alt], double header [tol],
//         then x, y, z, w (n doubles each; z ignored for Flat), then npatch*dim centres.
// output: npatch*dim refined centres, int64 iterations run.
// Exit status 0 on success, 3 on an engine error (message on stderr).
#include "tc_engine.h"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

namespace tc {
static std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(const std::string& m) { g_err = m; return 1; }
int hip_fail(hipError_t, const char* what) { g_err = std::string("emulated HIP failure in ") + what; return 2; }
hipStream_t stream() { return nullptr; }
hipStream_t stream2() { static int s2; return (hipStream_t)&s2; }
hipStream_t stream3() { static int s3; return (hipStream_t)&s3; }
bool timing_enabled() { return false; }
KernelStat& kernel_stat(int which) {
    static KernelStat s[KS_COUNT];
    return s[which];
}
void* dev_alloc(size_t bytes) {
    void* p = nullptr;
    return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? p : nullptr;
}
void dev_free(void* p, size_t) {
    if (p) (void)hipFree(p);
}
}  // namespace tc

int main(int argc, char** argv) {
    using namespace tc;
    if (argc != 3) return 2;
    std::ifstream in(argv[1], std::ios::binary);
    long long h[6];
    double tol;
    in.read((char*)h, sizeof h);
    in.read((char*)&tol, sizeof tol);
    const int C = (int)h[0], npatch = (int)h[2];
    const long long n = h[1];
    const int dim = C == Flat ? 2 : 3;
    std::vector<double> col[4];
    for (auto& c : col) {
        c.resize(n);
        in.read((char*)c.data(), 8 * n);
    }
    std::vector<double> cen((size_t)npatch * dim);
    in.read((char*)cen.data(), 8 * cen.size());
    if (!in) return 2;
    DeviceTree T;
    const double* z = C == Flat ? nullptr : col[2].data();
    if (build_tree(NData, C, col[0].data(), col[1].data(), z, nullptr, nullptr, col[3].data(),
                   nullptr, n, 0., std::numeric_limits<double>::infinity(), Mean, 0, 0,
                   (int)h[4], T) || settle_tree(T)) {
        std::fprintf(stderr, "emu_kmeans: tree build: %s\n", g_err.c_str());
        return 3;
    }
    // the objects as the tree holds them: unit vectors on the sphere, float weights
    std::vector<double> x = col[0], y = col[1], zz = col[2], w = col[3];
    for (long long i = 0; i < n; ++i) {
        if (C == Sphere) {
            const double r = std::sqrt(x[i] * x[i] + y[i] * y[i] + zz[i] * zz[i]);
            x[i] /= r; y[i] /= r; zz[i] /= r;
        }
        w[i] = (double)(float)w[i];
    }
    const double size = std::sqrt(T.info.sizesq);
    int niter = 0;
    const int rc = h[5] ? kmeans_run_tree(cen.data(), npatch, dim, x.data(), y.data(),
                                          dim == 3 ? zz.data() : nullptr, w.data(), n,
                                          (int)h[3], tol, size, C == Sphere ? 1 : 0, T.nodes,
                                          T.nnodes, T.dtop, (int)T.top.size(), &niter)
                        : kmeans_run(cen.data(), npatch, dim, x.data(), y.data(),
                                     dim == 3 ? zz.data() : nullptr, w.data(), n, (int)h[3],
                                     tol, size, C == Sphere ? 1 : 0, &niter);
    if (rc) {
        std::fprintf(stderr, "emu_kmeans: k-means: %s\n", g_err.c_str());
        return 3;
    }
    std::ofstream o(argv[2], std::ios::binary);
    o.write((const char*)cen.data(), 8 * cen.size());
    const long long it = niter;
    o.write((const char*)&it, 8);
    dev_free(T.nodes, T.nodes_bytes);
    dev_free(T.dtop, T.dtop_bytes);
    return o ? 0 : 2;
}
