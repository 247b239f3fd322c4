// Runs the library's device tree build (treecorr_amd/csrc/tc_field.hip) and the 2-point
// traversal (tc_corr2.hip) together on the CPU, through the HIP emulation in hip_emu/,
// in each of the traversal's modes (chunked launches, the persistent launch with jobs
// by a shared counter or by static shares, and its fallback).
//
//   emu_lib2 <input> <output>
// input:  int64 header [D, C, n1, n2, nbins, mode, levels, min_jobs, mintop, chunk]
//         double header [minsep, maxsep, binsize, b, a, minsize, maxsize]
//         then x, y, z, g1, g2, w of catalog 1 (n1 doubles each), the same of catalog 2.
// output: 7*nbins f64 sums, nbins u64 npairs, 4 u64 counters, int64 fallbacks.
// Exit status 0 on success, 3 on an engine error (message on stderr).
#include "tc_engine.h"

#include <cmath>
#include <cstdio>
#include <fstream>
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
    long long h[10];
    double p[7];
    in.read((char*)h, sizeof h);
    in.read((char*)p, sizeof p);
    const int D = (int)h[0], C = (int)h[1];
    const long long n[2] = {h[2], h[3]};
    std::vector<double> col[2][6];
    for (int s = 0; s < 2; ++s)
        for (auto& c : col[s]) {
            c.resize(n[s]);
            in.read((char*)c.data(), 8 * n[s]);
        }
    if (!in) return 2;
    DeviceTree T[2];
    for (int s = 0; s < 2 && n[s] > 0; ++s) {
        const double* z = C == Flat ? nullptr : col[s][2].data();
        const double* g1 = D == GData ? col[s][3].data() : nullptr;
        const double* g2 = D == GData ? col[s][4].data() : nullptr;
        int rc = build_tree(D, C, col[s][0].data(), col[s][1].data(), z, g1, g2, col[s][5].data(),
                            nullptr, n[s], p[5], p[6], Mean, 0, (int)h[8], 10, T[s]);
        if (rc == 0) rc = settle_tree(T[s]);
        if (rc) {
            std::fprintf(stderr, "emu_lib2: tree build: %s\n", g_err.c_str());
            return 3;
        }
    }
    C2Params P{};
    P.nbins = (int)h[4]; P.quick = 0;
    P.minsep = p[0]; P.maxsep = p[1]; P.binsize = p[2]; P.b = p[3]; P.a = p[4];
    P.logminsep = std::log(P.minsep); P.halfminsep = 0.5 * P.minsep;
    P.minsepsq = P.minsep * P.minsep; P.maxsepsq = P.maxsep * P.maxsep;
    P.bsq = P.b * P.b; P.asq = P.a * P.a;
    P.inv_binsize = 1. / P.binsize;
    P.metric = MetEuclidean;
    P.minrpar = -1.7976931348623157e308; P.maxrpar = 1.7976931348623157e308;
    if (set_traversal_mode((int)h[5]) || set_persist_params((int)h[6], h[7]) ||
        set_traversal_chunk(h[9])) {
        std::fprintf(stderr, "emu_lib2: %s\n", g_err.c_str());
        return 3;
    }
    Corr2Scratch* S = new_corr2_scratch();
    const int nb = P.nbins;
    std::vector<double> out((size_t)7 * nb);
    std::vector<unsigned long long> np(nb), cnt(4);
    const bool cross = n[1] > 0;
    int rc = corr2_run(*S, D, C, T[0].nodes, T[0].dtop, (int)T[0].top.size(), T[0].depth,
                       cross ? T[1].nodes : nullptr, cross ? T[1].dtop : nullptr,
                       cross ? (int)T[1].top.size() : 0, cross ? T[1].depth : 0, P, out.data(),
                       np.data(), cnt.data());
    if (rc) {
        std::fprintf(stderr, "emu_lib2: traversal: %s\n", g_err.c_str());
        return 3;
    }
    const long long fb = persist_fallbacks();
    std::ofstream o(argv[2], std::ios::binary);
    o.write((const char*)out.data(), 8 * out.size());
    o.write((const char*)np.data(), 8 * np.size());
    o.write((const char*)cnt.data(), 8 * cnt.size());
    o.write((const char*)&fb, 8);
    delete_corr2_scratch(S);
    for (auto& t : T) {
        dev_free(t.nodes, t.nodes_bytes);
        dev_free(t.dtop, t.dtop_bytes);
    }
    return o ? 0 : 2;
}
