// Runs the library's own device tree build (treecorr_amd/csrc/tc_field.hip) and the
// device multipole engine (tc_corr3.hip) together on the CPU, through the HIP
// emulation in hip_emu/: the node array, child links, counts and top list that the
// multipole kernels read are the ones the device builder wrote.
//
//   emu_lib3 <input> <output>
// input:  int64 header [D, C, n, nbins, maxn, quick, mintop, maxtop]
//         double header [minsep, maxsep, binsize, b, a, minsize, maxsize]
//         then x, y, z, g1, g2, w (n doubles each; z ignored for Flat).
// output: 15*ntot f64 sums, 4 u64 counters, int64 nnodes, int64 ntop, then the nodes
//         (nnodes DNode records of 48 bytes) and the top list (ntop int32).
// Exit status 0 on success, 3 on an engine error (message on stderr).
//
// Cross mode (first int64 of the input negative): int64 header [-1, D, C, nfields, nbins,
// maxn, mintop, maxtop, npass], the same double header, npass passes of 4 int64 (c1 field,
// list field, second list field or -1, ordered), then per field int64 n and its six
// columns.  Every field's tree is built by the device builder; each pass runs
// corr3_run_trees.  Output: the 15*ntot sums added over the passes, 4 u64 counters.
#include "tc_mp.h"

#include <cstdio>
#include <cstdlib>
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
// the matrix-core zeta (tc_zeta_mfma.hip) is GPU code: the emulation runs the FP64 one
ZetaMfmaFn pick_zeta_mfma(int, int, int, int& threads, size_t& bytes) { threads = 0; bytes = 0; return nullptr; }
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

static int cross_main(std::ifstream& in, const char* outpath) {
    using namespace tc;
    long long h[8];
    double p[7];
    in.read((char*)h, sizeof h);
    in.read((char*)p, sizeof p);
    const int D = (int)h[0], C = (int)h[1], nf = (int)h[2];
    std::vector<long long> passes(4 * h[7]);
    in.read((char*)passes.data(), 8 * passes.size());
    std::vector<DeviceTree> T(nf);
    for (int f = 0; f < nf; ++f) {
        long long n = 0;
        in.read((char*)&n, 8);
        std::vector<double> col[6];
        for (auto& c : col) {
            c.resize(n);
            in.read((char*)c.data(), 8 * n);
        }
        if (!in) return 2;
        const double* z = C == Flat ? nullptr : col[2].data();
        const double* g1 = D == GData ? col[3].data() : nullptr;
        const double* g2 = D == GData ? col[4].data() : nullptr;
        if (build_tree(D, C, col[0].data(), col[1].data(), z, g1, g2, col[5].data(), nullptr, n,
                       p[5], p[6], Mean, 0, (int)h[5], (int)h[6], T[f]) || settle_tree(T[f])) {
            std::fprintf(stderr, "emu_lib3: tree build: %s\n", g_err.c_str());
            return 3;
        }
    }
    C3Params P{};
    P.nbins = (int)h[3]; P.maxn = (int)h[4]; P.D = D; P.quick = 0;
    P.minsep = p[0]; P.maxsep = p[1]; P.binsize = p[2]; P.b = p[3]; P.a = p[4];
    P.logminsep = std::log(P.minsep);
    P.minsepsq = P.minsep * P.minsep; P.maxsepsq = P.maxsep * P.maxsep;
    P.bsq = P.b * P.b; P.asq = P.a * P.a;
    P.maxbsq_eff = P.maxsepsq * std::min(P.bsq, P.asq);
    const long long ntot = (long long)P.nbins * P.nbins * (2 * P.maxn + 1);
    std::vector<double> out(15 * ntot), acc(15 * ntot, 0.);
    unsigned long long tot[5] = {0, 0, 0, 0, 0};
    for (size_t q = 0; q < passes.size(); q += 4) {
        unsigned long long cnt[5] = {0, 0, 0, 0, 0};
        const long long a = passes[q], b = passes[q + 1], c = passes[q + 2];
        if (corr3_run_trees(D, C, T[a], T[b], c >= 0 ? &T[c] : nullptr, (int)passes[q + 3], P,
                            out.data(), cnt)) {
            std::fprintf(stderr, "emu_lib3: multipole pass: %s\n", g_err.c_str());
            return 3;
        }
        for (size_t j = 0; j < acc.size(); ++j) acc[j] += out[j];
        for (int k = 0; k < 5; ++k) tot[k] += cnt[k];
    }
    std::ofstream o(outpath, std::ios::binary);
    o.write((const char*)acc.data(), 8 * acc.size());
    o.write((const char*)tot, 4 * sizeof(unsigned long long));
    for (auto& t : T) {
        dev_free(t.nodes, t.nodes_bytes);
        dev_free(t.dtop, t.dtop_bytes);
    }
    return o ? 0 : 2;
}

int main(int argc, char** argv) {
    using namespace tc;
    if (argc != 3) return 2;
    // EMU_TREE_SPEC: the tree builder's speculation mode (tc_set_tree_speculation)
    if (const char* m = std::getenv("EMU_TREE_SPEC")) set_tree_speculation(std::atoi(m));
    std::ifstream in(argv[1], std::ios::binary);
    {
        long long mode = 0;
        in.read((char*)&mode, 8);
        if (mode < 0) return cross_main(in, argv[2]);
        in.seekg(0);
    }
    long long h[8];
    double p[7];
    in.read((char*)h, sizeof h);
    in.read((char*)p, sizeof p);
    const int D = (int)h[0], C = (int)h[1];
    const long long n = h[2];
    std::vector<double> col[6];
    for (auto& c : col) {
        c.resize(n);
        in.read((char*)c.data(), 8 * n);
    }
    if (!in) return 2;
    DeviceTree T;
    const double* z = C == Flat ? nullptr : col[2].data();
    const double* g1 = D == GData ? col[3].data() : nullptr;
    const double* g2 = D == GData ? col[4].data() : nullptr;
    int rc = build_tree(D, C, col[0].data(), col[1].data(), z, g1, g2, col[5].data(), nullptr, n,
                        p[5], p[6], Mean, 0, (int)h[6], (int)h[7], T);
    if (rc == 0) rc = settle_tree(T);
    if (rc) {
        std::fprintf(stderr, "emu_lib3: tree build: %s\n", g_err.c_str());
        return 3;
    }
    C3Params P{};
    P.nbins = (int)h[3]; P.maxn = (int)h[4]; P.D = D; P.quick = (int)h[5];
    P.minsep = p[0]; P.maxsep = p[1]; P.binsize = p[2]; P.b = p[3]; P.a = p[4];
    P.logminsep = std::log(P.minsep);
    P.minsepsq = P.minsep * P.minsep; P.maxsepsq = P.maxsep * P.maxsep;
    P.bsq = P.b * P.b; P.asq = P.a * P.a;
    P.maxbsq_eff = P.maxsepsq * std::min(P.bsq, P.asq);
    const long long ntot = (long long)P.nbins * P.nbins * (2 * P.maxn + 1);
    std::vector<double> out(15 * ntot, 0.);
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    rc = corr3_run(D, C, T.nodes, T.nnodes, T.dtop, (int)T.top.size(), P, out.data(), cnt);
    if (rc) {
        std::fprintf(stderr, "emu_lib3: multipole: %s\n", g_err.c_str());
        return 3;
    }
    std::ofstream o(argv[2], std::ios::binary);
    o.write((const char*)out.data(), 8 * out.size());
    o.write((const char*)cnt, 4 * sizeof(unsigned long long));
    const long long nn = T.nnodes, nt = (long long)T.top.size();
    o.write((const char*)&nn, 8);
    o.write((const char*)&nt, 8);
    o.write((const char*)T.nodes, sizeof(DNode) * nn);
    o.write((const char*)T.top.data(), 4 * nt);
    dev_free(T.nodes, T.nodes_bytes);
    dev_free(T.dtop, T.dtop_bytes);
    return o ? 0 : 2;
}
