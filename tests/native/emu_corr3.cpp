// Runs the device multipole engine (treecorr_amd/csrc/tc_corr3.hip) on the CPU
// through the HIP emulation in hip_emu/, on a tree handed over from the oracle.
// Built with g++ -fsanitize=address so that any out-of-bounds access of the
// kernels' indexing is reported with its source line.
#include "../../treecorr_amd/csrc/tc_corr3.hip"

#include <functional>
#include <string>
#include <vector>

namespace tc {
static std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(const std::string& m) { g_err = m; return 1; }
int hip_fail(hipError_t, const char* what) { g_err = std::string("emulated HIP failure in ") + what; return 2; }
hipStream_t stream() { return nullptr; }
bool timing_enabled() { return false; }
// the matrix-core zeta (tc_zeta_mfma.hip) is GPU code: the emulation runs the FP64 one
ZetaMfmaFn pick_zeta_mfma(int, int, int, int& threads, size_t& bytes) { threads = 0; bytes = 0; return nullptr; }
KernelStat& kernel_stat(int which) {
    static KernelStat s[KS_COUNT];
    return s[which];
}
void* dev_alloc(size_t bytes) {
    void* p = nullptr;
    return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
}
void dev_free(void* p, size_t) {
    if (p) (void)hipFree(p);
}
}  // namespace tc

extern "C" {

// Oracle tree (pre-order arrays with left/right links) -> device layout (siblings
// adjacent), then corr3_run.  out: 15 * nbins*nbins*(2*maxn+1) doubles.
int emu_corr3(int D, int C, long long nn, const double* pos, const float* size, const float* w,
              const float* g, const long* n, const int* left, const int* right, const int* top,
              int ntop, int nbins, int maxn, double minsep, double maxsep, double binsize,
              double b, double a, int quick, double* out, unsigned long long* cnt, char* err,
              int errlen) {
    using namespace tc;
    std::vector<DNode> nodes;
    nodes.reserve((size_t)nn);
    std::vector<int> dtop_h(ntop);
    std::function<void(int, int)> fill = [&](int src, int dst) {
        DNode d;
        d.x = pos[3 * src]; d.y = pos[3 * src + 1]; d.z = pos[3 * src + 2];
        d.size = size[src]; d.w = w[src]; d.a = g[2 * src]; d.b = g[2 * src + 1];
        d.n = (int)n[src];
        d.child = -1;
        nodes[dst] = d;
        if (left[src] >= 0) {
            const int c = (int)nodes.size();
            nodes.resize(nodes.size() + 2);
            nodes[dst].child = c;
            fill(left[src], c);
            fill(right[src], c + 1);
        }
    };
    for (int t = 0; t < ntop; ++t) {
        const int at = (int)nodes.size();
        nodes.resize(nodes.size() + 1);
        dtop_h[t] = at;
        fill(top[t], at);
    }
    DNode* dn = nullptr;
    int* dt = nullptr;
    if (hipMalloc(&dn, sizeof(DNode) * nodes.size()) || hipMalloc(&dt, sizeof(int) * ntop)) return 2;
    std::memcpy(dn, nodes.data(), sizeof(DNode) * nodes.size());
    std::memcpy(dt, dtop_h.data(), sizeof(int) * ntop);
    C3Params P{};
    P.nbins = nbins; P.maxn = maxn; P.D = D; P.quick = quick;
    P.minsep = minsep; P.maxsep = maxsep; P.binsize = binsize; P.b = b; P.a = a;
    P.logminsep = std::log(minsep);
    P.minsepsq = minsep * minsep; P.maxsepsq = maxsep * maxsep;
    P.bsq = b * b; P.asq = a * a;
    P.maxbsq_eff = P.maxsepsq * std::min(P.bsq, P.asq);
    const int rc = corr3_run(D, C, dn, (int)nodes.size(), dt, ntop, P, out, cnt);
    if (rc && err && errlen > 0) std::snprintf(err, errlen, "%s", g_err.c_str());
    hipFree(dn);
    hipFree(dt);
    return rc;
}

}  // extern "C"

#ifdef EMU_CORR3_MAIN
// Command-line driver (an executable links AddressSanitizer without preloading):
//   emu_corr3 <input> <output>
// input: int64 header [D, C, nnodes, ntop, nbins, maxn, quick], double header
// [minsep, maxsep, binsize, b, a], then pos (3n f64), size, w (n f32), g (2n f32),
// n (n i64), left, right (n i32), top (ntop i32).  output: 15*ntot f64 sums, 4 u64
// counters.  Exit status 0 on success, 3 on an engine error (message on stderr).
#include <fstream>
int main(int argc, char** argv) {
    if (argc != 3) return 2;
    std::ifstream in(argv[1], std::ios::binary);
    long long h[7];
    double p[5];
    in.read((char*)h, sizeof h);
    in.read((char*)p, sizeof p);
    const long long nn = h[2], ntop = h[3];
    std::vector<double> pos(3 * nn);
    std::vector<float> size(nn), w(nn), g(2 * nn);
    std::vector<long> n(nn);
    std::vector<int> left(nn), right(nn), top(ntop);
    in.read((char*)pos.data(), 8 * pos.size());
    in.read((char*)size.data(), 4 * nn);
    in.read((char*)w.data(), 4 * nn);
    in.read((char*)g.data(), 8 * nn);
    in.read((char*)n.data(), 8 * nn);
    in.read((char*)left.data(), 4 * nn);
    in.read((char*)right.data(), 4 * nn);
    in.read((char*)top.data(), 4 * ntop);
    if (!in) return 2;
    const long long ntot = h[4] * h[4] * (2 * h[5] + 1);
    std::vector<double> out(15 * ntot, 0.);
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    char err[256] = {0};
    const int rc = emu_corr3((int)h[0], (int)h[1], nn, pos.data(), size.data(), w.data(), g.data(),
                             n.data(), left.data(), right.data(), top.data(), (int)ntop, (int)h[4],
                             (int)h[5], p[0], p[1], p[2], p[3], p[4], (int)h[6], out.data(), cnt,
                             err, sizeof err);
    if (rc) {
        std::fprintf(stderr, "emu_corr3: %s\n", err);
        return 3;
    }
    std::ofstream o(argv[2], std::ios::binary);
    o.write((const char*)out.data(), 8 * out.size());
    o.write((const char*)cnt, 4 * sizeof(unsigned long long));
    return o ? 0 : 2;
}
#endif
