// treecorr_amd: the extern "C" boundary (include/treecorr_amd.h).
//
// Handles, argument validation, error reporting, host<->device staging and the
// accumulate-into-caller-arrays semantics of the reference's Corr objects
// (include/Corr2.h:210-213).  All compute runs in the HIP kernels of
// tc_field.hip / tc_corr2.hip / tc_corr3.hip; there is no CPU fallback.
#include "tc_engine.h"
#include "../../include/treecorr_amd.h"

#include <cmath>
#include <cstring>
#include <limits>
#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace tc {


static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(const std::string& msg) { g_err = msg; return TC_ERR_INVALID; }
int hip_fail(hipError_t e, const char* what) {
    g_err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
    return TC_ERR_HIP;
}

static hipStream_t g_stream = nullptr;
static std::once_flag g_stream_once;
hipStream_t stream() {
    std::call_once(g_stream_once, [] {
        // Default priority, non-blocking with respect to the null stream.
        if (hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) g_stream = nullptr;
    });
    return g_stream;
}
static hipStream_t g_stream2 = nullptr, g_stream3 = nullptr;
static std::once_flag g_stream2_once, g_stream3_once;
hipStream_t stream2() {
    std::call_once(g_stream2_once, [] {
        if (hipStreamCreateWithFlags(&g_stream2, hipStreamNonBlocking) != hipSuccess) g_stream2 = nullptr;
    });
    return g_stream2;
}
hipStream_t stream3() {
    std::call_once(g_stream3_once, [] {
        if (hipStreamCreateWithFlags(&g_stream3, hipStreamNonBlocking) != hipSuccess) g_stream3 = nullptr;
    });
    return g_stream3;
}

// ---- caching allocator for temporaries ----
namespace {
struct FreeBlock { void* p; size_t bytes; };
std::mutex g_pool_m;
std::vector<FreeBlock> g_pool;
size_t g_pool_bytes = 0;
constexpr size_t kPoolMax = size_t(24) << 30;   // bytes kept for reuse
}  // namespace

void* dev_alloc(size_t bytes) {
    if (bytes == 0) bytes = 256;
    {
        std::lock_guard<std::mutex> g(g_pool_m);
        // smallest cached block that fits and is not more than twice the request
        int best = -1;
        for (int i = 0; i < (int)g_pool.size(); ++i)
            if (g_pool[i].bytes >= bytes && g_pool[i].bytes <= 2 * bytes &&
                (best < 0 || g_pool[i].bytes < g_pool[best].bytes))
                best = i;
        if (best >= 0) {
            FreeBlock b = g_pool[best];
            g_pool.erase(g_pool.begin() + best);
            g_pool_bytes -= b.bytes;
            return b.p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        // release the cache and try once more
        std::lock_guard<std::mutex> g(g_pool_m);
        for (auto& b : g_pool) (void)hipFree(b.p);
        g_pool.clear();
        g_pool_bytes = 0;
        (void)hipGetLastError();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();   // the caller reports it (or does without)
            return nullptr;
        }
    }
    return p;
}

void dev_free(void* p, size_t bytes) {
    if (!p) return;
    if (bytes == 0) bytes = 256;
    std::lock_guard<std::mutex> g(g_pool_m);
    if (g_pool_bytes + bytes > kPoolMax) { (void)hipFree(p); return; }
    g_pool.push_back(FreeBlock{p, bytes});
    g_pool_bytes += bytes;
}

static bool g_timing = false;
static KernelStat g_stats[KS_COUNT];
bool timing_enabled() { return g_timing; }
KernelStat& kernel_stat(int which) { return g_stats[which]; }

}  // namespace tc

using namespace tc;

// Node array shared by the trees of a multi-root build (freed with the last of them).
struct SharedNodes {
    DNode* p = nullptr;
    size_t bytes = 0;
    ~SharedNodes() { tc::dev_free(p, bytes); }
};

struct tc_field {
    int d = 0, coords = 0;
    long long nobj = 0;
    DeviceTree T;
    std::shared_ptr<SharedNodes> shared;   // set when T.nodes belongs to a multi-root build
    ~tc_field() {
        (void)tc::settle_tree(T);    // the exact pass of the top levels writes T.nodes
        if (!shared) tc::dev_free(T.nodes, T.nodes_bytes);
        tc::dev_free(T.dtop, T.dtop_bytes);
    }
};

struct tc_corr2 {
    int d = 0, d2 = 0;     // field types of the two sides
    int kind = 0;          // corr2_kind(d, d2): NData, KData, GData or XK_*
    C2Params P{};
    double* xi[4] = {nullptr, nullptr, nullptr, nullptr};
    double *meanr = nullptr, *meanlogr = nullptr, *weight = nullptr, *npairs = nullptr;
    long long counters[4] = {0, 0, 0, 0};
};

// One traversal scratch (work stack, output regions, accumulation slots) for the
// process, reused by every correlation; library calls that run GPU work hold
// g_run_m, so they are serialized.
static std::mutex g_run_m;
static Corr2Scratch* g_s2 = nullptr;
static tc_cost_model* g_cost_model_set = nullptr;   // the model tc_set_cost_model installed

namespace {

int check_gpu() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return tc::fail("no usable GPU: treecorr_amd runs only on a HIP device");
    if (!tc::stream()) return tc::fail("could not create the HIP stream");
    return 0;
}

int validate_field_args(int d, int coords, const double* x, const double* y, const double* z,
                        const double* d1, const double* d2, const double* w, long long n,
                        int sm, int mintop, int maxtop) {
    if (d != NData && d != KData && !is_complex(d))
        return tc::fail("data_type must be 0 (N), 1 (K), 2 (Z), 3 (V), 4 (G), 5 (T) or 6 (Q)");
    if (coords != Flat && coords != ThreeD && coords != Sphere)
        return tc::fail("coords must be 1 (Flat), 2 (ThreeD) or 3 (Sphere)");
    if (n <= 0) return tc::fail("a field needs at least one object");
    if (n > (long long)std::numeric_limits<int>::max() / 2)
        return tc::fail("too many objects for one field (limit 2^30)");
    if (!x || !y || !w) return tc::fail("x, y and w are required");
    if (coords != Flat && !z) return tc::fail("z is required for ThreeD and Sphere coordinates");
    if (d == KData && !d1) return tc::fail("k values are required for a K field");
    if (d == GData && (!d1 || !d2)) return tc::fail("g1 and g2 are required for a G field");
    if (is_complex(d) && (!d1 || !d2))
        return tc::fail("both components are required for a complex (Z, V, G, T, Q) field");
    if (sm == Random) {
        set_error("split_method='random' draws from the C library rand() and is not supported");
        return TC_ERR_UNSUPPORTED;
    }
    if (sm != Middle && sm != Median && sm != Mean) return tc::fail("invalid split_method");
    if (maxtop < 0 || maxtop > 60) return tc::fail("max_top must be in [0, 60]");
    (void)mintop;
    return 0;
}

int make_field(bool device, int d, int coords, const double* x, const double* y, const double* z,
               const double* d1, const double* d2, const double* w, const double* wpos,
               long long n, double minsize, double maxsize, int sm, int brute, int mintop,
               int maxtop, tc_field** out) {
    if (!out) return tc::fail("out is NULL");
    *out = nullptr;
    if (int rc = check_gpu()) return rc;
    if (int rc = validate_field_args(d, coords, x, y, z, d1, d2, w, n, sm, mintop, maxtop)) return rc;
    const double* src[7] = {x, y, z, d1, d2, w, wpos};
    double* dev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = 0;
    if (!device) {
        for (int k = 0; k < 7 && rc == 0; ++k) {
            if (!src[k]) continue;
            if (coords == Flat && k == 2) continue;
            dev[k] = static_cast<double*>(tc::dev_alloc(sizeof(double) * n));
            if (!dev[k]) { rc = tc::fail("out of device memory staging catalog columns"); break; }
            hipError_t e = hipMemcpy(dev[k], src[k], sizeof(double) * n, hipMemcpyHostToDevice);
            if (e != hipSuccess) rc = tc::hip_fail(e, "staging catalog columns");
        }
    } else {
        for (int k = 0; k < 7; ++k) dev[k] = const_cast<double*>(src[k]);
        if (coords == Flat) dev[2] = nullptr;
    }
    tc_field* f = nullptr;
    if (rc == 0) {
        f = new tc_field;
        f->d = d; f->coords = coords; f->nobj = n;
        std::lock_guard<std::mutex> run_lock(g_run_m);
        rc = tc::build_tree(d, coords, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dev[6], n,
                            minsize, maxsize, sm, brute, mintop, maxtop, f->T);
    }
    if (!device)
        for (int k = 0; k < 7; ++k) if (dev[k]) tc::dev_free(dev[k], sizeof(double) * n);
    if (rc) { delete f; return rc; }
    *out = f;
    return 0;
}

}  // namespace

extern "C" {

int tc_version(void) { return 100; }
const char* tc_last_error(void) { return tc::g_err.c_str(); }

int tc_device_count(int* count) {
    if (!count) return tc::fail("count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return 0;
}

int tc_field_create(int d, int coords, const double* x, const double* y, const double* z,
                    const double* d1, const double* d2, const double* w, const double* wpos,
                    long long nobj, double minsize, double maxsize, int sm, long long seed,
                    int brute, int mintop, int maxtop, tc_field** out) {
    (void)seed;
    return make_field(false, d, coords, x, y, z, d1, d2, w, wpos, nobj, minsize, maxsize, sm,
                      brute, mintop, maxtop, out);
}

int tc_field_create_device(int d, int coords, const double* x, const double* y, const double* z,
                           const double* d1, const double* d2, const double* w,
                           const double* wpos, long long nobj, double minsize, double maxsize,
                           int sm, long long seed, int brute, int mintop, int maxtop,
                           tc_field** out) {
    (void)seed;
    return make_field(true, d, coords, x, y, z, d1, d2, w, wpos, nobj, minsize, maxsize, sm,
                      brute, mintop, maxtop, out);
}

int tc_field_create_multi(int d, int coords, const double* x, const double* y, const double* z,
                          const double* d1, const double* d2, const double* w, const double* wpos,
                          long long nobj, const long long* bounds, int nroot, double minsize,
                          double maxsize, int sm, long long seed, int brute, int mintop,
                          int maxtop, tc_field** out) {
    (void)seed;
    if (!out || !bounds || nroot <= 0) return tc::fail("NULL or invalid argument");
    for (int r = 0; r < nroot; ++r) out[r] = nullptr;
    if (int rc = check_gpu()) return rc;
    if (int rc = validate_field_args(d, coords, x, y, z, d1, d2, w, nobj, sm, mintop, maxtop))
        return rc;
    std::vector<long long> bnd(bounds, bounds + nroot + 1);
    if (bnd.front() != 0 || bnd.back() != nobj) return tc::fail("bounds must run from 0 to nobj");
    for (int r = 0; r < nroot; ++r)
        if (bnd[r + 1] <= bnd[r]) return tc::fail("every patch must hold objects");
    const double* src[7] = {x, y, z, d1, d2, w, wpos};
    double* dev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = 0;
    for (int k = 0; k < 7 && rc == 0; ++k) {
        if (!src[k] || (coords == Flat && k == 2)) continue;
        dev[k] = static_cast<double*>(tc::dev_alloc(sizeof(double) * nobj));
        if (!dev[k]) { rc = tc::fail("out of device memory staging catalog columns"); break; }
        hipError_t e = hipMemcpy(dev[k], src[k], sizeof(double) * nobj, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = tc::hip_fail(e, "staging catalog columns");
    }
    MultiTree M;
    if (rc == 0) {
        std::lock_guard<std::mutex> run_lock(g_run_m);
        rc = tc::build_trees(d, coords, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dev[6],
                             nobj, bnd, minsize, maxsize, sm, brute, mintop, maxtop, M);
    }
    for (int k = 0; k < 7; ++k) if (dev[k]) tc::dev_free(dev[k], sizeof(double) * nobj);
    auto shared = std::make_shared<SharedNodes>();
    shared->p = M.nodes;
    shared->bytes = M.nodes_bytes;
    if (rc) return rc;
    for (int r = 0; r < nroot; ++r) out[r] = nullptr;
    for (int r = 0; r < nroot && rc == 0; ++r) {
        tc_field* f = new tc_field;
        f->d = d; f->coords = coords; f->nobj = bnd[r + 1] - bnd[r];
        f->shared = shared;
        f->T.nodes = M.nodes;
        f->T.nodes_bytes = M.nodes_bytes;
        f->T.nnodes = M.nnodes;
        f->T.depth = M.depth;
        f->T.top = std::move(M.top[r]);
        f->T.info = M.info[r];
        f->T.dtop_bytes = sizeof(int) * std::max<size_t>(1, f->T.top.size());
        f->T.dtop = static_cast<int*>(tc::dev_alloc(f->T.dtop_bytes));
        out[r] = f;
        if (!f->T.dtop) {
            rc = tc::fail("out of device memory for the top cells");
            break;
        }
        const hipError_t e = hipMemcpy(f->T.dtop, f->T.top.data(),
                                       sizeof(int) * f->T.top.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = tc::hip_fail(e, "copying top cells");
    }
    if (rc)
        for (int r = 0; r < nroot; ++r) { delete out[r]; out[r] = nullptr; }
    return rc;
}

// The field's centre and size and its records above the top cells are complete once
// the exact pass of the build's speculative levels has finished.
static int settled(const tc_field* f) {
    static std::mutex m;
    std::lock_guard<std::mutex> g(m);
    return tc::settle_tree(const_cast<tc_field*>(f)->T);
}

int tc_field_nobj(const tc_field* f, long long* nobj) {
    if (!f || !nobj) return tc::fail("NULL argument");
    *nobj = f->nobj;
    return 0;
}
int tc_field_size(const tc_field* f, double* size) {
    if (!f || !size) return tc::fail("NULL argument");
    if (int rc = settled(f)) return rc;
    *size = std::sqrt(f->T.info.sizesq);
    return 0;
}
int tc_field_ntop(const tc_field* f, long long* ntop) {
    if (!f || !ntop) return tc::fail("NULL argument");
    *ntop = (long long)f->T.top.size();
    return 0;
}
int tc_field_center(const tc_field* f, double* c) {
    if (!f || !c) return tc::fail("NULL argument");
    if (int rc = settled(f)) return rc;
    c[0] = f->T.info.cx; c[1] = f->T.info.cy; c[2] = f->T.info.cz;
    return 0;
}
int tc_field_nnodes(const tc_field* f, long long* nn) {
    if (!f || !nn) return tc::fail("NULL argument");
    *nn = f->T.nnodes;
    return 0;
}
int tc_field_copy_nodes(const tc_field* f, void* nodes_out, int* top_out) {
    if (!f || !nodes_out || !top_out) return tc::fail("NULL argument");
    if (int rc = settled(f)) return rc;
    TC_HIP(hipMemcpy(nodes_out, f->T.nodes, sizeof(DNode) * f->T.nnodes, hipMemcpyDeviceToHost));
    std::memcpy(top_out, f->T.top.data(), sizeof(int) * f->T.top.size());
    return 0;
}
int tc_field_destroy(tc_field* f) {
    delete f;
    return 0;
}

int tc_corr2_create(int d1, int d2, int bin_type, double minsep, double maxsep, int nbins,
                    double binsize, double b, double a, double minrpar, double maxrpar,
                    double xp, double yp, double zp, double* xi0, double* xi1, double* xi2,
                    double* xi3, double* meanr, double* meanlogr, double* weight,
                    double* npairs, tc_corr2** out) {
    if (!out) return tc::fail("out is NULL");
    *out = nullptr;
    const int kind = corr2_kind(d1, d2);
    if (kind < 0) {
        set_error("2-point correlations: NN, KK, ZZ, VV, GG, TT, QQ, and N or K with K, Z, V, "
                  "G, T or Q");
        return TC_ERR_UNSUPPORTED;
    }
    if (bin_type != 0) {
        set_error("only Log binning is implemented for 2-point correlations");
        return TC_ERR_UNSUPPORTED;
    }
    if (!(xp >= 0.) || !(yp >= 0.) || !(zp >= 0.)) {
        set_error("the periods must be >= 0");
        return TC_ERR_UNSUPPORTED;
    }
    if (nbins <= 0 || !(minsep > 0.) || !(maxsep > minsep) || !(binsize > 0.))
        return tc::fail("invalid binning");
    if (!weight || !meanr || !meanlogr || !npairs)
        return tc::fail("meanr, meanlogr, weight and npairs are required");
    const int nxi = corr2_nxi(kind);
    if ((nxi > 0 && !xi0) || (nxi > 1 && !xi1) || (nxi > 2 && (!xi2 || !xi3)))
        return tc::fail("the correlation needs its xi arrays");
    tc_corr2* c = new tc_corr2;
    c->d = d1;
    c->d2 = d2;
    c->kind = kind;
    C2Params& P = c->P;
    P.nbins = nbins; P.quick = 0;
    P.minsep = minsep; P.maxsep = maxsep; P.binsize = binsize; P.b = b; P.a = a;
    // BaseCorr2::BaseCorr2, src/Corr2.cpp:51-77
    P.logminsep = std::log(minsep); P.halfminsep = 0.5 * minsep;
    P.minsepsq = minsep * minsep; P.maxsepsq = maxsep * maxsep;
    P.bsq = b * b; P.asq = a * a;
    P.inv_binsize = 1. / binsize;
    P.metric = MetEuclidean; P.xp = xp; P.yp = yp; P.zp = zp;   // periods: Periodic metric
    // BaseCorr2::nontrivialRPar, include/Corr2.h:58-62
    P.minrpar = minrpar; P.maxrpar = maxrpar;
    P.rpar_on = (minrpar != -1.7976931348623157e308 || maxrpar != 1.7976931348623157e308);
    c->xi[0] = xi0; c->xi[1] = xi1; c->xi[2] = xi2; c->xi[3] = xi3;
    c->meanr = meanr; c->meanlogr = meanlogr; c->weight = weight; c->npairs = npairs;
    *out = c;
    return 0;
}

// The metric of a process call (src/Corr2.cpp:886-993 dispatch on Metric, enum at
// include/Metric.h:29): Euclidean (1); Arc (4), the great-circle angle, for spherical
// and 3-d coordinates (include/Metric.h:654-655); or Periodic (6) with the
// correlation's periods, which wrap every separation (flat and 3-d coordinates).
static int set_metric(const tc_corr2* c, int metric, int coords, C2Params& P) {
    P.metric = MetEuclidean;
    if (P.rpar_on && coords != ThreeD)
        return tc::fail("min_rpar/max_rpar need 3-d coordinates");
    if (P.rpar_on && (metric == MetArc || metric == MetPeriodic)) {
        set_error("min_rpar/max_rpar are implemented with the Euclidean, Rperp, OldRperp and "
                  "Rlens metrics");
        return TC_ERR_UNSUPPORTED;
    }
    if (metric == MetEuclidean) return 0;
    if (los_metric(metric)) {
        // Rperp, Rlens, OldRperp: 3-d only (include/Metric.h:271, :410, :570)
        if (coords != ThreeD) return tc::fail("this metric needs 3-d coordinates");
        P.metric = metric;
        return 0;
    }
    if (metric == MetArc) {
        if (coords == Flat) return tc::fail("the Arc metric needs spherical or 3-d coordinates");
        P.metric = MetArc;
        return 0;
    }
    if (metric == MetPeriodic) {
        if (coords == Sphere) return tc::fail("the Periodic metric needs flat or 3-d coordinates");
        if (!(c->P.xp > 0.) || !(c->P.yp > 0.) || (coords == ThreeD && !(c->P.zp > 0.)))
            return tc::fail("the Periodic metric needs positive periods");
        P.metric = MetPeriodic;
        return 0;
    }
    set_error("unknown metric");
    return TC_ERR_UNSUPPORTED;
}

// BaseCorr2::process(f1,f2) early exit, src/Corr2.cpp:215-231: the two fields are
// entirely closer than minsep or farther than maxsep.
// (the fields' centres and sizes: both fields settled)
static bool fields_out_of_range(const C2Params& P, int coords, const tc_field* f1,
                                const tc_field* f2) {
    const FieldInfo& a = f1->T.info;
    const FieldInfo& b = f2->T.info;
    double s1 = std::sqrt(a.sizesq), s2 = std::sqrt(b.sizesq);
    const double rsq = metric_dist_sq(P.metric, coords, a.cx, a.cy, a.cz, b.cx, b.cy, b.cz,
                                      P.xp, P.yp, P.zp, s1, s2);
    const double s1ps2 = s1 + s2;
    bool too_small = rsq < P.minsepsq && s1ps2 < P.minsep &&
                     rsq < (P.minsep - s1ps2) * (P.minsep - s1ps2);
    bool too_large = rsq >= P.maxsepsq && rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
    if (P.rpar_on || los_metric(P.metric)) {
        bool rin;
        too_small = metric_prune(P.metric, P.rpar_on, P.minrpar, P.maxrpar, a.cx, a.cy, a.cz,
                                 b.cx, b.cy, b.cz, rsq, s1ps2, too_small, too_large, P.minsep,
                                 P.minsepsq, P.maxsep, P.maxsepsq, rin);
        too_large = false;
    }
    return too_small || too_large;
}

static int corr2_common(tc_corr2* c, tc_field* f1, tc_field* f2, int quick, int metric) {
    if (!c || !f1) return tc::fail("NULL argument");
    if (!f2 && c->d != c->d2) return tc::fail("a mixed correlation needs two fields");
    if (f1->d != c->d || (f2 && (f2->d != c->d2 || f2->coords != f1->coords)))
        return tc::fail("field types do not match the correlation");
    C2Params P = c->P;
    if (int rc = set_metric(c, metric, f1->coords, P)) return rc;
    if (int rc = check_gpu()) return rc;
    P.quick = quick ? 1 : 0;
    const int nb = P.nbins;
    for (int k = 0; k < 4; ++k) c->counters[k] = 0;
    if (f2) {
        if (int rc = settled(f1)) return rc;
        if (int rc = settled(f2)) return rc;
        if (fields_out_of_range(P, f1->coords, f1, f2)) return 0;
    }
    std::vector<double> out((size_t)7 * nb);
    std::vector<unsigned long long> np(nb), cnt(4);
    std::lock_guard<std::mutex> run_lock(g_run_m);
    if (!g_s2) g_s2 = new_corr2_scratch();
    int rc = tc::corr2_run(*g_s2, c->kind, f1->coords, f1->T.nodes, f1->T.dtop, (int)f1->T.top.size(),
                           f1->T.depth, f2 ? f2->T.nodes : nullptr, f2 ? f2->T.dtop : nullptr,
                           f2 ? (int)f2->T.top.size() : 0, f2 ? f2->T.depth : 0, P, out.data(),
                           np.data(), cnt.data());
    if (rc) return rc;
    const int nxi = corr2_nxi(c->kind);
    for (int q = 0; q < nxi; ++q)
        for (int k = 0; k < nb; ++k) c->xi[q][k] += out[(size_t)q * nb + k];
    for (int k = 0; k < nb; ++k) {
        c->meanr[k] += out[(size_t)4 * nb + k];
        c->meanlogr[k] += out[(size_t)5 * nb + k];
        c->weight[k] += out[(size_t)6 * nb + k];
        c->npairs[k] += double(np[k]);
    }
    for (int k = 0; k < 4; ++k) c->counters[k] = (long long)cnt[k];
    return 0;
}

int tc_corr2_process_auto(tc_corr2* c, tc_field* f, int dots, int quick, int metric) {
    (void)dots;
    return corr2_common(c, f, nullptr, quick, metric);
}

int tc_corr2_process_cross(tc_corr2* c, tc_field* f1, tc_field* f2, int dots, int quick,
                           int metric) {
    (void)dots;
    if (!f2) return tc::fail("NULL argument");
    return corr2_common(c, f1, f2, quick, metric);
}

int tc_corr2_process_pairs(tc_corr2* c, tc_field* const* fields1, int nfield1,
                           tc_field* const* fields2, int nfield2, const int* p1, const int* p2,
                           int npair, int quick, int metric, double* pair_out) {
    if (!c || !fields1 || (npair > 0 && (!p1 || !p2 || !pair_out)) || npair < 0 || nfield1 <= 0)
        return tc::fail("NULL or invalid argument");
    const bool same = fields2 == nullptr;
    if (!same && nfield2 <= 0) return tc::fail("nfield2 must be positive");
    if (same && c->d != c->d2) return tc::fail("a mixed correlation needs two field lists");
    const int coords = fields1[0] ? fields1[0]->coords : 0;
    auto check_side = [&](tc_field* const* fs, int n, int d) -> int {
        for (int k = 0; k < n; ++k) {
            if (!fs[k]) return tc::fail("NULL field in the patch list");
            if (fs[k]->d != d || fs[k]->coords != coords)
                return tc::fail("field types do not match the correlation");
        }
        return 0;
    };
    if (int rc = check_side(fields1, nfield1, c->d)) return rc;
    if (!same) if (int rc = check_side(fields2, nfield2, c->d2)) return rc;
    for (int k = 0; k < nfield1; ++k)
        if (int rc = settled(fields1[k])) return rc;
    for (int k = 0; !same && k < nfield2; ++k)
        if (int rc = settled(fields2[k])) return rc;
    const int n2 = same ? nfield1 : nfield2;
    for (int p = 0; p < npair; ++p)
        if (p1[p] < 0 || p1[p] >= nfield1 || p2[p] < 0 || p2[p] >= n2)
            return tc::fail("patch index out of range in the pair list");
    C2Params P = c->P;
    if (int rc = set_metric(c, metric, coords, P)) return rc;
    if (int rc = check_gpu()) return rc;
    P.quick = quick ? 1 : 0;
    const int nb = P.nbins;
    for (int k = 0; k < 4; ++k) c->counters[k] = 0;
    std::fill(pair_out, pair_out + (size_t)8 * nb * npair, 0.);
    // the pairs that reach the traversal
    std::vector<int> pi, pj, pa, which;
    for (int p = 0; p < npair; ++p) {
        const bool autop = same && p1[p] == p2[p];
        tc_field* fa = fields1[p1[p]];
        tc_field* fb = same ? fields1[p2[p]] : fields2[p2[p]];
        if (!autop && fields_out_of_range(P, coords, fa, fb)) continue;
        pi.push_back(p1[p]); pj.push_back(p2[p]); pa.push_back(autop ? 1 : 0);
        which.push_back(p);
    }
    std::vector<const DeviceTree*> t1(nfield1), t2;
    for (int k = 0; k < nfield1; ++k) t1[k] = &fields1[k]->T;
    if (!same) {
        t2.resize(nfield2);
        for (int k = 0; k < nfield2; ++k) t2[k] = &fields2[k]->T;
    }
    const size_t nh = (size_t)7 * nb;
    std::vector<double> out(nh * pi.size());
    std::vector<unsigned long long> np((size_t)nb * pi.size()), cnt(4);
    {
        std::lock_guard<std::mutex> run_lock(g_run_m);
        if (!g_s2) g_s2 = new_corr2_scratch();
        if (int rc = tc::corr2_run_pairs(*g_s2, c->kind, coords, t1, t2, pi, pj, pa, P, out.data(),
                                         np.data(), cnt.data()))
            return rc;
    }
    const int nxi = corr2_nxi(c->kind);
    for (size_t q = 0; q < which.size(); ++q) {
        const double* o = out.data() + nh * q;
        double* r = pair_out + (size_t)8 * nb * which[q];
        for (size_t t = 0; t < nh; ++t) r[t] = o[t];
        for (int k = 0; k < nb; ++k) r[(size_t)7 * nb + k] = double(np[(size_t)nb * q + k]);
    }
    // totals, added in pair order
    for (int p = 0; p < npair; ++p) {
        const double* r = pair_out + (size_t)8 * nb * p;
        for (int q = 0; q < nxi; ++q)
            for (int k = 0; k < nb; ++k) c->xi[q][k] += r[(size_t)q * nb + k];
        for (int k = 0; k < nb; ++k) {
            c->meanr[k] += r[(size_t)4 * nb + k];
            c->meanlogr[k] += r[(size_t)5 * nb + k];
            c->weight[k] += r[(size_t)6 * nb + k];
            c->npairs[k] += r[(size_t)7 * nb + k];
        }
    }
    for (int k = 0; k < 4; ++k) c->counters[k] = (long long)cnt[k];
    return 0;
}

int tc_corr2_trivially_zero(tc_corr2* c, int metric, int coords, double x1, double y1,
                            double z1, double s1, double x2, double y2, double z2, double s2,
                            int* result) {
    if (!c || !result) return tc::fail("NULL argument");
    C2Params P = c->P;
    if (int rc = set_metric(c, metric, coords, P)) return rc;
    // TriviallyZero3 builds Position<C>(x,y,z) (Sphere normalises) and calls
    // BaseCorr2::triviallyZero, src/Corr2.cpp:590-601, 1040-1048.
    if (coords == Sphere) {
        auto norm = [](double& x, double& y, double& z) {
            double n = std::sqrt(x * x + y * y + z * z);
            if (n == 0.) { x = 1; y = 0; z = 0; return; }
            double r = 1. / n; x *= r; y *= r; z *= r;
        };
        norm(x1, y1, z1); norm(x2, y2, z2);
    }
    const double rsq = metric_dist_sq(P.metric, coords, x1, y1, z1, x2, y2, z2, P.xp, P.yp,
                                      P.zp, s1, s2);
    const double s1ps2 = s1 + s2;
    bool large = rsq >= P.maxsepsq && rsq >= (P.maxsep + s1ps2) * (P.maxsep + s1ps2);
    if (large && los_metric(P.metric)) {   // the metric's tooLargeDist, no r_par range
        bool rin;
        large = metric_prune(P.metric, 0, 0., 0., x1, y1, z1, x2, y2, z2, rsq, s1ps2, false,
                             true, P.minsep, P.minsepsq, P.maxsep, P.maxsepsq, rin);
    }
    *result = large ? 1 : 0;
    return 0;
}

int tc_set_traversal_chunk(long long items) { return tc::set_traversal_chunk(items); }
int tc_set_traversal_mode(int mode) {
    if (mode < 0) return tc::traversal_mode();
    return tc::set_traversal_mode(mode);
}
long long tc_traversal_fallbacks(void) { return tc::persist_fallbacks(); }
int tc_persist_counters(long long* counters4) {
    if (!counters4) return tc::fail("counters4 is required");
    tc::persist_counters(counters4);
    return 0;
}
struct tc_cost_model {
    tc::CostModel* m = nullptr;
    ~tc_cost_model() { tc::cost_model_delete(m); }
};

int tc_cost_model_create(unsigned long long seed, tc_cost_model** out) {
    if (!out) return tc::fail("out is required");
    *out = nullptr;
    if (int rc = check_gpu()) return rc;
    tc::CostModel* m = tc::cost_model_new(seed);
    if (!m) return TC_ERR_HIP;
    *out = new tc_cost_model;
    (*out)->m = m;
    return 0;
}
int tc_cost_model_destroy(tc_cost_model* m) {
    std::lock_guard<std::mutex> run_lock(g_run_m);
    if (g_cost_model_set == m) {
        tc::set_cost_model(nullptr);
        g_cost_model_set = nullptr;
    }
    delete m;
    return 0;
}
int tc_cost_model_sizes(int* nfeatures, int* nparams) {
    if (nfeatures) *nfeatures = tc::CM_NF;
    if (nparams) *nparams = tc::CM_NP;
    return 0;
}
int tc_cost_model_params(tc_cost_model* m, float* params, int set) {
    if (!m || !params) return tc::fail("model and params are required");
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::cost_model_params(m->m, params, set != 0);
}
int tc_cost_model_predict(tc_cost_model* m, const float* X, long long n, float* yhat) {
    if (!m || (n > 0 && (!X || !yhat))) return tc::fail("model, X and yhat are required");
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::cost_model_predict(m->m, X, n, yhat);
}
int tc_cost_model_grad(tc_cost_model* m, const float* X, const float* y, long long n,
                       float* grad, double* loss) {
    if (!m || !X || !y || !grad || !loss || n <= 0) return tc::fail("invalid arguments");
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::cost_model_grad(m->m, X, y, n, grad, loss);
}
int tc_cost_model_train(tc_cost_model* m, const float* X, const float* y, long long n, int steps,
                        int batch, float lr, double* loss) {
    if (!m || !X || !y || !loss || n <= 0) return tc::fail("invalid arguments");
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::cost_model_train(m->m, X, y, n, steps, batch, lr, loss);
}
int tc_set_cost_model(tc_cost_model* m) {
    std::lock_guard<std::mutex> run_lock(g_run_m);
    tc::set_cost_model(m ? m->m : nullptr);
    g_cost_model_set = m;
    return 0;
}
int tc_set_job_log(long long max_samples) {
    if (max_samples < 0) return tc::fail("max_samples must be >= 0");
    std::lock_guard<std::mutex> run_lock(g_run_m);
    tc::set_job_log(max_samples);
    return 0;
}
int tc_job_log_fetch(float* X, float* visits, long long cap, long long* n) {
    if (!n) return tc::fail("n is required");
    std::lock_guard<std::mutex> run_lock(g_run_m);
    *n = tc::job_log_fetch(X, visits, cap);
    return 0;
}

int tc_cov_rowsums(int nrow, int npair, int nq, const double* W, const double* S, double* R) {
    if (nrow < 0 || npair < 0 || nq < 0) return tc::fail("sizes must be >= 0");
    if ((nrow > 0 && nq > 0) && (!R || (npair > 0 && (!W || !S))))
        return tc::fail("W, S and R are required");
    if (int rc = check_gpu()) return rc;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::cov_rowsums(nrow, npair, nq, W, S, R);
}
int tc_cov_gram(int nrow, int ncol, const double* V, const double* w, double* C) {
    if (!V || !C) return tc::fail("V and C are required");
    if (int rc = check_gpu()) return rc;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::cov_gram(nrow, ncol, V, w, C);
}

int tc_set_persist_grab(int div) {
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::set_persist_grab(div);
}
int tc_set_graph_mode(int on) {
    static int cur = 0;
    if (on < 0) return cur;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    cur = on ? 1 : 0;
    tc::set_graph_mode(cur != 0);
    return 0;
}
int tc_graph_stats(long long* captures, long long* replays) {
    if (!captures || !replays) return tc::fail("captures and replays are required");
    tc::graph_stats(captures, replays);
    return 0;
}

int tc_set_persist_params(int levels, long long min_jobs) {
    return tc::set_persist_params(levels, min_jobs);
}
int tc_set_tree_speculation(int mode) { return tc::set_tree_speculation(mode); }
int tc_set_multipole_plain(int on) { return tc::set_multipole_plain(on); }
int tc_set_zeta_mode(int mode) { return tc::set_zeta_mode(mode); }

int tc_quick_assign(const double* centers, int npatch, const double* x, const double* y,
                    const double* z, long long n, int* patch) {
    if (!centers || npatch <= 0 || !x || !y || n < 0 || (n > 0 && !patch))
        return tc::fail("NULL or invalid argument");
    if (n == 0) return 0;
    if (int rc = check_gpu()) return rc;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::quick_assign(centers, npatch, x, y, z, n, patch);
}

int tc_select_patch(int patch, const double* centers, int npatch, const double* x,
                    const double* y, const double* z, long long n, int* use) {
    if (!centers || npatch <= 0 || !x || !y || n < 0 || (n > 0 && !use))
        return tc::fail("NULL or invalid argument");
    if (patch < 0 || patch >= npatch) return tc::fail("patch must be in [0, npatch)");
    if (n == 0) return 0;
    if (int rc = check_gpu()) return rc;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::select_patch(patch, centers, npatch, x, y, z, n, use);
}

int tc_kmeans_run(double* centers, int npatch, int dim, const double* x, const double* y,
                  const double* z, const double* w, long long n, int max_iter, double tol,
                  double size, int sphere, int* niter) {
    if (!centers || npatch <= 0 || (dim != 2 && dim != 3) || !x || !y || (dim == 3 && !z) ||
        !w || n <= 0 || max_iter < 0 || !(tol >= 0.) || !(size >= 0.))
        return tc::fail("NULL or invalid argument");
    if (int rc = check_gpu()) return rc;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::kmeans_run(centers, npatch, dim, x, y, z, w, n, max_iter, tol, size, sphere,
                          niter);
}

// KMeansRun with alt=True (src/KMeans.cpp:774-833): each iteration first estimates every
// patch's inertia on the field's tree, then moves the centres with the inertia-penalised
// assignment of the objects.
int tc_kmeans_run_alt(const tc_field* f, double* centers, int npatch, const double* x,
                      const double* y, const double* z, const double* w, long long n,
                      int max_iter, double tol, double size, int* niter) {
    if (!f) return tc::fail("NULL field");
    const int dim = f->coords == Flat ? 2 : 3;
    if (!centers || npatch <= 0 || !x || !y || (dim == 3 && !z) || !w || n <= 0 ||
        max_iter < 0 || !(tol >= 0.) || !(size >= 0.))
        return tc::fail("NULL or invalid argument");
    if (f->T.top.empty()) return tc::fail("the field has no tree");
    if (int rc = check_gpu()) return rc;
    if (int rc = settled(f)) return rc;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::kmeans_run_tree(centers, npatch, dim, x, y, z, w, n, max_iter, tol, size,
                               f->coords == Sphere ? 1 : 0, f->T.nodes, f->T.nnodes, f->T.dtop,
                               (int)f->T.top.size(), niter);
}

int tc_kmeans_run_noalloc(double* centers, int npatch, int dim, const double* x,
                          const double* y, const double* z, const double* w, long long n,
                          int max_iter, double tol, double size, int sphere, int* niter) {
    if (!centers || npatch <= 0 || (dim != 2 && dim != 3) || !x || !y || (dim == 3 && !z) ||
        !w || n <= 0 || max_iter < 0 || !(tol >= 0.) || !(size >= 0.))
        return tc::fail("NULL or invalid argument");
    if (int rc = check_gpu()) return rc;
    std::lock_guard<std::mutex> run_lock(g_run_m);
    return tc::kmeans_run_noalloc(centers, npatch, dim, x, y, z, w, n, max_iter, tol, size,
                                  sphere, niter);
}

int tc_patch_split(const int* patch, long long n, int npatch, int ncol, const double* const* cols,
                   double* const* out, long long* counts) {
    if (!patch || n < 0 || npatch <= 0 || ncol < 0 || !counts || (ncol > 0 && (!cols || !out)))
        return tc::fail("NULL or invalid argument");
    for (int c = 0; c < ncol; ++c)
        if (!cols[c] || !out[c]) return tc::fail("NULL column");
    // Stable counting sort: thread t counts its contiguous chunk, the chunks' offsets
    // follow in order within each patch, then every thread scatters its chunk in order.
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::max<long long>(1, std::min<long long>({(long long)std::min(hw, 16u),
                                                                    n / 65536 + 1}));
    std::vector<long long> cnt((size_t)nt * npatch, 0);
    std::atomic<bool> bad{false};
    auto chunk = [&](int t, long long& b, long long& e) {
        b = n * t / nt;
        e = n * (t + 1) / nt;
    };
    auto run = [&](auto fn) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(fn, t);
        fn(0);
        for (auto& x : th) x.join();
    };
    run([&](int t) {
        long long b, e;
        chunk(t, b, e);
        long long* ct = cnt.data() + (size_t)t * npatch;
        for (long long i = b; i < e; ++i) {
            const int p = patch[i];
            if (p < 0 || p >= npatch) { bad = true; return; }
            ++ct[p];
        }
    });
    if (bad) return tc::fail("patch index out of range");
    long long run_off = 0;
    for (int p = 0; p < npatch; ++p) {
        long long tot = 0;
        for (int t = 0; t < nt; ++t) {
            long long& c = cnt[(size_t)t * npatch + p];
            const long long k = c;
            c = run_off + tot;       // becomes this chunk's next destination
            tot += k;
        }
        counts[p] = tot;
        run_off += tot;
    }
    run([&](int t) {
        long long b, e;
        chunk(t, b, e);
        long long* dst = cnt.data() + (size_t)t * npatch;
        for (long long i = b; i < e; ++i) {
            const long long d = dst[patch[i]]++;
            for (int c = 0; c < ncol; ++c) out[c][d] = cols[c][i];
        }
    });
    return 0;
}

int tc_corr2_counters(const tc_corr2* c, long long* counters4) {
    if (!c || !counters4) return tc::fail("NULL argument");
    for (int k = 0; k < 4; ++k) counters4[k] = c->counters[k];
    return 0;
}

int tc_corr2_destroy(tc_corr2* c) {
    delete c;
    return 0;
}

int tc_generate_xyz(double* x, double* y, double* z, const double* ra, const double* dec,
                    const double* r, long long n) {
    if (!x || !y || !z || !ra || !dec) return tc::fail("NULL argument");
    for (long long i = 0; i < n; ++i) {
        double sr, cr, sd, cd;
        ::sincos(ra[i], &sr, &cr);
        ::sincos(dec[i], &sd, &cd);
        x[i] = cd * cr; y[i] = cd * sr; z[i] = sd;
        if (r) { x[i] *= r[i]; y[i] *= r[i]; z[i] *= r[i]; }
    }
    return 0;
}

static int g_threads = 1;
int tc_set_num_threads(int n) { g_threads = n > 0 ? n : 1; return 0; }
int tc_get_num_threads(int* n) {
    if (!n) return tc::fail("NULL argument");
    *n = g_threads;
    return 0;
}

int tc_set_timing(int on) { tc::g_timing = on != 0; return 0; }
int tc_kernel_stats(int which, double* ms, long long* launches) {
    if (which < 0 || which >= KS_COUNT || !ms || !launches) return tc::fail("invalid argument");
    *ms = tc::g_stats[which].ms;
    *launches = tc::g_stats[which].launches;
    return 0;
}
int tc_reset_kernel_stats(void) {
    for (int k = 0; k < KS_COUNT; ++k) tc::g_stats[k] = KernelStat();
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// 3-point multipole
// ---------------------------------------------------------------------------
struct tc_corr3 {
    int d = 0;
    tc::C3Params P{};
    double* z[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double *meand2 = nullptr, *meanlogd2 = nullptr, *meand3 = nullptr, *meanlogd3 = nullptr;
    double *weight = nullptr, *weight_im = nullptr, *ntri = nullptr;
    long long counters[5] = {0, 0, 0, 0, 0};
};

extern "C" {

int tc_corr3_create(int d, int bin_type, double minsep, double maxsep, int nbins,
                    double binsize, double b, double a, int max_n, double* zeta0, double* zeta1,
                    double* zeta2, double* zeta3, double* zeta4, double* zeta5, double* zeta6,
                    double* zeta7, double* meand1, double* meanlogd1, double* meand2,
                    double* meanlogd2, double* meand3, double* meanlogd3, double* meanu,
                    double* meanv, double* weight, double* weight_im, double* ntri,
                    tc_corr3** out) {
    (void)meand1; (void)meanlogd1; (void)meanu; (void)meanv;
    if (!out) return tc::fail("out is NULL");
    *out = nullptr;
    if (d != NData && d != GData) {
        tc::set_error("only NNN and GGG multipole correlations are implemented");
        return TC_ERR_UNSUPPORTED;
    }
    if (bin_type != 5) {
        tc::set_error("only LogMultipole binning is implemented for 3-point correlations");
        return TC_ERR_UNSUPPORTED;
    }
    if (nbins <= 0 || max_n < 0 || !(minsep > 0.) || !(maxsep > minsep) || !(binsize > 0.))
        return tc::fail("invalid binning");
    if (!meand2 || !meanlogd2 || !meand3 || !meanlogd3 || !weight || !weight_im || !ntri)
        return tc::fail("meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri are required");
    if (d == GData && (!zeta0 || !zeta1 || !zeta2 || !zeta3 || !zeta4 || !zeta5 || !zeta6 || !zeta7))
        return tc::fail("GGG needs zeta0..zeta7");
    tc_corr3* c = new tc_corr3;
    c->d = d;
    tc::C3Params& P = c->P;
    P.nbins = nbins; P.maxn = max_n; P.D = d; P.quick = 0;
    P.minsep = minsep; P.maxsep = maxsep; P.binsize = binsize; P.b = b; P.a = a;
    // BaseCorr3::BaseCorr3, src/Corr3.cpp:108-120
    P.logminsep = std::log(minsep);
    P.minsepsq = minsep * minsep; P.maxsepsq = maxsep * maxsep;
    P.bsq = b * b; P.asq = a * a;
    P.maxbsq_eff = P.maxsepsq * std::min(P.bsq, P.asq);   // getEffectiveBSq, BinType.h:50-51
    double* zz[8] = {zeta0, zeta1, zeta2, zeta3, zeta4, zeta5, zeta6, zeta7};
    for (int k = 0; k < 8; ++k) c->z[k] = zz[k];
    c->meand2 = meand2; c->meanlogd2 = meanlogd2; c->meand3 = meand3; c->meanlogd3 = meanlogd3;
    c->weight = weight; c->weight_im = weight_im; c->ntri = ntri;
    *out = c;
    return 0;
}

}  // extern "C"

namespace {

// One multipole pass: c1 from fields[a], lists from fields[b] (and fields[c] when c >= 0,
// with the ordered code `o`).
struct Pass3 { int a, b, c, o; };

// Runs the passes and adds their sums into the correlation's arrays.
int corr3_passes(tc_corr3* c, tc_field* const* fs, int nfs, const Pass3* passes, int npass,
                 int quick, int metric) {
    if (!c) return tc::fail("NULL argument");
    for (int i = 0; i < nfs; ++i)
        if (!fs[i]) return tc::fail("NULL argument");
    if (metric != 1) {
        tc::set_error("only the Euclidean metric is implemented");
        return TC_ERR_UNSUPPORTED;
    }
    for (int i = 0; i < nfs; ++i) {
        if (fs[i]->d != c->d) return tc::fail("field type does not match the correlation");
        if (fs[i]->coords != fs[0]->coords) return tc::fail("the fields' coordinate systems differ");
    }
    if (int rc = check_gpu()) return rc;
    for (int i = 0; i < nfs; ++i)
        if (int rc = settled(fs[i])) return rc;
    tc::C3Params P = c->P;
    P.quick = quick ? 1 : 0;
    const int ntot = P.nbins * P.nbins * (2 * P.maxn + 1);
    std::vector<double> out((size_t)15 * ntot), acc((size_t)15 * ntot, 0.);
    unsigned long long tot[5] = {0, 0, 0, 0, 0};
    std::lock_guard<std::mutex> run_lock(g_run_m);
    for (int p = 0; p < npass; ++p) {
        const Pass3& ps = passes[p];
        unsigned long long cnt[5] = {0, 0, 0, 0, 0};
        int rc = tc::corr3_run_trees(c->d, fs[0]->coords, fs[ps.a]->T, fs[ps.b]->T,
                                     ps.c >= 0 ? &fs[ps.c]->T : nullptr, ps.o, P, out.data(), cnt);
        if (rc) return rc;
        for (size_t j = 0; j < acc.size(); ++j) acc[j] += out[j];
        for (int k = 0; k < 5; ++k) tot[k] += cnt[k];
    }
    if (c->d == GData)
        for (int q = 0; q < 8; ++q)
            for (int j = 0; j < ntot; ++j) c->z[q][j] += acc[(size_t)q * ntot + j];
    double* dst[7] = {c->meand2, c->meanlogd2, c->meand3, c->meanlogd3, c->weight, c->weight_im,
                      c->ntri};
    for (int q = 0; q < 7; ++q)
        for (int j = 0; j < ntot; ++j) dst[q][j] += acc[(size_t)(8 + q) * ntot + j];
    for (int k = 0; k < 5; ++k) c->counters[k] = (long long)tot[k];
    return 0;
}

}  // namespace

extern "C" {

// BaseCorr3::multipole(field), src/Corr3.cpp:1124-1175 (via ProcessAuto :3654-3720).
int tc_corr3_process_auto(tc_corr3* c, tc_field* f, int dots, int quick, int metric) {
    (void)dots;
    tc_field* fs[1] = {f};
    const Pass3 p{0, 0, -1, 0};
    return corr3_passes(c, fs, 1, &p, 1, quick, metric);
}

// ProcessCross12c, src/Corr3.cpp:3731-3768: ordered 1 is multipole(f1, f2); ordered 0
// first adds multipole(f2, f1, f2, ordered 1).
int tc_corr3_process_cross12(tc_corr3* c, tc_field* f1, tc_field* f2, int ordered, int dots,
                             int quick, int metric) {
    (void)dots;
    if (ordered != 0 && ordered != 1) return tc::fail("processCross12: ordered must be 0 or 1");
    tc_field* fs[2] = {f1, f2};
    const Pass3 p0[2] = {{1, 0, 1, 1}, {0, 1, -1, 0}};
    return ordered == 0 ? corr3_passes(c, fs, 2, p0, 2, quick, metric)
                        : corr3_passes(c, fs, 2, p0 + 1, 1, quick, metric);
}

// ProcessCross21c, src/Corr3.cpp:3831-3870: ordered 3 is multipole(f1, f1, f2, 4);
// ordered 0 the three placements of f2 with ordered 1.
int tc_corr3_process_cross21(tc_corr3* c, tc_field* f1, tc_field* f2, int ordered, int dots,
                             int quick, int metric) {
    (void)dots;
    if (ordered != 0 && ordered != 3) return tc::fail("processCross21: ordered must be 0 or 3");
    tc_field* fs[2] = {f1, f2};
    const Pass3 p0[3] = {{1, 0, 0, 1}, {0, 1, 0, 1}, {0, 0, 1, 1}};
    const Pass3 p3[1] = {{0, 0, 1, 4}};
    return ordered == 0 ? corr3_passes(c, fs, 2, p0, 3, quick, metric)
                        : corr3_passes(c, fs, 2, p3, 1, quick, metric);
}

// ProcessCrossc, src/Corr3.cpp:3938-3985: the passes of each ordered code 0-4.
int tc_corr3_process_cross(tc_corr3* c, tc_field* f1, tc_field* f2, tc_field* f3, int ordered,
                           int dots, int quick, int metric) {
    (void)dots;
    tc_field* fs[3] = {f1, f2, f3};
    static const Pass3 o0[3] = {{1, 0, 2, 1}, {2, 0, 1, 1}, {0, 1, 2, 1}};
    static const Pass3 o1[1] = {{0, 1, 2, 1}};
    static const Pass3 o2[2] = {{0, 1, 2, 4}, {2, 1, 0, 4}};
    static const Pass3 o3[2] = {{0, 1, 2, 4}, {1, 0, 2, 4}};
    static const Pass3 o4[1] = {{0, 1, 2, 4}};
    switch (ordered) {
        case 0: return corr3_passes(c, fs, 3, o0, 3, quick, metric);
        case 1: return corr3_passes(c, fs, 3, o1, 1, quick, metric);
        case 2: return corr3_passes(c, fs, 3, o2, 2, quick, metric);
        case 3: return corr3_passes(c, fs, 3, o3, 2, quick, metric);
        case 4: return corr3_passes(c, fs, 3, o4, 1, quick, metric);
        default: return tc::fail("processCross: ordered must be 0 to 4");
    }
}

int tc_corr3_counters(const tc_corr3* c, long long* counters4) {
    if (!c || !counters4) return tc::fail("NULL argument");
    for (int k = 0; k < 4; ++k) counters4[k] = c->counters[k];
    return 0;
}

// TriviallyZero3 → BaseCorr3::triviallyZero, src/Corr3.cpp:275-299, 4069-4130, for the
// Euclidean metric: whether no triangle with one vertex in each of three cells (centre,
// size) can land in a LogMultipole bin.  p13: cell 2 is cell 1 or cell 3, so only the
// distance from cell 1 to cell 3 counts (BinTypeHelper::tooLargeDist; the Euclidean
// metric's own test is always true).  Otherwise trivial_stop, include/BinType.h (the
// LogMultipole helper): ordered, any of d2, d3 out of range stops; unordered, two of
// d1, d2, d3 below min_sep or two above max_sep.
int tc_corr3_trivially_zero(tc_corr3* c, int coords, double x1, double y1, double z1,
                            double s1, double x2, double y2, double z2, double s2, double x3,
                            double y3, double z3, double s3, int ordered, int p13,
                            int* result) {
    if (!c || !result) return tc::fail("NULL argument");
    if (coords != Flat && coords != ThreeD && coords != Sphere)
        return tc::fail("coords must be 1 (Flat), 2 (ThreeD) or 3 (Sphere)");
    const C3Params& P = c->P;
    if (coords == Flat) { z1 = z2 = z3 = 0.; }
    if (coords == Sphere) {   // Position<Sphere> puts its point on the unit sphere
        auto norm = [](double& x, double& y, double& z) {
            double n = std::sqrt(x * x + y * y + z * z);
            if (n == 0.) { x = 1; y = 0; z = 0; return; }
            double r = 1. / n; x *= r; y *= r; z *= r;
        };
        norm(x1, y1, z1); norm(x2, y2, z2); norm(x3, y3, z3);
    }
    auto dsq = [](double ax, double ay, double az, double bx, double by, double bz) {
        double dx = ax - bx, dy = ay - by, dz = az - bz;
        return dx * dx + dy * dy + dz * dz;
    };
    auto sq = [](double v) { return v * v; };
    if (p13) {
        double d = dsq(x1, y1, z1, x3, y3, z3), s = s1 + s3;
        *result = (d >= P.maxsepsq && d >= sq(P.maxsep + s)) ? 1 : 0;
        return 0;
    }
    const double d1sq = dsq(x2, y2, z2, x3, y3, z3);
    const double d2sq = dsq(x1, y1, z1, x3, y3, z3);
    const double d3sq = dsq(x1, y1, z1, x2, y2, z2);
    const double mn = P.minsep, mnsq = P.minsepsq, mx = P.maxsep, mxsq = P.maxsepsq;
    auto small = [&](double d, double s) {
        return d < mnsq && s < mn && (s == 0. || d < sq(mn - s));
    };
    auto large = [&](double d, double s) { return d >= mxsq && (s == 0. || d >= sq(mx + s)); };
    *result = 1;
    int nsmall = 0, nlarge = 0;
    if (small(d2sq, s1 + s3)) { if (ordered) return 0; ++nsmall; }
    if (small(d3sq, s1 + s2)) { if (ordered) return 0; ++nsmall; }
    if (!ordered && small(d1sq, s2 + s3)) ++nsmall;
    if (!ordered && nsmall >= 2) return 0;
    if (large(d2sq, s1 + s3)) { if (ordered) return 0; ++nlarge; }
    if (large(d3sq, s1 + s2)) { if (ordered) return 0; ++nlarge; }
    if (!ordered && large(d1sq, s2 + s3)) ++nlarge;
    if (!ordered && nlarge >= 2) return 0;
    *result = 0;
    return 0;
}

int tc_corr3_zeta_cells(const tc_corr3* c, long long* cells) {
    if (!c || !cells) return tc::fail("NULL argument");
    *cells = c->counters[4];
    return 0;
}

int tc_corr3_destroy(tc_corr3* c) {
    delete c;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// HBM counter calibration: a plain streaming copy of known size.
// ---------------------------------------------------------------------------
__global__ void k_copy_calibration(const double* __restrict__ src, double* __restrict__ dst,
                                   long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

extern "C" int tc_copy_calibration(const double* src, double* dst, long long n) {
    if (!src || !dst || n <= 0) return tc::fail("invalid argument");
    if (int rc = check_gpu()) return rc;
    hipLaunchKernelGGL(k_copy_calibration, dim3(2048), dim3(256), 0, tc::stream(), src, dst, n);
    TC_HIP(hipGetLastError());
    TC_HIP(hipStreamSynchronize(tc::stream()));
    return 0;
}
