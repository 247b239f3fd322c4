// treecorr_amd: engine-level types and entry points shared by the host-side
// translation units (tree build, 2-point traversal, 3-point multipole, C ABI).
#pragma once
#include "tc_common.h"
#include <memory>
#include <vector>

namespace tc {

// A tree resident in GPU memory.
// Device arrays come from the library's caching allocator (dev_alloc/dev_free), so
// they carry their sizes.
// The exact pass of a build's speculative top levels, still running on its own streams:
// it writes the node records of those levels (above the top cells, which the
// traversals never read) and the field's centre and size.  settle_tree waits for it.
struct TreeCheck;

struct DeviceTree {
    std::shared_ptr<TreeCheck> pending;   // set while the exact pass runs
    DNode* nodes = nullptr;
    size_t nodes_bytes = 0;
    int nnodes = 0;
    int depth = 0;            // levels of the tree (a leaf at level depth-1 at most)
    std::vector<int> top;     // top-level node ids in TreeCorr's order
    int* dtop = nullptr;      // the same on the device
    size_t dtop_bytes = 0;
    FieldInfo info{};
};

int build_tree(int D, int C, const double* x, const double* y, const double* z, const double* d1,
               const double* d2, const double* w, const double* wpos, long long n,
               double minsize, double maxsize, int sm, int brute, int mintop, int maxtop,
               DeviceTree& T);

// Trees of several root ranges [bounds[r], bounds[r+1]) built together, sharing one
// node array (node ids, child links and top lists are indices into it).
struct MultiTree {
    std::shared_ptr<TreeCheck> pending;   // single-root builds only
    DNode* nodes = nullptr;     // owned by the caller after the build (dev_free)
    size_t nodes_bytes = 0;
    int nnodes = 0;
    int depth = 0;                       // levels of the deepest tree
    std::vector<std::vector<int>> top;   // per root, in TreeCorr's order
    std::vector<FieldInfo> info;         // per root
};
// Speculative centroids for the top tree levels: 0 off, 1 on for levels of large
// segments (default), 2 on every top level with the check forced to fail, so the exact
// rebuild runs, 3 on every top level (2 and 3 for tests).  Only levels above min_top
// are speculated on.  A negative mode changes nothing: -2 returns the number of builds
// that returned before their exact pass ended (validated decisions), any other the
// number of exact rebuilds so far.
int set_tree_speculation(int mode);
// Waits for T's exact pass, if one is running, and fills T.info; an error if that pass
// found a decision different from the validated speculative one (which the validation
// rules out).  Everything that reads node records above the top cells or T.info calls
// it first; the tree's owner calls it before freeing the nodes.
int settle_tree(DeviceTree& T);
int build_trees(int D, int C, const double* x, const double* y, const double* z, const double* d1,
                const double* d2, const double* w, const double* wpos, long long n,
                const std::vector<long long>& bounds, double minsize, double maxsize, int sm,
                int brute, int mintop, int maxtop, MultiTree& M);

// Log-binning parameters of one 2-point correlation (BaseCorr2, src/Corr2.cpp:51-77).
struct C2Params {
    int nbins, quick;
    double minsep, maxsep, binsize, b, a;
    double logminsep, halfminsep, minsepsq, maxsepsq, bsq, asq;
    double inv_binsize;         // 1/binsize rounded, for bin coordinates away from an edge
    int metric;                 // Metric: Euclidean, Arc or Periodic
    double xp, yp, zp;          // the Periodic metric's periods
    int rpar_on;                // a [minrpar, maxrpar] range was given (3-d)
    double minrpar, maxrpar;
};

// Lloyd iterations of k-means over the objects (centers in/out, npatch x dim).
int kmeans_run(double* centers, int npatch, int dim, const double* x, const double* y,
               const double* z, const double* w, long long n, int max_iter, double tol,
               double size, int sphere, int* niter);
// kmeans_run with the alternate algorithm's inertia penalty (KMeansRun1 with alt,
// src/KMeans.cpp:791-814) computed on the field's tree (nodes, top cells on the device)
int kmeans_run_tree(double* centers, int npatch, int dim, const double* x, const double* y,
                    const double* z, const double* w, long long n, int max_iter, double tol,
                    double size, int sphere, const DNode* nodes, int nnodes, const int* dtop,
                    int ntop, int* niter);
// kmeans_run on device-resident columns and centres, with a fixed static scratch array
// (no allocation): at most 8192 centres.
int kmeans_run_noalloc(double* centers, int npatch, int dim, const double* x, const double* y,
                       const double* z, const double* w, long long n, int max_iter, double tol,
                       double size, int sphere, int* niter);
// QuickAssign on the device: patch[i] = the nearest of npatch centres (2 or 3 values
// each; z == nullptr means 2), host arrays in and out.
int quick_assign(const double* centers, int npatch, const double* x, const double* y,
                 const double* z, long long n, int* patch);
// SelectPatch on the device: use[i] = 1 unless a centre other than `patch` is strictly
// nearer (host arrays in and out).
int select_patch(int patch, const double* centers, int npatch, const double* x, const double* y,
                 const double* z, long long n, int* use);

// The learned cost model of the traversal's jobs (tc_cost.hip): a perceptron of
// CM_NF inputs, two hidden layers of CM_H ReLU units and one output, FP32.
constexpr int CM_NF = 11, CM_H = 64;
constexpr int CM_NP = CM_H * CM_NF + CM_H + CM_H * CM_H + CM_H + CM_H + 1;
struct CostModel;
CostModel* cost_model_new(unsigned long long seed);
void cost_model_delete(CostModel* m);
int cost_model_params(CostModel* m, float* host, bool set);
int cost_model_predict(CostModel* m, const float* X, long long n, float* yhat);
int cost_model_predict_device(const CostModel* m, const float* dX, long long n, float* dyhat,
                              hipStream_t st);
int cost_model_grad(CostModel* m, const float* X, const float* Y, long long n, float* grad,
                    double* loss);
int cost_model_train(CostModel* m, const float* X, const float* Y, long long n, int steps,
                     int batch, float lr, double* loss);

// Covariance linear algebra (tc_cov.hip): R = W S (the design matrix's rows from the
// patch-pair sums) and C = sum_r w_r (v_r - mean)^T (v_r - mean); host arrays, row major.
int cov_rowsums(int nrow, int npair, int nq, const double* W, const double* S, double* R);
int cov_gram(int nrow, int ncol, const double* V, const double* w, double* C);

int set_traversal_chunk(long long items);   // 0 restores the default
// Traversal modes: chunked launches driven from the host, or one persistent launch whose
// workgroups take jobs by a shared counter or from static equal-count shares.
// The learned modes take the cost model's predictions: static shares of equal predicted
// work, or jobs handed out by the counter in order of decreasing predicted work.
enum { TRAV_CHUNKED = 0, TRAV_PERSIST_DYNAMIC = 1, TRAV_PERSIST_STATIC = 2,
       TRAV_PERSIST_LEARNED_STATIC = 3, TRAV_PERSIST_LEARNED_DYNAMIC = 4 };
void set_cost_model(const CostModel* m);
void set_job_log(long long max_samples);     // 0: off
long long job_log_fetch(float* X, float* V, long long cap);   // returns the samples kept
int set_traversal_mode(int mode);
int traversal_mode();
long long persist_fallbacks();
void persist_counters(long long* c4);   // visits, accepted, leaf, refined of the last persistent launch
int set_persist_params(int levels, long long min_jobs);   // 0, 0: the defaults
int set_persist_grab(int div);          // a shared-counter grab: 1/div of a fair share
void set_graph_mode(bool on);           // capture persistent launches as hipGraphs
void graph_stats(long long* captures, long long* replays);
struct Corr2Scratch;
Corr2Scratch* new_corr2_scratch();
void delete_corr2_scratch(Corr2Scratch*);
int corr2_run(Corr2Scratch& S, int D, int C, const DNode* nodes1, const int* dtop1, int ntop1,
              int depth1, const DNode* nodes2, const int* dtop2, int ntop2, int depth2,
              const C2Params& P, double* out, unsigned long long* np, unsigned long long* cnt);
// All patch pairs (pi[p], pj[p]) of a patch-based correlation in one traversal;
// trees2 empty means side 2 is side 1.  Per-pair sums (7*nbins doubles, nbins
// npairs each) are written to host arrays.
int corr2_run_pairs(Corr2Scratch& S, int D, int C, const std::vector<const DeviceTree*>& trees1,
                    const std::vector<const DeviceTree*>& trees2, const std::vector<int>& pi,
                    const std::vector<int>& pj, const std::vector<int>& pauto, const C2Params& P,
                    double* pair_out, unsigned long long* pair_npairs, unsigned long long* cnt);

// LogMultipole parameters of one 3-point correlation (BaseCorr3, src/Corr3.cpp:70-127).
struct C3Params {
    int nbins, maxn, D, quick;
    int plain;      // Gn accumulation with every lane in multipole order (the untuned form)
    // Two lists (multipole(f1, f2, f3, ordered), src/Corr3.cpp:1238-1305): order3 is the
    // ordered code, 1 or 4; list entries with node index >= base3 belong to the second
    // list.  order3 = 0: one list (the auto-correlation and multipole(f1, f2)).
    int order3, base3, pad;
    double minsep, maxsep, binsize, b, a;
    double logminsep, minsepsq, maxsepsq, bsq, asq, maxbsq_eff;
};

// Multipole auto-correlation of one device tree: writes 15*ntot sums (zeta0..7,
// meand2, meanlogd2, meand3, meanlogd3, weight, weight_im, ntri) and 5 counters (Gn
// calls, zeta calls, zeta rows, stages, zeta cells).
int corr3_run(int D, int C, const DNode* nodes, int nnodes, const int* dtop, int ntop,
              const C3Params& P, double* out, unsigned long long* cnt);
// One multipole pass over several trees laid out in one node array: c1 runs over the nc1
// nodes c1top (device), the list starts as the nl nodes ltop (device); P.order3 and
// P.base3 select the two-list form.  Writes and counts as corr3_run does.
int corr3_run_x(int D, int C, const DNode* nodes, int nnodes, const int* c1top, int nc1,
                const int* ltop, int nl, const C3Params& P, double* out,
                unsigned long long* cnt);
// One multipole pass of BaseCorr3::multipole (src/Corr3.cpp:1124-1305): c1 from t1, the
// list from t2 (t3 NULL), or from t2 and t3 with ordered 1 or 4.  The trees are copied
// into one node array unless t3 is NULL and t1, t2 are the same tree.
int corr3_run_trees(int D, int C, const DeviceTree& t1, const DeviceTree& t2,
                    const DeviceTree* t3, int ordered, const C3Params& P, double* out,
                    unsigned long long* cnt);
int set_multipole_plain(int on);   // < 0: returns the current setting
// zeta contraction: 0 FP64 (default), 1 FP32 MFMA, 2 split-BF16 MFMA; < 0 returns it
int set_zeta_mode(int mode);

}  // namespace tc
