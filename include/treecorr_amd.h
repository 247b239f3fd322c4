/* treecorr_amd: C ABI of the MI355X ball-tree correlation engine.
 *
 * This is the drop-in boundary for TreeCorr's compiled engine, the pybind11
 * module `_treecorr` (src/module.cpp:21-26).  Every entry point replaces one
 * reference binding, cited below, with plain pointers and sizes.  The Python
 * package treecorr_amd binds these with ctypes; INTEGRATION.md shows the
 * binding a TreeCorr maintainer would add.
 *
 * Conventions
 *   - Every function returns an int status: TC_OK (0) or a TC_ERR_* code; the
 *     message of the last failure on the calling thread is tc_last_error().
 *   - Arrays given to *_create are host arrays unless the name says _device.
 *   - Output arrays given to tc_corr2_create / tc_corr3_create are borrowed
 *     (the caller owns them) and every process call ADDS into them, exactly as
 *     the reference does (include/Corr2.h:210-213).
 *   - Calls are synchronous.  A handle must not be used from two threads at once.
 */
#ifndef TREECORR_AMD_H
#define TREECORR_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define TC_OK 0
#define TC_ERR_INVALID 1
#define TC_ERR_HIP 2
#define TC_ERR_UNSUPPORTED 3

/* Enumerations, values as in the reference:
 *   data types  N=0, K=1, Z=2, V=3, G=4, T=5, Q=6 (include/Cell.h:34 DataType)
 *   coords      Flat=1, ThreeD=2, Sphere=3 (include/Position.h:25)
 *   bin types   Log=0, LogMultipole=5    (include/BinType.h:35)
 *   metrics     Euclidean=1, Rperp=2, Rlens=3, Arc=4, OldRperp=5, Periodic=6
 *               (include/Metric.h:29)
 *   split       Middle=0, Median=1, Mean=2, Random=3 (include/Cell.h:26) */

typedef struct tc_field tc_field;
typedef struct tc_corr2 tc_corr2;
typedef struct tc_corr3 tc_corr3;

/* Library version (major*10000 + minor*100 + patch). */
int tc_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* tc_last_error(void);

/* Number of visible GPUs (0 when no GPU is usable). */
int tc_device_count(int* count);

/* ---- Fields: replace {N,K,Z,G}Field{Flat,Sphere,ThreeD}(...) ----
 * Reference: BuildNField/BuildKField/BuildAnyZField, src/Field.cpp:430-506,
 * bound at src/Field.cpp:684-690; the tree itself is Field<D,C>::Field +
 * BuildCells, src/Field.cpp:163-277 (built on the GPU here, eagerly).
 * data_type 0 (N), 1 (K), 2 (Z, spin-0 complex) or 4 (G); d1 is k (K fields), z1 or g1;
 * d2 is z2 or g2.  z and wpos may be NULL. */
int tc_field_create(int data_type, int coords, const double* x, const double* y,
                    const double* z, const double* d1, const double* d2, const double* w,
                    const double* wpos, long long nobj, double minsize, double maxsize,
                    int split_method, long long seed, int brute, int mintop, int maxtop,
                    tc_field** out);

/* The same with the catalog columns already resident in GPU memory. */
int tc_field_create_device(int data_type, int coords, const double* x, const double* y,
                           const double* z, const double* d1, const double* d2,
                           const double* w, const double* wpos, long long nobj,
                           double minsize, double maxsize, int split_method, long long seed,
                           int brute, int mintop, int maxtop, tc_field** out);

/* BaseField::getNObj / getSize / getNTopLevel, src/Field.cpp:624-628. */
/* The fields of several patches in one call: objects [bounds[r], bounds[r+1]) of the
 * columns form patch r (bounds[0] = 0, bounds[nroot] = nobj, every patch non-empty).
 * All trees are built in one level-synchronous pass, the patches being the level-0
 * segments, and share one node array.  Replaces one Field construction per patch
 * catalog in TreeCorr's patch loop (treecorr/corr2base.py:855-1050 via
 * catalog.py:2230-2316, src/Field.cpp:430-506); each out[r] is the same tree as
 * tc_field_create on patch r's columns.  The shared node array is freed with the
 * last of the nroot fields. */
int tc_field_create_multi(int data_type, int coords, const double* x, const double* y,
                          const double* z, const double* d1, const double* d2, const double* w,
                          const double* wpos, long long nobj, const long long* bounds, int nroot,
                          double minsize, double maxsize, int split_method, long long seed,
                          int brute, int mintop, int maxtop, tc_field** out);
int tc_field_nobj(const tc_field* f, long long* nobj);
int tc_field_size(const tc_field* f, double* size);
int tc_field_ntop(const tc_field* f, long long* ntop);
/* Field::_center (3 doubles; z = 0 for Flat), src/Field.cpp:199-203. */
int tc_field_center(const tc_field* f, double* center3);
/* Number of tree nodes, and a copy of them (48-byte records: x,y,z double;
 * size,w,a,b float; child,n int) plus the ordered top-level node ids. */
int tc_field_nnodes(const tc_field* f, long long* nnodes);
int tc_field_copy_nodes(const tc_field* f, void* nodes_out, int* top_out);
/* ~BaseField, src/Field.cpp:279-285. */
int tc_field_destroy(tc_field* f);

/* ---- 2-point: replace {NN,KK,GG}Corr(...) + processAuto/processCross ----
 * Reference: BuildCorr2 src/Corr2.cpp:842-864 bound at :1135-1145;
 * ProcessAuto/ProcessCross :886-993 bound at :1151-1157 (BaseCorr2::process,
 * :146-288).  (d1, d2) is a field type with itself: NN (0,0), KK (1,1), ZZ (2,2), VV (3,3),
 * GG (4,4), TT (5,5), QQ (6,6); or the mixed N or K (0 or 1) with K, Z, V, G, T or Q
 * (src/Corr2.cpp DirectHelper<1>/<2>, :408-442; cross only, f1/fields1 the d1 side);
 * bin_type Log.  xp, yp, zp are the periods of the
 * Periodic metric (0 when unused).  minrpar/maxrpar: the r_par range (3-d only; -DBL_MAX
 * and DBL_MAX for none, BaseCorr2::nontrivialRPar, include/Corr2.h:58-62; ParHelper,
 * include/Metric.h:37-106).  The process calls take metric 1 (Euclidean), 2 (Rperp), 3
 * (Rlens) or 5 (OldRperp) (3-d only, include/Metric.h:268-651), 4 (Arc,
 * spherical or 3-d coordinates, include/Metric.h:656-810: the great-circle angle; in 3-d
 * atan2(|p1 x p2|, p1.p2) with the sizes scaled by 1/|p|) or 6 (Periodic, flat or 3-d
 * coordinates, include/Metric.h:813-838: every separation wrapped into [-p/2, p/2]).  xi arrays: KK and NK xi0; the other mixed kinds xi0 (xi) and xi1
 * (xi_im); ZZ, VV, GG, TT, QQ xi0..xi3 (xip, xip_im, xim, xim_im); the unused ones may be
 * NULL. */
int tc_corr2_create(int d1, int d2, int bin_type, double minsep, double maxsep, int nbins,
                    double binsize, double b, double a, double minrpar, double maxrpar,
                    double xp, double yp, double zp, double* xi0, double* xi1, double* xi2,
                    double* xi3, double* meanr, double* meanlogr, double* weight,
                    double* npairs, tc_corr2** out);
int tc_corr2_process_auto(tc_corr2* c, tc_field* f, int dots, int quick, int metric);
int tc_corr2_process_cross(tc_corr2* c, tc_field* f1, tc_field* f2, int dots, int quick,
                           int metric);
/* All patch pairs of a patch-based correlation in one device traversal.  Replaces
 * the Python loop of corr2base.py:855-1050 (_process_all_auto/_process_all_cross)
 * that calls ProcessAuto/ProcessCross (src/Corr2.cpp:886-993) once per patch pair
 * through _single_process12 (:836-853).  Pair p correlates fields1[p1[p]] with
 * fields2[p2[p]]; with fields2 == NULL both sides are fields1 and a pair with
 * p1[p] == p2[p] is that patch's auto-correlation.  A cross pair whose fields are
 * entirely out of range is skipped as BaseCorr2::process does (:215-231).
 * pair_out receives 8*nbins doubles per pair (xi0..xi3, meanr, meanlogr, weight,
 * npairs; raw sums, zeros for unused xi), and the same sums are added to the
 * correlation's arrays.  Results equal per-pair process calls to 1e-12. */
int tc_corr2_process_pairs(tc_corr2* c, tc_field* const* fields1, int nfield1,
                           tc_field* const* fields2, int nfield2, const int* p1, const int* p2,
                           int npair, int quick, int metric, double* pair_out);
/* Patch of every object from the patch centres: the nearest centre, computed on the
 * GPU.  Replaces QuickAssign, src/KMeans.cpp:890-942 (bound at src/Field.cpp:700), which
 * Catalog._assign_patches calls for a catalog given patch_centers
 * (treecorr/catalog.py:1556-1566).  centers holds npatch rows of (x,y) or, when z is
 * not NULL, (x,y,z); the squared distance is (dx^2+dy^2)+dz^2 and the lowest index
 * wins a tie, as in the reference.  patch receives int32 labels. */
int tc_quick_assign(const double* centers, int npatch, const double* x, const double* y,
                    const double* z, long long n, int* patch);
/* SelectPatch, src/KMeans.cpp:944-1011 (bound in src/module.cpp; called by
 * treecorr/catalog.py:1576-1594 _get_patch_index for a catalog given patch_centers and a
 * single patch): use[i] = 1 unless a centre other than `patch` is strictly nearer object
 * i (a tie keeps it); centres npatch x 2 (z == NULL) or x 3.  On the GPU. */
int tc_select_patch(int patch, const double* centers, int npatch, const double* x,
                    const double* y, const double* z, long long n, int* use);
/* Lloyd iterations of k-means on the GPU: KMeansRun (src/KMeans.cpp:773-833, bound
 * at src/Field.cpp:633) with alt = false, which treecorr/field.py:552-560
 * kmeans_refine_centers calls.  centers (npatch rows of dim = 2 or 3 values) is refined
 * in place over the n objects (x, y[, z], weights w): each object joins its nearest
 * centre, centres move to their objects' w-weighted means (normalised to the unit
 * sphere when sphere != 0; an empty patch's centre becomes the origin, as in the
 * reference), until the rms shift is below tol * size or max_iter steps ran; niter
 * receives the steps.  The reference iterates over tree cells, whose float weights
 * and summation order make its centres differ from these at the ~1e-7 level. */
int tc_kmeans_run(double* centers, int npatch, int dim, const double* x, const double* y,
                  const double* z, const double* w, long long n, int max_iter, double tol,
                  double size, int sphere, int* niter);
/* KMeansRun with alt=True (src/KMeans.cpp:774-833, bound at :1066-1067; called from
 * treecorr/field.py:454-560): every iteration first estimates each patch's inertia on
 * the field's tree (CalculateInertia over the cells where the candidate pruning of
 * FindCellsInPatches stops, :388-447, 544-643), then moves the centres with the
 * assignment penalised by that inertia.  The objects (x, y, z, w: host arrays, positions
 * as the tree holds them) and the field must be the same catalog's. */
int tc_kmeans_run_alt(const tc_field* f, double* centers, int npatch, const double* x,
                      const double* y, const double* z, const double* w, long long n,
                      int max_iter, double tol, double size, int* niter);
/* tc_kmeans_run without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * tc_kmeans_run, except that centers, x, y, z and w are DEVICE pointers (centers is
 * refined in place on the device).  The per-centre sums live in one fixed static device
 * array of 8192 * 5 + 8 doubles (about 320 KB, held by each GPU the library runs on
 * until the process ends); npatch above 8192 is rejected with an error.  It calls no
 * hipMalloc/hipFree.  Calls that overlap on one GPU share that array and race: run them
 * one after another. */
int tc_kmeans_run_noalloc(double* centers, int npatch, int dim, const double* x,
                          const double* y, const double* z, const double* w, long long n,
                          int max_iter, double tol, double size, int sphere, int* niter);
/* Splits a catalog's columns by patch: a stable counting sort on host threads.
 * Replaces the per-patch np.where + fancy indexing of Catalog.get_patches
 * (treecorr/catalog.py:2966-3050).  out[c] receives cols[c] ordered by patch
 * (objects of one patch keep their order); counts[p] the size of patch p.
 * patch[i] must be in [0, npatch). */
int tc_patch_split(const int* patch, long long n, int npatch, int ncol, const double* const* cols,
                   double* const* out, long long* counts);
/* BaseCorr2::triviallyZero / TriviallyZero, src/Corr2.cpp:590-601, 1107-1128. */
int tc_corr2_trivially_zero(tc_corr2* c, int metric, int coords, double x1, double y1,
                            double z1, double s1, double x2, double y2, double z2, double s2,
                            int* result);
/* Work counters of the last process call: node-pair visits, accepted pairs,
 * accepted leaf-leaf pairs, refined single-bin tests (no reference counterpart;
 * used for roofline accounting). */
int tc_corr2_counters(const tc_corr2* c, long long* counters4);
/* Node pairs per traversal launch (default 2^26; clamped to [256, 2^26]; 0 restores
 * the default).  Results do not depend on it; tests use small values to exercise
 * the multi-launch work stack.  No reference counterpart. */
int tc_set_traversal_chunk(long long items);
/* How the 2-point traversal runs.  mode 1 (default): one persistent launch per
 * correlation; each workgroup keeps its own LIFO stack of node pairs and takes jobs
 * (initial node pairs, src/Corr2.cpp:158-203) by a shared counter, as the reference's
 * OpenMP schedule(dynamic) does.  2: the same launch with a static, equal-count
 * partition of the jobs.  0: chunked launches driven from the host (one read-back per
 * launch).  3: static shares of equal work as the cost model predicts it; 4: jobs
 * handed out by the counter in order of decreasing predicted work (3 and 4 need
 * tc_set_cost_model).  Results do not depend on the mode (npairs identical, sums to
 * rounding).
 * Returns 0, or with mode < 0 the current mode.  No reference counterpart. */
int tc_set_traversal_mode(int mode);
/* Persistent traversals redone with chunked launches because a workgroup's stack would
 * have overflowed (expected 0).  No reference counterpart. */
long long tc_traversal_fallbacks(void);
/* Work counters (as tc_corr2_counters) of the last persistent launch alone, without the
 * chunked split before it; for roofline accounting.  No reference counterpart. */
int tc_persist_counters(long long* counters4);
/* Patch-based covariances (treecorr/corr2base.py:2321-2538): every row of a jackknife,
 * sample, bootstrap or marked-bootstrap design matrix is a weighted sum of patch-pair
 * sums, R = W S (W: nrow x npair row weights, S: npair x nq stacked pair sums, R: nrow x
 * nq; host arrays, row major), replacing the reference's per-row Python sums
 * (_make_cov_design_matrix_core :2321-2368, Corr2._sum :1333-1352); and the covariance of
 * the finished rows V (nrow x ncol), C = sum_r w_r (v_r - mean)^T (v_r - mean) with mean
 * the plain row mean and w null for unit weights (the v.conj().T.dot(v) of :2432-2538).
 * FP64, a fixed summation order per element. */
int tc_cov_rowsums(int nrow, int npair, int nq, const double* W, const double* S, double* R);
int tc_cov_gram(int nrow, int ncol, const double* V, const double* w, double* C);
/* The learned cost model of the persistent traversal's jobs (no reference counterpart:
 * TreeCorr hands out its top-level pairs with OpenMP's schedule(dynamic),
 * src/Corr2.cpp:158-203).  A perceptron of three FP32 layers (nfeatures inputs, two
 * hidden layers of 64 ReLU units, one output) predicts y = log2(1 + items processed) of
 * a job from its two nodes; trained on the GPU (forward, backward and Adam kernels of
 * this library) on the traversal's own job logs.  Traversal modes 3 and 4 use it:
 * static shares of equal predicted work, or jobs handed out in order of decreasing
 * predicted work.  Results never depend on it. */
typedef struct tc_cost_model tc_cost_model;
int tc_cost_model_create(unsigned long long seed, tc_cost_model** out);
int tc_cost_model_destroy(tc_cost_model* m);
int tc_cost_model_sizes(int* nfeatures, int* nparams);
/* params (nparams floats: W1[64][nf], b1[64], W2[64][64], b2[64], W3[64], b3) are read
 * (set = 0) or written (set = 1). */
int tc_cost_model_params(tc_cost_model* m, float* params, int set);
int tc_cost_model_predict(tc_cost_model* m, const float* X, long long n, float* yhat);
/* The gradient of the mean squared error over the n samples (X: n x nfeatures), and
 * that loss. */
int tc_cost_model_grad(tc_cost_model* m, const float* X, const float* y, long long n,
                       float* grad, double* loss);
/* steps Adam steps on mini-batches of `batch` samples drawn from the n; loss: the mean
 * squared error over all n afterwards. */
int tc_cost_model_train(tc_cost_model* m, const float* X, const float* y, long long n, int steps,
                        int batch, float lr, double* loss);
/* The model traversal modes 3 and 4 use (null clears it). */
int tc_set_cost_model(tc_cost_model* m);
/* With max_samples > 0, each following persistent 2-point traversal (NN, KK or GG, the
 * Euclidean metric) counts the items of every job and keeps the features and counts of
 * an evenly strided sample of at most max_samples jobs; 0 stops logging and clears the
 * log.  tc_job_log_fetch copies at most cap samples (X: cap x nfeatures, visits: cap)
 * and sets n to the number kept. */
int tc_set_job_log(long long max_samples);
int tc_job_log_fetch(float* X, float* visits, long long cap, long long* n);
/* Parameters of the persistent traversal.  levels: stack room of each workgroup, in
 * levels of 4 x 256 node pairs (0, the default: the two trees' depths + 4).  min_jobs:
 * fewer initial node pairs than this are first split by chunked launches until there
 * are at least this many jobs (0, the default: 2^26).  Tests set
 * small values to exercise the fallback and the persistent launch on small inputs.
 * No reference counterpart. */
int tc_set_persist_params(int levels, long long min_jobs);
/* A persistent workgroup that runs short takes 1/div of a fair share of the jobs at a
 * time from the shared counter (default 16: the queue depth the autotuner varies).
 * No reference counterpart. */
int tc_set_persist_grab(int div);
/* Graph mode: each persistent launch (its control words cleared, then the kernel) is
 * captured as a hipGraph on first use and replayed whenever a launch has the same
 * arguments again (the arguments of a graph's kernels are fixed at capture; up to 8
 * signatures are kept).  on < 0 returns the current setting.  tc_graph_stats reports the
 * captures and replays so far.  No reference counterpart. */
int tc_set_graph_mode(int on);
int tc_graph_stats(long long* captures, long long* replays);
/* Speculative top tree levels.  The partitions near the root proceed from centroids
 * summed in parallel while the reference-order sums run on a second stream; those
 * then write the node records and check every decision and every object's partition
 * side, and any difference makes the build start again without speculation, so the
 * trees are always the reference's.  mode 0: off; 1 (default): levels whose segments
 * hold 16384 objects or more; 3: every level of at most 64 segments; 2: as 3 with the
 * check forced to fail (2 and 3 are for tests).  Only levels above min_top are
 * speculated on.  When a single tree's speculative decisions pass a validation by
 * rounding bounds (no object within the bound of a split value, no cell near size 0),
 * the build returns while the exact pass still runs: it writes only the records above
 * the top cells and the field's centre and size, which the traversals do not read, and
 * every call that does read them (tc_field_size, tc_field_center, tc_field_copy_nodes,
 * cross-correlation early exits, k-means, the multipole) waits for it.  Returns 0, or
 * with mode -2 the number of such builds so far, with another mode < 0 the number of
 * exact rebuilds so far (nothing changes).  No reference counterpart: TreeCorr builds
 * its trees on the host (src/Cell.cpp:556-615). */
int tc_set_tree_speculation(int mode);
/* The multipole's Gn accumulation in its plain form: every lane of a wavefront adds the
 * multipoles m = 0, 1, ... in order, so lanes with pairs in the same bin contend for one
 * LDS address (the untuned baseline of the benchmark).  on = 0 (default) restores the
 * tuned form, in which each lane starts at its own multipole.  Results agree to the
 * last bits.  Returns 0, or with on < 0 the current setting.  No reference counterpart. */
int tc_set_multipole_plain(int on);
/* Precision of the multipole's zeta contraction for the cells k3 != k2 (the weight and
 * zeta0..7 sums; the diagonal cells, ntri and the mean distances always stay FP64):
 * mode 0 (default) FP64 on the vector ALUs, results within 1e-12 of the reference;
 * 1: FP32 operands on 16x16x4 FP32 MFMAs, FP32 sums per block of 32 centres added in
 * FP64; 2: each FP32 operand split into three BF16 parts, six products per pair on
 * 16x16x32 BF16 MFMAs, the same block sums.  Modes 1 and 2 need nbins <= 32 (else 0 is
 * used); their error against mode 0 is about 1e-7 of each multipole's largest value.
 * Returns 0, or with mode < 0 the current mode.  No reference counterpart. */
int tc_set_zeta_mode(int mode);
int tc_corr2_destroy(tc_corr2* c);

/* ---- 3-point multipole: replace {NNN,GGG}Corr(...) + processAuto ----
 * Reference: BuildCorr3 src/Corr3.cpp:3599-3645 bound at :4205-4226 (the
 * LogMultipole subset: nubins = max_n, u/v bins unused); ProcessAuto
 * :3710-3720 -> BaseCorr3::multipole :1124-1175.  d in {0 (NNN), 4 (GGG)};
 * bin_type must be LogMultipole (5).  Every array has nbins*nbins*(2*max_n+1)
 * entries in TreeCorr's layout ((k2*nbins + k3)*(2*max_n+1) + max_n + n);
 * zeta0..7 are gam0r,gam0i,..,gam3i for GGG and may be NULL for NNN;
 * meand1, meanlogd1, meanu, meanv are accepted and left untouched, as in the
 * reference's multipole path. */
int tc_corr3_create(int d, int bin_type, double minsep, double maxsep, int nbins,
                    double binsize, double b, double a, int max_n, double* zeta0, double* zeta1,
                    double* zeta2, double* zeta3, double* zeta4, double* zeta5, double* zeta6,
                    double* zeta7, double* meand1, double* meanlogd1, double* meand2,
                    double* meanlogd2, double* meand3, double* meanlogd3, double* meanu,
                    double* meanv, double* weight, double* weight_im, double* ntri,
                    tc_corr3** out);
int tc_corr3_process_auto(tc_corr3* c, tc_field* f, int dots, int quick, int metric);
/* Multipole cross-correlations: replace BaseCorr3.processCross12 / processCross21 /
 * processCross (bound at src/Corr3.cpp:4199-4201; ProcessCross12c :3731-3768,
 * ProcessCross21c :3831-3870, ProcessCrossc :3938-3985).  `ordered` is the code
 * treecorr/corr3base.py passes (:1968, 2027, 2086): cross12 0 or 1 (one point from f1,
 * two from f2), cross21 0 or 3 (two from f1, one from f2), cross 0 to 4.  Each runs the
 * reference's multipole passes (multipole(f1, f2) :1177-1236, multipole(f1, f2, f3,
 * ordered) :1238-1305) and ADDS to the arrays, as processAuto does.  Every field must
 * have the correlation's type and one coordinate system. */
int tc_corr3_process_cross12(tc_corr3* c, tc_field* f1, tc_field* f2, int ordered, int dots,
                             int quick, int metric);
int tc_corr3_process_cross21(tc_corr3* c, tc_field* f1, tc_field* f2, int ordered, int dots,
                             int quick, int metric);
int tc_corr3_process_cross(tc_corr3* c, tc_field* f1, tc_field* f2, tc_field* f3, int ordered,
                           int dots, int quick, int metric);
/* Work counters of the last process call: Gn calls, zeta calls, zeta rows,
 * multipoleFinish stages (no reference counterpart; roofline accounting). */
int tc_corr3_counters(const tc_corr3* c, long long* counters4);
/* Replaces TriviallyZeroCorr3 (src/Corr3.cpp:4069-4130, BaseCorr3::triviallyZero
 * :275-299; Euclidean metric): *result = 1 when no triangle with a vertex in each of
 * three cells (centre x, y, z and size s; z ignored for Flat) can fall in a bin.
 * ordered is the reference's ordered code (0 = any vertex order); p13 != 0 when
 * cell 2 is cell 1 or cell 3 (only the 1-3 distance is tested). */
int tc_corr3_trivially_zero(tc_corr3* c, int coords, double x1, double y1, double z1,
                            double s1, double x2, double y2, double z2, double s2, double x3,
                            double y3, double z3, double s3, int ordered, int p13,
                            int* result);
/* Zeta cells of the last process call: the (k2, k3 >= k2) entries that its zeta
 * contractions added to, summed over the contractions (the unit of the zeta work,
 * times 2*max_n+1 multipoles).  No reference counterpart. */
int tc_corr3_zeta_cells(const tc_corr3* c, long long* cells);
int tc_corr3_destroy(tc_corr3* c);

/* ---- Host utilities ---- */
/* GenerateXYZ, src/KMeans.cpp:1013-1050 (glibc sincos; r may be NULL). */
int tc_generate_xyz(double* x, double* y, double* z, const double* ra, const double* dec,
                    const double* r, long long n);
/* SetOMPThreads / GetOMPThreads, src/Corr2.cpp:995-1003: accepted for API
 * compatibility; the GPU engine ignores the host thread count. */
int tc_set_num_threads(int n);
int tc_get_num_threads(int* n);

/* Kernel timing with HIP events on the library stream (bench/profiling only;
 * no reference counterpart).  which: 0 = 2-pt traversal (all its launches), 1 = tree
 * build, 2 = multipole Gn, 3 = zeta contraction, 4 = the persistent traversal launches
 * alone. */
int tc_set_timing(int on);
/* Streams n doubles from src to dst (device pointers) with a plain copy kernel;
 * used only to calibrate HBM traffic counters (FETCH_SIZE/WRITE_SIZE). */
int tc_copy_calibration(const double* src, double* dst, long long n);
int tc_kernel_stats(int which, double* total_ms, long long* launches);
int tc_reset_kernel_stats(void);

#ifdef __cplusplus
}
#endif
#endif /* TREECORR_AMD_H */
