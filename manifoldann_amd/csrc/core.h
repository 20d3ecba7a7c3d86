// Device-resident operands and the search pipelines shared by the C ABI.
#pragma once

#include <map>
#include <vector>

#include "runtime.h"
#include "util.h"

struct mann_dataset {
  int device = 0;
  int64_t d = 0, n = 0, d16 = 0, kp = 0, n_pad = 0;
  int normalized = 0;            // bf16 copy holds unit-normalised rows (cosine)
  float* x32 = nullptr;          // n_pad x d16 FP32 rows (exact operand)
  mann::bf16_t* xb = nullptr;    // n_pad x kp bf16 rows (MFMA operand)
  float* xn = nullptr;           // |x~|^2, +inf on padding
  float* err = nullptr;          // |x - x~| per row (rounded up)
  float* nrm = nullptr;          // |x~| per row (rounded up)
  float* xn2 = nullptr;          // A16 |x|^2 per row
  float* center = nullptr;       // d floats subtracted before the bf16 rounding (or null)
  float ex = 0.0f, rx = 0.0f;    // max err, max nrm
  // stride sample for threshold estimation
  int m = 0, m_pad = 0;
  mann::bf16_t* sb = nullptr;
  float* sxn = nullptr;
  // the sample as a query operand (folded form; null when not folded: sb)
  mann::bf16_t* sbq = nullptr;
  // columns d..d+2 of xb (and sb) hold |x~|^2 as three bf16 parts summing to
  // it exactly, d+3..d+5 hold 1 (first pass with folded norms); qfold: the
  // rows as folded query operands (-2 x~, 1, 1, 1, parts of |x~|^2) for
  // searches of the dataset against itself
  int fold = 0;
  mann::bf16_t* qfold = nullptr;
  // FP8 (E4M3) first pass, when selected: rows x^ = e4m3(t 2^sexp8) of kp8
  // bytes (t: the vectors of the bf16 copy before rounding), |x~|^2 (+inf on
  // padding), |t - x~| and |x~| rounded up, and their maxima
  int first_pass = 0;            // 0 bf16, 1 FP8
  int64_t kp8 = 0;
  int sexp8 = 0;
  unsigned char* x8 = nullptr;
  float* xn8 = nullptr;
  float* err8 = nullptr;
  float* nrm8 = nullptr;
  float ex8 = 0.0f, rx8 = 0.0f;
  ~mann_dataset();
};

// IVF index handle: list-contiguous copy of the base plus routing operands.
struct mann_ivf {
  int device = 0;
  int64_t d = 0, n = 0, d16 = 0, kp = 0, n_pad = 0;
  int nlist = 0;   // routing distances per query (centroids)
  int64_t max_list = 0;  // rows of the largest list
  int nb = 0;      // buckets = max assignment (partition_by_transform)
  int flat = 0;    // IVF-Flat: nlist lists, bounded-heap merge (ivf_flat.jl:138-167)
  int normalized = 0;
  // routing metric of the centroids: EUCLIDEAN (KMeansTransform's norms form,
  // kmeans/distance.jl:213-242), or for IVF-Flat its centroid_metric:
  // SQEUCLIDEAN (no square root) or COSINE (ivf_flat.jl:103-131)
  int route_metric = 0;
  float* c32 = nullptr;
  float* cn2 = nullptr;
  // bf16 first-pass operands of the first min(nb, nlist) centroids for the
  // certified TopKRouting (null when it does not apply: cosine rows or
  // centroids, or more centroids than one wavefront selects from): rows
  // padded to r_pad with |c~|^2 = +inf, the maxima of |c - c~| and |c~|, and
  // an upper bound of the centroids' norms
  mann::bf16_t* rb = nullptr;
  float* rn = nullptr;
  int64_t r_pad = 0;
  float rex = 0.0f, rrx = 0.0f, rra = 0.0f;
  // residual centres of the bf16 rows: the centroids, or for cosine (unit
  // rows) the unit centroids c / |c|
  float* cres = nullptr;
  int* list_off = nullptr;  // nb + 1
  int* perm = nullptr;      // reordered row -> point id (0-based)
  int* bucket_of = nullptr; // point id (0-based) -> bucket
  int* lorder = nullptr;    // lists by size, largest first
  std::vector<int> h_off;
  float* x32 = nullptr;
  mann::bf16_t* xb = nullptr;
  float* xn = nullptr;
  float* xn2 = nullptr;
  float ex = 0.0f, rx = 0.0f;
  // HNSW terminals (mann_ivf_hnsw_build): one graph per list over the
  // list-contiguous rows; base-layer lists hn_nbr0[row][M + 1] and hn_deg0[row],
  // upper layers in compact slots of the hn_n1 rows of level >= 1
  // (hn_nbru[(layer - 1) n1 + hn_upslot[row]][M + 1], hn_degu likewise); ids in
  // the lists are local to the list
  int hn_M = 0, hn_layers = 0, hn_policy = 1, hn_ef_c = 0;
  int64_t hn_n1 = 0;
  int* hn_nbr0 = nullptr;
  int* hn_deg0 = nullptr;
  int* hn_upslot = nullptr;
  int* hn_nbru = nullptr;
  int* hn_degu = nullptr;
  int* hn_level = nullptr;
  int* hn_entry = nullptr;   // per bucket, local id (-1: empty)
  int* hn_maxl = nullptr;
  void free_hnsw() {
    void* hs[] = {hn_nbr0, hn_deg0, hn_upslot, hn_nbru, hn_degu, hn_level, hn_entry, hn_maxl};
    for (void* p : hs)
      if (p) (void)hipFree(p);
    hn_nbr0 = hn_deg0 = hn_upslot = hn_nbru = hn_degu = hn_level = hn_entry = hn_maxl = nullptr;
    hn_M = hn_layers = hn_ef_c = 0;
    hn_n1 = 0;
  }
  ~mann_ivf() {
    (void)hipSetDevice(device);
    free_hnsw();
    void* ps[] = {c32, cn2, cres != c32 ? cres : nullptr, list_off, perm, bucket_of, lorder,
                  x32, xb, xn, xn2, rb, rn};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
};

// Learned probe router handle: FP32 master parameters + Adam moments.
struct mann_router {
  int device = 0;
  int d = 0, h = 0, c = 0;
  int step = 0;
  float* w = nullptr;   // all parameters: W1 (h x d), b1 (h), W2 (c x h), b2 (c)
  float* m1 = nullptr;  // Adam moments
  float* m2 = nullptr;
  int64_t nparam = 0;
  // optional centroid prior: logits += gamma * (2 x.c - |c|^2) (c x pld floats)
  float* pc = nullptr;
  float* pcn = nullptr;
  int64_t pld = 0;
  float gamma = 0.0f;
  // the same prior as a GEMM: logits += x.(2 gamma c) + (-gamma |c|^2)
  float* pc2 = nullptr;
  float* pb = nullptr;
  float* W1() const { return w; }
  float* b1() const { return w + (int64_t)h * d; }
  float* W2() const { return b1() + h; }
  float* b2() const { return W2() + (int64_t)c * h; }
  ~mann_router() {
    (void)hipSetDevice(device);
    if (w) (void)hipFree(w);
    if (m1) (void)hipFree(m1);
    if (m2) (void)hipFree(m2);
    if (pc) (void)hipFree(pc);
    if (pcn) (void)hipFree(pcn);
    if (pc2) (void)hipFree(pc2);
    if (pb) (void)hipFree(pb);
  }
};

namespace mann {

// Query operands (scratch of the calling thread, or views of a dataset).
struct QueryOps {
  int64_t nq = 0;
  const float* q32 = nullptr;
  const bf16_t* qb = nullptr;
  const float* qn = nullptr;
  const float* qe = nullptr;
  const float* qr = nullptr;
  const float* qn2 = nullptr;
};

// Upload (if host) and convert a query matrix; scratch of this thread.
QueryOps prepare_queries(const float* q, int64_t d, int64_t nq, int64_t ldq, int normalize,
                         int d16, int kp, const char* tag, const float* center = nullptr);
QueryOps dataset_as_queries(const mann_dataset& ds);

mann_dataset* create_dataset(const float* data, int64_t d, int64_t n, int64_t ld, int metric);
// select the first-pass arithmetic of the exact search: 0 bf16 (frees the FP8
// copy), 1 FP8 E4M3 (builds it)
void dataset_set_first_pass(mann_dataset& ds, int kind);

struct SearchStats {
  long long reranked = 0, overflow = 0, pass2 = 0, exact = 0;
  long long probes = 0;  // IVF: (list, query) pairs scanned
  long long emitted = 0; // first-pass candidates read by the selection (all passes)
  long long heap_replays = 0;  // IVF-Flat queries merged through the bounded heap
  long long hnsw_distances = 0;  // IVF-HNSW: exact distances the graph walks computed
};
SearchStats& last_stats();
// adds the selection counters of a kStatSlots x 4 host copy to s
void add_stats(const int* slots, SearchStats& s);

// Exact k nearest rows of `ds` for each query (bf16 MFMA first pass, certified
// FP32 re-rank, refinement pass, exact fallback).  out_* are device pointers,
// nq x k.  exclude_self drops the row whose id equals the query index.
// Exact fallback for the nflag queries listed in e.qlist (count also in
// *e.nq_dev): segmented exact top-k over their row ranges (at most total_rows
// rows in at most nranges + 1 ranges per query).
// Returns the arguments as launched (partials and segment count set).
ExactArgs run_exact_fallback(ExactArgs e, int nflag, int64_t total_rows, int nranges);

// exact search for k + exclude_self above the selection's capacity: every
// distance, a segmented sort per query (core_bf_all.hip)
void bf_search_all(const mann_dataset& ds, const QueryOps& Q, int k, int metric, int exclude_self,
                   int64_t* out_ids, float* out_dist, int* out_count);
void bf_search(const mann_dataset& ds, const QueryOps& Q, int k, int metric, int exclude_self,
               int64_t* out_ids, float* out_dist, int* out_count);

// Knobs chosen by the autotuner, per search shape (under the tuning mutex).
struct ShapeKey {
  int64_t path = 0;       // 0 exact search, 1 IVF search
  int64_t device = 0, metric = 0, first_pass = 0, self_join = 0, k = 0;
  int64_t n = 0, d = 0;
  int64_t nq_bucket = 0;  // query count rounded up to a power of two
  int64_t nprobe = 0;     // IVF: lists probed per query
  int64_t lists = 0;      // IVF: lists of the index
  bool operator<(const ShapeKey& o) const;
};
ShapeKey bf_shape_key(const mann_dataset& ds, int64_t nq, int k_req, int metric, bool self_join);
void shape_tuning_put(const ShapeKey& key, const Tuning& t);
bool shape_tuning_apply(const ShapeKey& key, Tuning& t);  // false: no entry
void shape_tuning_clear();
bool& shape_tuning_disabled();  // this thread: ignore the shape entries
// Replaces the calling thread's knobs for one scope.
struct TuningScope {
  Tuning saved;
  bool active = false;
  void set(const Tuning& t) {
    if (!active) saved = tuning();
    tuning() = t;
    active = true;
  }
  ~TuningScope() {
    if (active) tuning() = saved;
  }
};
struct AutotuneResult {
  Tuning best;
  double ms = 0.0;
  int trials = 0;
};
AutotuneResult autotune_bf(const mann_dataset& ds, const QueryOps& Q, int k, int metric,
                           int exclude_self, int64_t* ids, float* dist);
// IVF search (core_ivf.hip): the list-scan tile configuration, the rows of a
// work item, the first-list threshold tile and the shortlist, per shape
ShapeKey ivf_shape_key(const mann_ivf& iv, int64_t nq, int k, int nprobe, int metric);
AutotuneResult autotune_ivf(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int metric,
                            const float* route, int64_t* ids, float* dist);

// k-means (core_kmeans.hip).  C is k x d16 device FP32; assign 0-based.
void kmeans_estep(const mann_dataset& ds, const float* C, int k, int* assign, float* mindist,
                  bool primed, long long* overflow_out);
int64_t kmeans_lloyd(const mann_dataset& ds, float* C, int k, int64_t max_iters, double tol,
                     int* assign, float* mindist, long long* overflow);
void kmeans_pp(const mann_dataset& ds, int k, uint64_t seed, float* C, int* picked);
std::vector<int> random_indices(int64_t n, int k, uint64_t seed);

// IVF (core_ivf.hip)
// _symmetrize_adjacency on device arrays: adj (n x k, 1-based) -> CSR offsets
// (n + 1) and ids (capacity 2 n k); *total edges, *max_degree the longest row
void knn_graph_symmetrize_gpu(const int64_t* adj, int64_t n, int64_t k, int64_t* offsets,
                              int64_t* ids, int64_t* total, int64_t* max_degree);
mann_ivf* ivf_build(const mann_dataset& ds, const float* C, int nlist, const int64_t* assign_host,
                    bool all_lists = false, int route_metric = 0);
void ivf_route(const mann_ivf& iv, const QueryOps& Q, float* R);
// mass > 0: probability-mass probing (R = negated logits), at most nprobe lists
void ivf_probe(const mann_ivf& iv, const float* R, int nq, int nprobe, int* probes, int* pcount,
               int* prows, float mass = 0.0f);
// merge: 0 SimpleMerge, 1 DisjointMerge; probe_order: per-list results in
// probe order (the single-query path) instead of bucket order (the batch).
// An IVF-Flat handle always merges through its bounded heap in probe order.
void ivf_search(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int metric,
                const float* route_in, int64_t* out_ids, float* out_dist, float mass = 0.0f,
                int merge = 0, int probe_order = 0);
// IVF-HNSW (core_hnsw.hip): build the HNSW terminals of every list (policy 0
// heuristic, 1 diversified; levels from the splitmix64 stream of seed over the
// list-contiguous rows), and search them (route_in: optional routing
// distances; merge 0 SimpleMerge, 1 DisjointMerge; probe_order as ivf_search)
void ivf_hnsw_build(mann_ivf& iv, int M, int ef_construction, int policy, uint64_t seed);
// install graphs built elsewhere (dense [layers][n][M + 1] / [layers][n] host
// arrays over the list-contiguous rows), validated so that every access the
// kernels make stays inside the handle; level null: inferred
void ivf_hnsw_set_graph(mann_ivf& iv, int layers, int M, const int* nbr, const int* deg,
                        const int* entry, const int* maxl, const int* level);
void ivf_hnsw_search(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int ef_search,
                     const float* route_in, int merge, int probe_order, int64_t* out_ids,
                     float* out_dist);

// learned router (core_router.hip)
mann_router* router_create(int d, int h, int c, uint64_t seed);
// attach (gamma != 0) or drop (gamma == 0) the centroid prior of an IVF index
void router_set_prior(mann_router& r, const mann_ivf* iv, float gamma);
// negate: out = -scores (the routing distances of a routed search)
void router_scores(const mann_router& r, const float* X, int64_t ldx, int64_t n, float* out,
                   bool negate = false);
void router_grad(const mann_router& r, const float* X, int64_t ldx, int B, const int* labels,
                 int L, float* grad, float* loss);
double router_train(mann_router& r, const float* X, int64_t ldx, int64_t n, const int* labels,
                    int L, int epochs, int batch, float lr, uint64_t seed);

}  // namespace mann
