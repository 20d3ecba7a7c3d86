/*
 * mann.h -- C ABI of the MI355X (gfx950) search core of ManifoldANN.jl.
 *
 * Every entry point replaces one reference Julia method on the data-parallel
 * search path (citations are ManifoldANN.jl source files).  Conventions:
 *  - A matrix argument is a Julia column-major d x n Float32 matrix: n points,
 *    each d contiguous floats, point stride `ld` (>= d) in elements.
 *  - Pointers may be host memory (staged through the call) or device memory
 *    of the current HIP device (used in place); the library checks which.
 *    Device memory is read on the calling thread's own non-blocking stream:
 *    the work that produced an input must be complete when the call is made
 *    (synchronize the producing stream first).  Outputs are complete when
 *    the call returns.
 *  - Ids are 1-based int64 like Julia's Neighbor.id; 0 marks an empty slot.
 *    Per-query results are k consecutive slots (a k x nq column-major matrix,
 *    the layout of benchmarking/wrappers/manifoldann.py:57-70), ascending
 *    distance, ties to the lower id (Julia partialsortperm, bruteforce.jl:65).
 *  - Return value: 0 on success, or a negative status whose class matches the
 *    Julia exception the reference throws at that point; mann_last_error()
 *    gives the message (thread-local).
 *  - Calls are re-entrant: each host thread has its own stream and scratch.
 */
#ifndef MANN_H
#define MANN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MANN_OK 0
#define MANN_ERR_DIMENSION_MISMATCH (-1) /* Julia DimensionMismatch */
#define MANN_ERR_ARGUMENT (-2)           /* Julia ArgumentError */
#define MANN_ERR_ASSERTION (-3)          /* Julia AssertionError */
#define MANN_ERR_DEVICE (-4)             /* HIP runtime failure */
#define MANN_ERR_ALLOC (-5)              /* device allocation failure */
#define MANN_ERR_NULL (-6)               /* null handle or pointer */
#define MANN_ERR_UNSUPPORTED (-7)

/* Metrics: Distances.Euclidean / SqEuclidean / CosineDist
 * (src/utils/distances.jl:15-18). */
#define MANN_EUCLIDEAN 0
#define MANN_SQEUCLIDEAN 1
#define MANN_COSINE 2

typedef struct mann_dataset* mann_dataset_t;
typedef struct mann_ivf* mann_ivf_t;
typedef struct mann_router* mann_router_t;
typedef struct mann_graph* mann_graph_t;

/* Library status. */
int mann_version(void);
const char* mann_last_error(void);
int mann_device_count(void);

/* A device-resident copy of a data matrix (the `data` argument every
 * reference call receives, ann_index.jl:1-10; the reference never stores it,
 * so this handle is explicit and owned by the caller).  `metric` selects the
 * first-pass representation (COSINE keeps unit-normalised bf16 rows). */
int mann_dataset_create(const float* data, int64_t d, int64_t n, int64_t ld, int metric,
                        mann_dataset_t* out);
/* The same from IEEE binary16 data (a Matrix{Float16} base, as in the large
 * cosine search over 10M x 256 Float16 rows):
 * the rows are upcast exactly to Float32 on the device.  The reference would
 * compute in Float16 (src/indices/bruteforce.jl:53, S = float(T)); this
 * library computes in Float32 on the upcast values (the documented oracle of
 * that configuration). */
int mann_dataset_create_f16(const void* data, int64_t d, int64_t n, int64_t ld, int metric,
                            mann_dataset_t* out);
int mann_dataset_destroy(mann_dataset_t ds);
int mann_dataset_info(mann_dataset_t ds, int64_t* d, int64_t* n);

/* First-pass arithmetic of the exact searches over `ds` (mann_bf_knn,
 * mann_knn_graph, mann_knn_candidates).  No counterpart in the reference,
 * whose searches are FP32 throughout: results are the same either way (the
 * FP32 re-rank is certified against the bounds of the pass used); only the
 * speed differs.  MANN_FIRST_PASS_BF16 (default) frees the FP8 copy;
 * MANN_FIRST_PASS_FP8 builds an OCP E4M3 copy of the rows (one byte per
 * element, d rounded up to 128) with a power-of-two scale, scanned by the
 * gfx950 FP8 MFMA; the FP8 survivors are re-scored on bf16 before the
 * re-rank.  The allocation-free entry point always uses bf16. */
#define MANN_FIRST_PASS_BF16 0
#define MANN_FIRST_PASS_FP8 1
int mann_dataset_set_first_pass(mann_dataset_t ds, int kind);

/* query(::BruteForceIndex, data, Q, k) -- batch form (src/ann_index.jl:71-101
 * over src/indices/bruteforce.jl:46-71, validation src/utils/validation.jl:8-34).
 * k_eff = min(k, n); slots past k_eff are empty.  nq = 0 is valid.  Any k:
 * up to 256 (255 for the self-joins of mann_knn_graph / mann_knn_candidates)
 * the certified selection runs; above, every exact distance is sorted per
 * query (same results, one full pass and sort per query). */
int mann_bf_knn(mann_dataset_t base, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                int64_t k, int metric, int64_t* out_ids, float* out_dist);

/* mann_bf_knn without allocation -- no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks and results as mann_bf_knn (bit-identical ids and distances), but
 * queries, out_ids and out_dist must be device memory, and every scratch
 * buffer of the search comes from one 64 MiB array that the library keeps per
 * GPU for the life of the process: a batch whose scratch would exceed it is
 * an ArgumentError (the message names the bytes needed; split the batch), and
 * so is k > 256 (the all-distances sort's scratch grows with n).
 * The first call on a host thread and device does one-time set-up (the
 * thread's stream, the array's device address); no call allocates after that.
 * Calls on one GPU share the array: the library runs them one at a time, and
 * work that overlaps them on that GPU by other means races. */
int mann_bf_knn_noalloc(mann_dataset_t base, const float* queries, int64_t d, int64_t nq,
                        int64_t ldq, int64_t k, int metric, int64_t* out_ids, float* out_dist);

/* An exact search (mann_bf_knn) captured in a HIP graph, for serving loops
 * that repeat one batch shape.  No counterpart in the reference.  queries,
 * out_ids and out_dist must be device memory; the graph reads the queries and
 * writes the outputs at those addresses on every launch, so the caller keeps
 * them (and `base`) alive and refills the queries between launches.  Creation
 * runs the search once eagerly (its answer is written), then captures it with
 * all scratch in one allocation the graph owns.  mann_graph_launch replays it
 * on the calling thread's stream and waits; flagged queries beyond the 64 the
 * graph's exact fallback covers are finished eagerly, so every launch returns
 * the exact answer.  Statistics of the launch are in mann_last_stats. */
int mann_bf_knn_graph_create(mann_dataset_t base, const float* queries, int64_t d, int64_t nq,
                             int64_t ldq, int64_t k, int metric, int64_t* out_ids,
                             float* out_dist, mann_graph_t* out);
int mann_graph_launch(mann_graph_t g);
int mann_graph_destroy(mann_graph_t g);

/* The batched MultiLevelIndex search (mann_ivf_search: routing, TopKRouting
 * probes, the persistent list scan, the certified selection in the batched
 * SimpleMerge order of query.jl:279-399, merge.jl:79-124) captured in a HIP
 * graph for one batch shape.  No counterpart in the reference.  Every step
 * runs device-gated: the refinement pass and its selection cover whatever the
 * device flags, the exact fallback its first 64 queries; mann_graph_launch
 * finishes any further fallback queries eagerly, so every launch returns the
 * exact answer.  queries (d x nq, stride ldq), route_dist (optional nq x
 * nlist routing distances, read at every launch, e.g. a router's scores) and
 * the outputs must be device memory the caller keeps alive.  Creation runs the
 * device-gated search once to size the graph's own scratch buffers, then
 * captures it; the outputs are defined after the first mann_graph_launch.
 * IVF-Flat handles (heap merge) are rejected with ArgumentError. */
int mann_ivf_search_graph_create(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq,
                                 int64_t ldq, int64_t k, int64_t nprobe, int metric,
                                 const float* route_dist, int64_t* out_ids, float* out_dist,
                                 mann_graph_t* out);

/* Autotuner of the exact search (no counterpart in the reference): times the
 * tile configuration, the emission target and the shortlist (for an FP8
 * dataset its tile configuration and emission target) on this batch by
 * coordinate descent, and keeps the best for the shape (device, dataset size
 * and dimension, metric, first pass, k, query count rounded up to a power of
 * two; queries == NULL tunes the kNN graph's self-join).  Later searches of
 * the same shape use it; mann_tune_set keeps setting the defaults of the
 * others.  out_choice (optional, 6 values): scan_cfg, emit_per_k,
 * shortlist_mult, scan8_cfg, fp8_emit_per_k, searches timed; out_ms: the best
 * search time.  mann_autotune_clear forgets every shape. */
int mann_autotune_bf_knn(mann_dataset_t base, const float* queries, int64_t d, int64_t nq,
                         int64_t ldq, int64_t k, int metric, int64_t* out_choice,
                         double* out_ms);
/* Autotuner of the batched MultiLevelIndex search (no counterpart in the
 * reference): times the list-scan tile configuration, the rows of a work item
 * (the persistent scan's work partition), the first-list threshold tile and
 * the shortlist on the caller's batch, and keeps the fastest for this shape
 * (index, k, nprobe, metric, query-count bucket); later searches of the shape
 * use it.  route_dist as in mann_ivf_search_ex.  out_choice (5 entries):
 * ivf_cfg, ivf_rows_item, ivf_tau_cfg, shortlist_mult, trials; out_ms: the
 * best search time.  Results are exact whatever the knobs; mann_autotune_clear
 * forgets these shapes too. */
int mann_autotune_ivf(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                      int64_t k, int64_t nprobe, int metric, const float* route_dist,
                      int64_t* out_choice, double* out_ms);
int mann_autotune_clear(void);

/* build_knn_graph(index, data; k, include_self) -- directed adjacency
 * (src/graphs/knn_graph.jl:42-81, _prepare_neighbors :105-130): each point is
 * queried with k+1 (k if include_self), itself dropped by id, first k kept.
 * out_ids is k x n.  MANN_ERR_ARGUMENT unless 0 < k <= n - !include_self. */
int mann_knn_graph(mann_dataset_t data, int64_t k, int include_self, int metric,
                   int64_t* out_ids, float* out_dist);

/* The candidate re-query of build_weighted_graph with candidate_k > k
 * (src/graphs/weighted_knn_graph.jl:497-503, :549-553): every point queried
 * with candidate_k + 1, its own id dropped, the first candidate_k kept.  Unlike
 * build_knn_graph a short row is not an error: a point gets min(candidate_k,
 * n - 1) candidates (fewer only if it is not among its own first
 * candidate_k + 1), and empty slots hold id 0.  out_ids (and out_dist when
 * given) are candidate_k x n; out_count (optional) the candidates per point.
 * Any candidate_k (above 255 through the all-distances sort of mann_bf_knn). */
int mann_knn_candidates(mann_dataset_t data, int64_t candidate_k, int metric, int64_t* out_ids,
                        float* out_dist, int64_t* out_count);

/* _symmetrize_adjacency (src/graphs/knn_graph.jl:196-210): union of i->j and
 * j->i, each row sorted ascending, as CSR (out_offsets n+1, out_ids 2*n*k
 * capacity).  *out_total receives the edge count; *out_max_degree the new
 * graph.k (:74).  Runs on the GPU (radix sort of both edge directions, then
 * duplicate removal); adj and the two output arrays are all host or all
 * device memory.  A neighbour id outside [1, n] is an ArgumentError. */
int mann_knn_graph_symmetrize(const int64_t* adj, int64_t n, int64_t k, int64_t* out_offsets,
                              int64_t* out_ids, int64_t* out_total, int64_t* out_max_degree);

/* assign_clusters!(pairwise_euclidean!(X, C)) -- the k-means E-step
 * (src/transforms/kmeans/distance.jl:68-107, 287-310): nearest centroid by the
 * FP32 norms form sqrt(max((|c|^2 + |x|^2) - 2 c.x, 0)), ties to the lower
 * centroid.  centroids is a d x k matrix; out_assign (n) is 1-based;
 * out_mindist (optional) receives the distance to the assigned centroid. */
int mann_kmeans_assign(mann_dataset_t X, const float* centroids, int64_t k, int64_t* out_assign,
                       float* out_mindist);

/* init_kmeans_plus_plus (src/transforms/kmeans/init.jl:56-128) with draws from
 * the library's splitmix64 stream of `seed` (Julia's RNG streams cannot be
 * reproduced): u_0 picks floor(u_0 n); round i picks the first point whose
 * blocked cumulative D^2 reaches u_i * total.  out_idx (optional) 1-based. */
int mann_kmeanspp(mann_dataset_t X, int64_t k, uint64_t seed, float* out_centroids,
                  int64_t* out_idx);

/* fit!(::KMeansTransform, X; rng) -> lloyd! (src/transforms/KMeansTransform.jl:
 * 106-162, src/transforms/kmeans/lloyd.jl:36-134): init 0 = the given
 * init_centroids (d x k), 1 = kmeans++ (seed), 2 = random distinct points
 * (seed).  M-step sums in FP64 over a fixed order; empty clusters take the
 * farthest point (lloyd.jl:81-104); stops when the max centroid movement < tol.
 * subsample (KMeansTransform's subsample_size; 0 = nothing): when 0 < subsample
 * < n, Lloyd runs on max(subsample, k) distinct points (the first of a
 * splitmix64 Fisher-Yates shuffle of `seed`, in that order, as
 * X[:, randperm(rng, n)[1:m]], :113-120), the initialisation then drawing from
 * the stream of seed * 0x9E3779B97F4A7C15 + 1, and out_assign (n) comes from one
 * E-step over all points (:137-146).
 * Errors: MANN_ERR_ASSERTION for k <= 0, k > n, max_iters <= 0, tol <= 0. */
int mann_kmeans_fit(mann_dataset_t X, int64_t k, int64_t max_iters, double tol, int init,
                    uint64_t seed, const float* init_centroids, int64_t subsample,
                    float* out_centroids, int64_t* out_assign, int64_t* out_iters);

/* transform(::KMeansTransform, x) for a batch (src/transforms/KMeansTransform.jl:
 * 181-190 -> src/transforms/kmeans/distance.jl:213-242): out_dist[q][c] =
 * sqrt(max((|c|^2 + |q|^2) - 2 c.q, 0)) in FP32 (A16 order); centroids d x nlist. */
int mann_route_distances(const float* centroids, int64_t nlist, int64_t d, const float* queries,
                         int64_t nq, int64_t ldq, float* out_dist);

/* build_index(MultiLevelIndex, X, TransformedConfig(KMeansTransform, TopKRouting,
 * TerminalConfig(BruteForceIndex))) after the k-means fit
 * (src/indices/multilevel/builder.jl:61-218, partition_by_transform
 * src/transforms/utils.jl:75-174): buckets = max(assign); each bucket's points
 * in ascending id; the base is copied list-contiguously.  assign is 1-based. */
int mann_ivf_build(mann_dataset_t X, const float* centroids, int64_t nlist,
                   const int64_t* assign, mann_ivf_t* out);
/* build_index(IVFFlatIndex, data; nlist, ...) after the k-means fit
 * (src/indices/ivf_flat.jl:30-93): one list per centroid, nlist lists, empty
 * ones included (:66-77), each in ascending id.  Queries go through
 * mann_ivf_search with nprobe clamped to [1, nlist] (:146); the lists probed
 * are the nprobe nearest centroids, empty ones taking their slots (:149-150).
 * mann_ivf_search on such a handle merges as the reference's bounded max-heap
 * (:158-167): the GPU takes the (dist, id) top-(k+1); a query whose k-th and
 * (k+1)-th distances tie replays its probed lists through the heap on the GPU,
 * so ties at the k-th place resolve as the reference's do.  k <= 255.
 * centroid_metric (IVFFlatIndex's centroid_metric, ivf_flat.jl:103-131):
 * MANN_EUCLIDEAN / MANN_SQEUCLIDEAN (the norms form, with and without the
 * square root) or MANN_COSINE (CosineDist of centroid and query).  The point
 * metric is the dataset's (a MANN_COSINE dataset serves CosineDist).
 * assign is 1-based and must lie in [1, nlist]. */
int mann_ivf_flat_build(mann_dataset_t X, const float* centroids, int64_t nlist,
                        const int64_t* assign, int centroid_metric, mann_ivf_t* out);
int mann_ivf_destroy(mann_ivf_t ivf);
/* number of buckets and (optional) the size of each */
int mann_ivf_info(mann_ivf_t ivf, int64_t* n_buckets, int64_t* list_sizes);
/* the point ids of the lists, concatenated in bucket order (1-based) */
int mann_ivf_list_ids(mann_ivf_t ivf, int64_t* out_ids);
/* routing distances of the index's centroids (as mann_route_distances) */
int mann_ivf_route(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                   float* out_dist);
/* select_indices(TopKRouting(nprobe), KMeansAssignment(route_dist[q]), buckets)
 * followed by _resolve_child_positions (src/indices/multilevel/routing.jl:94-118,
 * query.jl:108-121): 1-based buckets in probe order, empty buckets dropped,
 * 0-padded to nprobe. */
int mann_ivf_probe(mann_ivf_t ivf, const float* route_dist, int64_t nq, int64_t nprobe,
                   int64_t* out_buckets);
/* query(::MultiLevelIndex, data, Q, k) -- batched (src/indices/multilevel/
 * query.jl:279-512 with SimpleMerge, merge.jl:79-124): exact top-k over the
 * probed lists, ties by (distance, bucket, id) as the batched per-child merge
 * produces.  route_dist (optional, nq x nlist, smaller = probe first) replaces
 * the centroid distances (a learned router passes its negated scores). */
int mann_ivf_search(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                    int64_t k, int64_t nprobe, int metric, const float* route_dist,
                    int64_t* out_ids, float* out_dist);

/* The same with the index's merge strategy and query form (query.jl:25-39 for
 * one query, :279-300 for a batch; merge.jl:79-169):
 *  merge = MANN_MERGE_SIMPLE (SimpleMerge) or MANN_MERGE_DISJOINT
 *  (DisjointMerge: each probed list's top-min(k, |list|) by (dist, id) pushed
 *  through the bounded max-heap of neighbor_heaps.jl:84-100, ties at the k-th
 *  place resolved as that heap resolves them, output by (dist, id));
 *  probe_order = 1 concatenates the per-list results in probe order, as the
 *  single-query path does (_query_recursive, query.jl:66-106): SimpleMerge then
 *  orders ties by (dist, probe rank, id); 0 uses the batched child order
 *  (dist, bucket, id).
 * On an IVF-Flat handle both are ignored (its query is the bounded heap over
 * every probed row in probe order).  k <= 255 when the merge is not the
 * batched SimpleMerge. */
#define MANN_MERGE_SIMPLE 0
#define MANN_MERGE_DISJOINT 1
int mann_ivf_search_ex(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                       int64_t k, int64_t nprobe, int metric, const float* route_dist, int merge,
                       int probe_order, int64_t* out_ids, float* out_dist);

/* build_ivf_hnsw_index's terminals (src/indices/multilevel/ivf_hnsw.jl:46-88:
 * TerminalConfig(HNSWIndex, (M, ef_construction, ef_search, neighbor_policy)),
 * built per child by build_index(HNSWIndex, child data) src/indices/hnsw/
 * index.jl:10-67): one HNSW graph per non-empty list of an IVF handle
 * (mann_ivf_build), over the list's points in ascending id, by the serial
 * insertion of src/indices/hnsw/query.jl:264-325 (threaded = false; the
 * threaded build is not deterministic).  neighbor_policy: 0 :heuristic, 1
 * :diversified (neighbor_policy.jl:22-112).  Levels: the reference's children
 * draw from Random.default_rng(); here level = floor(-log(max(u, eps)) /
 * log(max(M, 2))) (planner.jl:19-30) over the splitmix64 stream of `seed`, one
 * draw per point in list order (bucket order, ascending id).  Euclidean
 * distance (the HNSWIndex default).  M <= 63, ef_construction <= 512.  A
 * second call replaces the graphs; like the reference's insert! (hnsw/query.jl:
 * 161-163), it must not run while searches of the same handle are in flight
 * (searches from several threads may share a handle). */
int mann_ivf_hnsw_build(mann_ivf_t ivf, int64_t M, int64_t ef_construction, int neighbor_policy,
                        uint64_t seed);
/* query(::MultiLevelIndex{KMeans, TopK, HNSW}, data, Q, k; ef_search)
 * (src/indices/multilevel/query.jl:279-512, single query :25-106): TopK routing
 * (route_dist as in mann_ivf_search), empty buckets dropped, each probed list's
 * HNSW query (src/indices/hnsw/query.jl:95-123: entry point marked, greedy
 * descent, base-layer search with ef = max(ef_search, k), first min(k, found)),
 * then the merge (MANN_MERGE_SIMPLE / MANN_MERGE_DISJOINT, merge.jl:79-169) of the
 * per-list results in bucket order, or in probe order when probe_order = 1.
 * Output k x nq as mann_ivf_search, 0 / +inf padded.  max(ef_search, k) <= 512. */
int mann_ivf_hnsw_search(mann_ivf_t ivf, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                         int64_t k, int64_t nprobe, int64_t ef_search, const float* route_dist,
                         int merge, int probe_order, int64_t* out_ids, float* out_dist);
/* Install HNSW graphs built elsewhere -- the reference's own HNSWIndex children,
 * whose HNSWLayer slabs (hnsw/types.jl:18-49) the shim passes -- as the
 * handle's terminals, in the layout mann_ivf_hnsw_graph returns (host arrays;
 * level may be null: then a row's level is the highest layer where it has
 * neighbours, the entry point's its bucket's top layer).  Every id, degree
 * (<= M <= 63) and layer membership is checked before anything is installed. */
int mann_ivf_hnsw_set_graph(mann_ivf_t ivf, int64_t layers, int64_t M, const int32_t* nbr,
                            const int32_t* deg, const int32_t* entry, const int32_t* max_layer,
                            const int32_t* level);
/* The graphs, as the reference's HNSWLayer slabs (hnsw/types.jl:18-49) over the
 * list-contiguous rows: layers and M, then (each optional) out_nbr
 * [layers][n][M + 1] local 0-based ids (the live prefix is out_deg's count; rows
 * absent from a layer have degree 0), out_deg [layers][n], entry point (local
 * id) and top layer per bucket (-1 when empty), level per row. */
int mann_ivf_hnsw_graph(mann_ivf_t ivf, int64_t* out_layers, int64_t* out_M, int32_t* out_nbr,
                        int32_t* out_deg, int32_t* out_entry, int32_t* out_max_layer,
                        int32_t* out_level);

/* Learned probe router -- no counterpart in the reference: an MLP
 * d -> hidden (ReLU) -> nlist whose scores rank the IVF lists of a query
 * (higher = probe first).  It plugs into the reference's AbstractTransform seam
 * (src/transforms/Transform.jl:36-83) as a transform returning
 * KMeansAssignment(-scores), with TopKRouting unchanged
 * (src/indices/multilevel/routing.jl:94-118).  Parameters are FP32 (W1 hidden x d,
 * b1, W2 nlist x hidden, b2, row-major, concatenated); GEMMs run on bf16
 * operands with FP32 accumulation.  labels: n x L matrix of 1-based list ids
 * (0 = none); the target of a row is uniform over its labels. */
int mann_router_create(int64_t d, int64_t hidden, int64_t nlist, uint64_t seed,
                       mann_router_t* out);
int mann_router_destroy(mann_router_t r);
int mann_router_params(mann_router_t r, float* out, int64_t* count);
int mann_router_set_params(mann_router_t r, const float* params);
/* Centroid prior (gamma != 0): the logits gain gamma * (2 q.c_l - |c_l|^2), i.e.
 * -gamma |q - c_l|^2 up to a per-query constant, with c_l the index's centroids
 * (copied).  The MLP then learns a correction of the centroid routing.
 * gamma == 0 removes the prior (ivf may be null). */
int mann_router_set_prior(mann_router_t r, mann_ivf_t ivf, double gamma);
/* gradient of the mean softmax cross-entropy over one batch (no update) */
int mann_router_grad(mann_router_t r, const float* X, int64_t d, int64_t n, int64_t ldx,
                     const int64_t* labels, int64_t L, float* out_grad, double* out_loss);
/* mini-batch Adam over shuffled rows; out_loss = mean loss of the last epoch */
int mann_router_train(mann_router_t r, const float* X, int64_t d, int64_t n, int64_t ldx,
                      const int64_t* labels, int64_t L, int64_t epochs, int64_t batch, double lr,
                      uint64_t seed, double* out_loss);
int mann_router_scores(mann_router_t r, const float* Q, int64_t d, int64_t nq, int64_t ldq,
                       float* out_scores);
/* Probability-mass probing (no reference counterpart; for a learned router):
 * route_dist holds negated logits; per query the lists are taken in ascending
 * route_dist order (ties to the lower bucket, empty buckets dropped) until the
 * softmax probabilities of the lists taken reach `mass`, at most nprobe_max
 * non-empty lists (nprobe_max <= 64, at most 1,024 lists).  Output as
 * mann_ivf_probe: nq x nprobe_max, 1-based buckets, 0 = unused slot. */
int mann_ivf_probe_mass(mann_ivf_t ivf, const float* route_dist, int64_t nq, int64_t nprobe_max,
                        double mass, int64_t* out_buckets);
/* mann_ivf_search with the router's negated scores as routing distances */
int mann_ivf_search_routed(mann_ivf_t ivf, mann_router_t r, const float* queries, int64_t d,
                           int64_t nq, int64_t ldq, int64_t k, int64_t nprobe, int metric,
                           int64_t* out_ids, float* out_dist);
/* the same with probability-mass probing of the router's softmax (see
 * mann_ivf_probe_mass); the per-query list count varies */
int mann_ivf_search_routed_mass(mann_ivf_t ivf, mann_router_t r, const float* queries, int64_t d,
                                int64_t nq, int64_t ldq, int64_t k, int64_t nprobe_max,
                                double mass, int metric, int64_t* out_ids, float* out_dist);
/* (list, query) pairs scanned by the last IVF search of the calling thread */
int mann_last_probe_count(int64_t* total);
/* First-pass candidates the selection read in the last exact or IVF search of
 * the calling thread (all passes).  No reference counterpart: a diagnostic of
 * the emission threshold. */
int mann_last_emitted(int64_t* total);

/* Tuning knobs of the search kernels (process-wide): "scan_cfg", "ivf_cfg",
 * "kmeans_cfg" (first-pass tile of the exact kNN scan, the IVF list scan and
 * the k-means assignment: 0 = 128 rows x 128 queries, 1 = 256 x 128,
 * 2 = 128 x 256, 3 = 256 x 256 on 4 wavefronts, 4 and 5 = 256 x 256 on 8
 * wavefronts, 6 and 7 = configs 5 and 0 with four LDS stages instead of
 * three, 8 = 128 x 64, 9 = 128 x 32 on 2 wavefronts), "ivf_tau_cfg" (the
 * same for the IVF first-list threshold scan, default 9), "ivf_rows_item"
 * (rows of one IVF work item, a multiple of 256, default 2,048),
 * "ivf_first_rows" (rows of each query's first list that the IVF threshold
 * scan covers: 0 = the whole list, the default, else at least 16), "xcd_map"
 * (grid order: 0 plain, 1 query-major per XCD,
 * 2 row-partition-major per XCD, 3 bands of 8 row partitions per XCD),
 * "target_groups" (workgroups the row partitioning aims for), "emit_per_k"
 * (exact kNN: first-pass emissions per query the threshold sample aims at, per
 * neighbour requested), "sample_rank" (multiplier of that target),
 * "shortlist_mult" (first exact re-rank batch = that multiple of k),
 * "ivf_tau_mode" (IVF emission threshold: 0 = exact FP32 sample of the probed
 * rows, 1 = bf16 scan of each query's first list, the default), "ivf_tau_grp"
 * (that scan stores the minimum of every 4 consecutive rows of a first list
 * of at least max(64, 4 k) rows instead of every value: 1, the default, or
 * 0), "ivf_route_cert" (centroid TopKRouting of the IVF searches from a bf16
 * first pass over the centroids, with FP32 norms-form re-rank of the
 * centroids its bound cannot exclude: 1, the default, or 0 = every routing
 * distance in FP32; the probes are the same), "ivf_route_cfg" (tile
 * configuration of that first pass, default 0), "tau_rows"
 * (rows of that exact sample, 16..1024), "fold_norms" (exact kNN: 1 = when a
 * dataset has 6 spare padded columns, carry the norms there so that the MFMA
 * produces the first-pass distances itself; read when a dataset is created),
 * "sample_rows" (exact kNN: minimum
 * rows of the stride sample a dataset keeps for the emission threshold; read
 * when the dataset is created), "scan8_cfg" (FP8 first-pass tile
 * configuration, 0..4), "fp8_emit_per_k" (FP8 first pass: the bf16 sample
 * rank it starts from, per neighbour requested, before the FP8 band widens
 * it).  mann_tune_get also reads
 * "last_heap_replays": the queries of the last IVF-Flat search on this thread
 * whose k-th place had a distance tie and went through the bounded-heap merge
 * (read only), and "last_hnsw_distances": the exact distances the graph walks
 * of the last IVF-HNSW search on this thread computed (read only).  No
 * reference counterpart: the reference has no tunable
 * kernels. */
int mann_tune_set(const char* key, int64_t value);
int mann_tune_get(const char* key, int64_t* value);

/* Kernel timing of the calling process (HIP events on the launching stream). */
int mann_profile_enable(int on);
int mann_profile_reset(void);
/* total milliseconds and launch count of kernel `name` since the last reset */
int mann_profile_query(const char* name, int64_t* count, double* total_ms);
/* Statistics of the last search call on this thread: [0] candidates
 * re-ranked, [1] queries whose candidate buffer overflowed, [2] queries sent
 * to the second pass, [3] queries sent to the exact fallback. */
int mann_last_stats(int64_t* out4);

#ifdef __cplusplus
}
#endif

#endif /* MANN_H */
