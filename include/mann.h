/*
 * mann.h -- C ABI of the MI355X (gfx950) search core of ManifoldANN.jl.
 *
 * Every entry point replaces one reference Julia method on the data-parallel
 * search path (citations are ManifoldANN.jl source files).  Conventions:
 *  - A matrix argument is a Julia column-major d x n Float32 matrix: n points,
 *    each d contiguous floats, point stride `ld` (>= d) in elements.
 *  - Pointers may be host memory (staged through the call) or device memory
 *    of the current HIP device (used in place); the library checks which.
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
int mann_dataset_destroy(mann_dataset_t ds);
int mann_dataset_info(mann_dataset_t ds, int64_t* d, int64_t* n);

/* query(::BruteForceIndex, data, Q, k) -- batch form (src/ann_index.jl:71-101
 * over src/indices/bruteforce.jl:46-71, validation src/utils/validation.jl:8-34).
 * k_eff = min(k, n); slots past k_eff are empty.  nq = 0 is valid. */
int mann_bf_knn(mann_dataset_t base, const float* queries, int64_t d, int64_t nq, int64_t ldq,
                int64_t k, int metric, int64_t* out_ids, float* out_dist);

/* build_knn_graph(index, data; k, include_self) -- directed adjacency
 * (src/graphs/knn_graph.jl:42-81, _prepare_neighbors :105-130): each point is
 * queried with k+1 (k if include_self), itself dropped by id, first k kept.
 * out_ids is k x n.  MANN_ERR_ARGUMENT unless 0 < k <= n - !include_self. */
int mann_knn_graph(mann_dataset_t data, int64_t k, int include_self, int metric,
                   int64_t* out_ids, float* out_dist);

/* _symmetrize_adjacency (src/graphs/knn_graph.jl:196-210): union of i->j and
 * j->i, each row sorted ascending, as CSR (out_offsets n+1, out_ids 2*n*k
 * capacity).  *out_total receives the edge count; *out_max_degree the new
 * graph.k (:74). */
int mann_knn_graph_symmetrize(const int64_t* adj, int64_t n, int64_t k, int64_t* out_offsets,
                              int64_t* out_ids, int64_t* out_total, int64_t* out_max_degree);

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
