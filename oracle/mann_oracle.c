/*
 * mann_oracle.c -- CPU restatement of ManifoldANN.jl's data-parallel search core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle (the checker) and
 * the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path (manifoldann_amd) never does.
 *
 * The reference is pure Julia (no Julia in this image), so every function here
 * restates the reference algorithm from its source, cited per function.  Where
 * the reference leaves an arithmetic order unpinned (Distances.jl's SIMD loop,
 * OpenBLAS SGEMM/SGEMV, thread-count-dependent chunk sums), this file fixes
 * one order and the HIP library reproduces exactly that order:
 *
 *   "A16" order for a length-d FP32 reduction of terms a_i * b_i:
 *     16 accumulators acc[0..15] = +0.0f; element i goes to acc[i & 15] via
 *     acc = fmaf(a_i, b_i, acc) in increasing i; then the tree
 *     acc[j] += acc[j + 8] (j < 8), += acc[j + 4] (j < 4), += acc[j + 2]
 *     (j < 2), += acc[j + 1] (j < 1); the result is acc[0].
 *   Zero padding past d adds exact zeros and leaves the result unchanged, so a
 *   16-lane GPU group (lane j owns i = j mod 16) reproduces it bit for bit.
 *
 * Compile with -ffp-contract=off (no silent FMA contraction) and without
 * -ffast-math: every rounding is the one written here.
 *
 * Ids crossing this ABI are 1-based int64 (Julia `Neighbor.id`); 0 marks an
 * empty slot.  Matrices are Julia column-major d x n, i.e. n rows of d floats
 * with row stride `ld` (elements).
 *
 * Parity pins (see tests/test_oracle_*.py): the reference's own known-answer
 * tests, closed forms, an FP64 brute force, and tiny-input exhaustive checks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EUCLIDEAN 0
#define ORC_SQEUCLIDEAN 1
#define ORC_COSINE 2

/* ------------------------------------------------------------------------ */
/* A16 reductions                                                          */
/* ------------------------------------------------------------------------ */

static float tree16(float* acc) {
  for (int s = 8; s >= 1; s >>= 1)
    for (int j = 0; j < s; ++j) acc[j] = acc[j] + acc[j + s];
  return acc[0];
}

/* sum_i (x_i - q_i)^2 ; Distances.jl SqEuclidean (utils/distances.jl:16) */
float orc_sumsq_diff(const float* x, const float* q, int64_t d) {
  float acc[16];
  for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
  int64_t i = 0;
  for (; i + 16 <= d; i += 16)
    for (int j = 0; j < 16; ++j) {
      float t = x[i + j] - q[i + j];
      acc[j] = fmaf(t, t, acc[j]);
    }
  for (int j = 0; i < d; ++i, ++j) {
    float t = x[i] - q[i];
    acc[j] = fmaf(t, t, acc[j]);
  }
  return tree16(acc);
}

/* sum_i x_i * y_i in A16 order (dot and squared norms of the norms form) */
float orc_dot(const float* x, const float* y, int64_t d) {
  float acc[16];
  for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
  int64_t i = 0;
  for (; i + 16 <= d; i += 16)
    for (int j = 0; j < 16; ++j) acc[j] = fmaf(x[i + j], y[i + j], acc[j]);
  for (int j = 0; i < d; ++i, ++j) acc[j] = fmaf(x[i], y[i], acc[j]);
  return tree16(acc);
}

/* Per-pair metric of BruteForceIndex.query (bruteforce.jl:60-62).
 * Euclidean = sqrt(sum (x-q)^2) (utils/distances.jl:15), SqEuclidean = the sum
 * (:16), CosineDist = max(1 - x.q / (sqrt(|x|^2) sqrt(|q|^2)), 0) (:18). */
float orc_metric(int metric, const float* x, const float* q, int64_t d) {
  if (metric == ORC_EUCLIDEAN) return sqrtf(orc_sumsq_diff(x, q, d));
  if (metric == ORC_SQEUCLIDEAN) return orc_sumsq_diff(x, q, d);
  float ab = orc_dot(x, q, d), a2 = orc_dot(x, x, d), b2 = orc_dot(q, q, d);
  float v = 1.0f - ab / (sqrtf(a2) * sqrtf(b2));
  return v > 0.0f ? v : 0.0f;
}

/* ------------------------------------------------------------------------ */
/* (dist, id) ordering: Julia partialsortperm over Perm ordering -> ties go to */
/* the lower index (bruteforce.jl:65).                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
  float dist;
  int64_t id; /* 0-based here */
} cand_t;

static int cand_less(const cand_t* a, const cand_t* b) {
  return a->dist < b->dist || (a->dist == b->dist && a->id < b->id);
}

/* keep the k smallest of a stream in a sorted array (insertion) */
static void topk_push(cand_t* best, int64_t* cnt, int64_t k, cand_t c) {
  int64_t m = *cnt;
  if (m == k && !cand_less(&c, &best[k - 1])) return;
  int64_t pos = (m < k) ? m : k - 1;
  while (pos > 0 && cand_less(&c, &best[pos - 1])) {
    best[pos] = best[pos - 1];
    --pos;
  }
  best[pos] = c;
  if (m < k) *cnt = m + 1;
}

static int nthreads_or_default(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* BruteForceIndex batch query (ann_index.jl:71-101 -> bruteforce.jl:46-71).
 * out_ids / out_dist are k x nq column-major (query qi's results at qi*k).
 * k_eff = min(k, n); slots beyond k_eff get id 0 / +inf. */
int orc_bf_knn(const float* base, int64_t d, int64_t n, int64_t ld, const float* Q,
               int64_t nq, int64_t ldq, int64_t k, int metric, int64_t* out_ids,
               float* out_dist, int nthreads) {
  if (k <= 0 || nq <= 0) return 0;
  int64_t keff = k < n ? k : n;
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t qi = 0; qi < nq; ++qi) {
    cand_t* best = (cand_t*)malloc(sizeof(cand_t) * (size_t)(keff > 0 ? keff : 1));
    int64_t cnt = 0;
    const float* q = Q + qi * ldq;
    for (int64_t j = 0; j < n; ++j) {
      cand_t c = {orc_metric(metric, base + j * ld, q, d), j};
      topk_push(best, &cnt, keff, c);
    }
    for (int64_t r = 0; r < k; ++r) {
      out_ids[qi * k + r] = r < cnt ? best[r].id + 1 : 0;
      out_dist[qi * k + r] = r < cnt ? best[r].dist : INFINITY;
    }
    free(best);
  }
  return 0;
}

/* Directed build_knn_graph (knn_graph.jl:42-81, _prepare_neighbors :105-130):
 * query every column with k+1 (k if include_self), drop the point itself by id,
 * keep the first k.  Returns -2 (ArgumentError) if a row cannot be filled. */
int orc_knn_graph(const float* X, int64_t d, int64_t n, int64_t ld, int64_t k,
                  int include_self, int metric, int64_t* out_ids, int nthreads) {
  int64_t maxnb = n - (include_self ? 0 : 1);
  if (!(k > 0 && k <= maxnb)) return -2;
  int64_t req = k + (include_self ? 0 : 1);
  int64_t* ids = (int64_t*)malloc(sizeof(int64_t) * (size_t)(req * n));
  float* dist = (float*)malloc(sizeof(float) * (size_t)(req * n));
  orc_bf_knn(X, d, n, ld, X, n, ld, req, metric, ids, dist, nthreads);
  int bad = 0;
  for (int64_t c = 0; c < n; ++c) {
    int64_t m = 0;
    for (int64_t r = 0; r < req && m < k; ++r) {
      int64_t id = ids[c * req + r];
      if (id == 0) break;
      if (!include_self && id == c + 1) continue;
      out_ids[c * k + m++] = id;
    }
    if (m != k) bad = 1;
  }
  free(ids);
  free(dist);
  return bad ? -2 : 0;
}

static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

/* _symmetrize_adjacency (knn_graph.jl:196-210): union of i->j and j->i per
 * node, ids sorted ascending.  CSR output; out_ids needs room for 2*n*k. */
int64_t orc_symmetrize(const int64_t* adj, int64_t n, int64_t k, int64_t* out_offsets,
                       int64_t* out_ids) {
  int64_t* deg = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t r = 0; r < k; ++r) {
      deg[i]++;
      deg[adj[i * k + r] - 1]++;
    }
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  off[0] = 0;
  for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + deg[i];
  int64_t* fill = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(off[n] > 0 ? off[n] : 1));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t r = 0; r < k; ++r) {
      int64_t j = adj[i * k + r] - 1;
      tmp[off[i] + fill[i]++] = j + 1;
      tmp[off[j] + fill[j]++] = i + 1;
    }
  int64_t w = 0;
  out_offsets[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    qsort(tmp + off[i], (size_t)deg[i], sizeof(int64_t), cmp_i64);
    for (int64_t t = 0; t < deg[i]; ++t) {
      int64_t v = tmp[off[i] + t];
      if (t > 0 && v == tmp[off[i] + t - 1]) continue;
      out_ids[w++] = v;
    }
    out_offsets[i + 1] = w;
  }
  free(deg);
  free(off);
  free(fill);
  free(tmp);
  return w;
}

/* ------------------------------------------------------------------------ */
/* k-means (transforms/kmeans/...)                                            */
/* ------------------------------------------------------------------------ */

/* squared norms, A16 (sum(abs2, X; dims=1), kmeans/distance.jl:81-84) */
void orc_norms(const float* X, int64_t d, int64_t n, int64_t ld, float* out) {
  for (int64_t j = 0; j < n; ++j) out[j] = orc_dot(X + j * ld, X + j * ld, d);
}

/* Norms-form Euclidean (kmeans/distance.jl:99-101, :237-238):
 * sqrt(max((|c|^2 + |x|^2) - 2 c.x, 0)) in FP32. */
static float norms_form(float cn, float xn, float dot) {
  float t = cn + xn;
  float sq = t - 2.0f * dot;
  return sqrtf(sq > 0.0f ? sq : 0.0f);
}

/* E-step: pairwise_euclidean! + assign_clusters! (kmeans/distance.jl:68-107,
 * 287-310): argmin over centroids with strict '<' (lowest index wins).
 * out_assign 1-based; out_mindist = D[assign[j], j]. */
int orc_kmeans_assign(const float* X, int64_t d, int64_t n, int64_t ld, const float* C,
                      int64_t k, int64_t* out_assign, float* out_mindist, int nthreads) {
  float* cn = (float*)malloc(sizeof(float) * (size_t)k);
  orc_norms(C, d, k, d, cn);
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t j = 0; j < n; ++j) {
    const float* x = X + j * ld;
    float xn = orc_dot(x, x, d);
    float best = norms_form(cn[0], xn, orc_dot(C, x, d));
    int64_t bi = 0;
    for (int64_t i = 1; i < k; ++i) {
      float v = norms_form(cn[i], xn, orc_dot(C + i * d, x, d));
      if (v < best) {
        best = v;
        bi = i;
      }
    }
    out_assign[j] = bi + 1;
    if (out_mindist) out_mindist[j] = best;
  }
  free(cn);
  return 0;
}

#define ORC_MSTEP_CHUNK 256

/* M-step sums in the fixed order the HIP library uses: members of a cluster in
 * ascending point id, sequential FP64 sums over chunks of 256 members, then a
 * sequential FP64 sum of the chunk partials.  (The reference's
 * _m_step_parallel!, lloyd.jl:149-212, sums FP32 per thread chunk, so its
 * bits depend on the thread count; this order is fixed and FP64.) */
static void mstep_sums(const float* X, int64_t d, int64_t n, int64_t ld,
                       const int64_t* assign, int64_t k, double* sums, int64_t* counts) {
  int64_t* off = (int64_t*)calloc((size_t)(k + 1), sizeof(int64_t));
  for (int64_t j = 0; j < n; ++j) off[assign[j]]++; /* assign is 1-based */
  for (int64_t c = 0; c < k; ++c) off[c + 1] += off[c];
  int64_t* members = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t* fill = (int64_t*)calloc((size_t)k, sizeof(int64_t));
  for (int64_t j = 0; j < n; ++j) {
    int64_t c = assign[j] - 1;
    members[off[c] + fill[c]++] = j;
  }
  double* part = (double*)malloc(sizeof(double) * (size_t)d);
  for (int64_t c = 0; c < k; ++c) {
    int64_t lo = off[c], hi = off[c + 1];
    counts[c] = hi - lo;
    double* s = sums + c * d;
    for (int64_t r = 0; r < d; ++r) s[r] = 0.0;
    for (int64_t t = lo; t < hi; t += ORC_MSTEP_CHUNK) {
      int64_t te = t + ORC_MSTEP_CHUNK < hi ? t + ORC_MSTEP_CHUNK : hi;
      for (int64_t r = 0; r < d; ++r) part[r] = 0.0;
      for (int64_t m = t; m < te; ++m) {
        const float* x = X + members[m] * ld;
        for (int64_t r = 0; r < d; ++r) part[r] += (double)x[r];
      }
      for (int64_t r = 0; r < d; ++r) s[r] += part[r];
    }
  }
  free(part);
  free(off);
  free(members);
  free(fill);
}

/* lloyd! (kmeans/lloyd.jl:36-134): E-step, M-step, empty-cluster repair by
 * farthest point with a -Inf mark (:81-104), divide (:109; FP64 sum / count
 * rounded once to FP32 here), max centroid movement (Euclidean, :113-118)
 * < tol -> stop, final E-step (:130-131).  C is d x k, updated in place. */
int orc_lloyd(const float* X, int64_t d, int64_t n, int64_t ld, float* C, int64_t k,
              int64_t max_iters, double tol, int64_t* out_assign, int64_t* out_iters,
              int nthreads) {
  if (k > n) return -3;
  float* mind = (float*)malloc(sizeof(float) * (size_t)n);
  double* sums = (double*)malloc(sizeof(double) * (size_t)(k * d));
  int64_t* counts = (int64_t*)malloc(sizeof(int64_t) * (size_t)k);
  float* newc = (float*)malloc(sizeof(float) * (size_t)(k * d));
  int64_t it = 0;
  for (int64_t iter = 1; iter <= max_iters; ++iter) {
    it = iter;
    orc_kmeans_assign(X, d, n, ld, C, k, out_assign, mind, nthreads);
    mstep_sums(X, d, n, ld, out_assign, k, sums, counts);
    for (int64_t i = 0; i < k; ++i) {
      if (counts[i] == 0) {
        float maxd = -INFINITY;
        int64_t far = 0;
        for (int64_t j = 0; j < n; ++j)
          if (mind[j] > maxd) {
            maxd = mind[j];
            far = j;
          }
        memcpy(newc + i * d, X + far * ld, sizeof(float) * (size_t)d);
        counts[i] = 1;
        mind[far] = -INFINITY;
      } else {
        for (int64_t r = 0; r < d; ++r)
          newc[i * d + r] = (float)(sums[i * d + r] / (double)counts[i]);
      }
    }
    double maxmove = 0.0;
    for (int64_t i = 0; i < k; ++i) {
      double mv = (double)sqrtf(orc_sumsq_diff(C + i * d, newc + i * d, d));
      if (mv > maxmove) maxmove = mv;
    }
    memcpy(C, newc, sizeof(float) * (size_t)(k * d));
    if (maxmove < tol) break;
  }
  orc_kmeans_assign(X, d, n, ld, C, k, out_assign, mind, nthreads);
  if (out_iters) *out_iters = it;
  free(mind);
  free(sums);
  free(counts);
  free(newc);
  return 0;
}

/* The library's documented RNG: splitmix64; uniform double in [0, 1) from the
 * top 53 bits.  (Julia's MersenneTwister/Xoshiro streams are not reproducible
 * outside Julia; kmeans++ draws come from this generator instead.) */
uint64_t orc_splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double orc_uniform(uint64_t* state) {
  return (double)(orc_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

#define ORC_PP_BLOCK 1024

/* init_kmeans_plus_plus (kmeans/init.jl:56-102, update :110-128) driven by
 * explicit uniform draws u[0..k-1] in [0, 1).
 * Draw u_0 picks the first centre floor(u_0 * n) (rand(rng, 1:n), :69); draw
 * u_i (i >= 1) is the threshold fraction of round i (rand(rng) * total, :92).
 * The squared distances follow _update_distances_euclidean! (:116-127): the
 * FP32 norms form, clamped at 0, widened to the FP64 running minimum.  The
 * cumulative sum uses blocks of 1024: sequential FP64 sums inside a block, a
 * sequential exclusive prefix over the block totals, cumsum[j] =
 * prefix[block] + local inclusive sum (the reference's cumsum is one
 * sequential pass, :89).  next = the first j with cumsum[j] >= u * total
 * (searchsortedfirst, :93), clamped to the last point (:96).  out_idx 1-based. */
int orc_kmeanspp_draws(const float* X, int64_t d, int64_t n, int64_t ld, int64_t k,
                       const double* u, float* out_C, int64_t* out_idx, int nthreads) {
  if (k > n) return -3;
  if (k <= 0) return 0;
  double* mins = (double*)malloc(sizeof(double) * (size_t)n);
  double* cum = (double*)malloc(sizeof(double) * (size_t)n);
  float* xn = (float*)malloc(sizeof(float) * (size_t)n);
  orc_norms(X, d, n, ld, xn);
  for (int64_t j = 0; j < n; ++j) mins[j] = INFINITY;
  int64_t first = (int64_t)(u[0] * (double)n);
  if (first >= n) first = n - 1;
  memcpy(out_C, X + first * ld, sizeof(float) * (size_t)d);
  if (out_idx) out_idx[0] = first + 1;
  int nt = nthreads_or_default(nthreads);
  for (int64_t i = 1; i < k; ++i) {
    const float* c = out_C + (i - 1) * d;
    float cn = orc_dot(c, c, d);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t j = 0; j < n; ++j) {
      float sq = (cn + xn[j]) - 2.0f * orc_dot(c, X + j * ld, d);
      double v = sq > 0.0f ? (double)sq : 0.0;
      if (v < mins[j]) mins[j] = v;
    }
    double prefix = 0.0;
    for (int64_t b = 0; b < n; b += ORC_PP_BLOCK) {
      int64_t be = b + ORC_PP_BLOCK < n ? b + ORC_PP_BLOCK : n;
      double loc = 0.0;
      for (int64_t j = b; j < be; ++j) {
        loc += mins[j];
        cum[j] = prefix + loc;
      }
      prefix += loc;
    }
    double thr = u[i] * prefix;
    int64_t lo = 0, hi = n; /* first j with cum[j] >= thr */
    while (lo < hi) {
      int64_t mid = lo + (hi - lo) / 2;
      if (cum[mid] >= thr) hi = mid;
      else lo = mid + 1;
    }
    int64_t next = lo < n ? lo : n - 1;
    memcpy(out_C + i * d, X + next * ld, sizeof(float) * (size_t)d);
    if (out_idx) out_idx[i] = next + 1;
  }
  free(mins);
  free(cum);
  free(xn);
  return 0;
}

/* The same with the draws taken from the splitmix64 stream of `seed`, in
 * order u_0, u_1, ... (the library's mann_kmeanspp). */
int orc_kmeanspp(const float* X, int64_t d, int64_t n, int64_t ld, int64_t k, uint64_t seed,
                 float* out_C, int64_t* out_idx, int nthreads) {
  if (k > n) return -3;
  if (k <= 0) return 0;
  double* u = (double*)malloc(sizeof(double) * (size_t)k);
  uint64_t st = seed;
  for (int64_t i = 0; i < k; ++i) u[i] = orc_uniform(&st);
  int rc = orc_kmeanspp_draws(X, d, n, ld, k, u, out_C, out_idx, nthreads);
  free(u);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* IVF: routing, partition, batched list scans, merges                       */
/* ------------------------------------------------------------------------ */

/* transform(::KMeansTransform) -> compute_distances_euclidean
 * (KMeansTransform.jl:181-190, kmeans/distance.jl:213-242), nq x nlist. */
int orc_route_distances(const float* Q, int64_t d, int64_t nq, int64_t ldq, const float* C,
                        int64_t nlist, float* out, int nthreads) {
  float* cn = (float*)malloc(sizeof(float) * (size_t)nlist);
  orc_norms(C, d, nlist, d, cn);
  int nt = nthreads_or_default(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
  for (int64_t qi = 0; qi < nq; ++qi) {
    const float* q = Q + qi * ldq;
    float qn = orc_dot(q, q, d);
    for (int64_t i = 0; i < nlist; ++i)
      out[qi * nlist + i] = norms_form(cn[i], qn, orc_dot(C + i * d, q, d));
  }
  free(cn);
  return 0;
}

/* select_indices(::TopKRouting) (routing.jl:94-118): the kk = min(nprobe,
 * min(n_domain, n_dist)) smallest of the first `available` distances, ties to
 * the lower bucket.  Returns kk (or -2 on no distances), out_sel 1-based. */
int64_t orc_select_topk(const float* dist, int64_t n_dist, int64_t n_domain, int64_t nprobe,
                        int64_t* out_sel) {
  if (n_domain <= 0) return 0;
  int64_t avail = n_domain < n_dist ? n_domain : n_dist;
  if (avail <= 0) return -2;
  int64_t kk = nprobe < avail ? nprobe : avail;
  cand_t* best = (cand_t*)malloc(sizeof(cand_t) * (size_t)kk);
  int64_t cnt = 0;
  for (int64_t i = 0; i < avail; ++i) {
    cand_t c = {dist[i], i};
    topk_push(best, &cnt, kk, c);
  }
  for (int64_t r = 0; r < kk; ++r) out_sel[r] = best[r].id + 1;
  free(best);
  return kk;
}

/* partition_by_transform / _finalize_partitions (transforms/utils.jl:142-174):
 * num_buckets = max(assign); stable counting sort -> CSR (offsets[nb+1], ids
 * 1-based ascending inside each bucket).  Returns nb or -2. */
int64_t orc_partition(const int64_t* assign, int64_t n, int64_t* offsets, int64_t* ids) {
  int64_t nb = 0;
  for (int64_t j = 0; j < n; ++j) nb = assign[j] > nb ? assign[j] : nb;
  if (nb <= 0) return -2;
  for (int64_t b = 0; b <= nb; ++b) offsets[b] = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (assign[j] < 1) return -2;
    offsets[assign[j]]++;
  }
  for (int64_t b = 0; b < nb; ++b) offsets[b + 1] += offsets[b];
  int64_t* fill = (int64_t*)calloc((size_t)nb, sizeof(int64_t));
  for (int64_t j = 0; j < n; ++j) {
    int64_t b = assign[j] - 1;
    ids[offsets[b] + fill[b]++] = j + 1;
  }
  free(fill);
  return nb;
}

/* SimpleMerge (indices/multilevel/merge.jl:79-124): concatenate, stable sort by
 * dist, drop repeated ids, keep k.  lists are concatenated already. */
int64_t orc_simple_merge(const int64_t* ids, const float* dist, int64_t total, int64_t k,
                         int64_t* out_ids, float* out_dist) {
  if (total == 0 || k <= 0) return 0;
  int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * (size_t)total);
  for (int64_t i = 0; i < total; ++i) ord[i] = i;
  /* stable insertion-merge sort by dist (positions break ties) */
  for (int64_t i = 1; i < total; ++i) {
    int64_t v = ord[i], p = i;
    while (p > 0 && dist[ord[p - 1]] > dist[v]) {
      ord[p] = ord[p - 1];
      --p;
    }
    ord[p] = v;
  }
  int64_t cap = k < total ? k : total, kept = 0;
  for (int64_t i = 0; i < total && kept < cap; ++i) {
    int64_t c = ord[i], dup = 0;
    for (int64_t j = 0; j < kept; ++j)
      if (out_ids[j] == ids[c]) {
        dup = 1;
        break;
      }
    if (dup) continue;
    out_ids[kept] = ids[c];
    out_dist[kept] = dist[c];
    ++kept;
  }
  free(ord);
  return kept;
}

/* DisjointMerge (merge.jl:150-169) over BoundedMaxHeap (utils/neighbor_heaps.jl:
 * 84-102 push!, :156-190 sift, :138-141 to_sorted_vector sort by (dist, id)). */
int64_t orc_disjoint_merge(const int64_t* ids, const float* dist, int64_t total, int64_t k,
                           int64_t* out_ids, float* out_dist) {
  if (total == 0 || k <= 0) return 0;
  int64_t cap = k < total ? k : total, len = 0;
  cand_t* h = (cand_t*)malloc(sizeof(cand_t) * (size_t)cap);
  for (int64_t i = 0; i < total; ++i) {
    cand_t c = {dist[i], ids[i]};
    if (len < cap) {
      int64_t x = len++;
      h[x] = c;
      while (x > 0) { /* 0-based parent of x is (x-1)/2 */
        int64_t p = (x - 1) / 2;
        if (h[x].dist > h[p].dist) {
          cand_t t = h[x]; h[x] = h[p]; h[p] = t;
          x = p;
        } else break;
      }
    } else if (c.dist < h[0].dist) {
      h[0] = c;
      int64_t x = 0;
      for (;;) {
        int64_t l = 2 * x + 1, r = l + 1, g = x;
        if (l < len && h[l].dist > h[g].dist) g = l;
        if (r < len && h[r].dist > h[g].dist) g = r;
        if (g == x) break;
        cand_t t = h[x]; h[x] = h[g]; h[g] = t;
        x = g;
      }
    }
  }
  /* to_sorted_vector: sort by (dist, id) */
  for (int64_t i = 1; i < len; ++i) {
    cand_t v = h[i];
    int64_t p = i;
    while (p > 0 && cand_less(&v, &h[p - 1])) {
      h[p] = h[p - 1];
      --p;
    }
    h[p] = v;
  }
  for (int64_t i = 0; i < len; ++i) {
    out_ids[i] = h[i].id;
    out_dist[i] = h[i].dist;
  }
  free(h);
  return len;
}

/* Batched MultiLevelIndex{KMeans, TopK, BruteForce} query
 * (multilevel/query.jl:279-399, :441-474; BruteForce children bruteforce.jl:46-71).
 * route = nq x nlist routing distances (centroid distances, or -scores of a
 * learned router); nb = number of buckets (max assignment); CSR offsets/ids from
 * orc_partition.  Per query: probe the TopK buckets, drop empty ones
 * (query.jl:108-121), scan each probed list with the per-pair metric, keep
 * min(k, |list|) per list by (dist, id), concatenate the per-list results in
 * ascending bucket order (batched path) or probe order (single-query path,
 * probe_order=1), and merge (merge_kind 0 = SimpleMerge, 1 = DisjointMerge).
 * Output k slots per query, padded with id 0 / +inf. */
int orc_ivf_search(const float* X, int64_t d, int64_t ld, const int64_t* offsets,
                   const int64_t* ids, int64_t nb, const float* Q, int64_t nq, int64_t ldq,
                   const float* route, int64_t nlist, int64_t k, int64_t nprobe, int metric,
                   int merge_kind, int probe_order, int64_t* out_ids, float* out_dist,
                   int nthreads) {
  int nt = nthreads_or_default(nthreads);
  int err = 0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
  for (int64_t qi = 0; qi < nq; ++qi) {
    int64_t* sel = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nprobe > 0 ? nprobe : 1));
    int64_t ns = orc_select_topk(route + qi * nlist, nlist, nb, nprobe, sel);
    if (ns < 0) {
      err = 1;
      ns = 0;
    }
    /* drop buckets without a child, then order by bucket id when batched */
    int64_t m = 0;
    for (int64_t s = 0; s < ns; ++s)
      if (offsets[sel[s]] > offsets[sel[s] - 1]) sel[m++] = sel[s];
    if (!probe_order) qsort(sel, (size_t)m, sizeof(int64_t), cmp_i64);
    int64_t cap = m * k, tot = 0;
    int64_t* cid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(cap > 0 ? cap : 1));
    float* cd = (float*)malloc(sizeof(float) * (size_t)(cap > 0 ? cap : 1));
    cand_t* best = (cand_t*)malloc(sizeof(cand_t) * (size_t)(k > 0 ? k : 1));
    const float* q = Q + qi * ldq;
    for (int64_t s = 0; s < m; ++s) {
      int64_t b = sel[s] - 1, lo = offsets[b], hi = offsets[b + 1];
      int64_t kk = k < hi - lo ? k : hi - lo, cnt = 0;
      for (int64_t t = lo; t < hi; ++t) { /* local index t-lo, ascending global id */
        cand_t c = {orc_metric(metric, X + (ids[t] - 1) * ld, q, d), t - lo};
        topk_push(best, &cnt, kk, c);
      }
      for (int64_t r = 0; r < cnt; ++r) {
        cid[tot] = ids[lo + best[r].id];
        cd[tot] = best[r].dist;
        ++tot;
      }
    }
    int64_t* oi = out_ids + qi * k;
    float* od = out_dist + qi * k;
    int64_t got = merge_kind == 0 ? orc_simple_merge(cid, cd, tot, k, oi, od)
                                  : orc_disjoint_merge(cid, cd, tot, k, oi, od);
    for (int64_t r = got; r < k; ++r) {
      oi[r] = 0;
      od[r] = INFINITY;
    }
    free(sel);
    free(cid);
    free(cd);
    free(best);
  }
  return err ? -2 : 0;
}

/* recall_at_k (utils/eval_utils.jl:7-35), averaged over queries; ids 1-based,
 * zeros in `found` never match. */
double orc_recall_at_k(const int64_t* found, int64_t kf, const int64_t* truth, int64_t kt,
                       int64_t nq) {
  if (nq == 0 || kt <= 0) return 0.0;
  double total = 0.0;
  for (int64_t qi = 0; qi < nq; ++qi) {
    int64_t hits = 0;
    for (int64_t a = 0; a < kf; ++a) {
      int64_t v = found[qi * kf + a];
      if (v == 0) continue;
      for (int64_t b = 0; b < kt; ++b)
        if (truth[qi * kt + b] == v) {
          ++hits;
          break;
        }
    }
    total += (double)hits / (double)kt;
  }
  return total / (double)nq;
}

/* query(::IVFFlatIndex, data, q, k; nprobe) (src/indices/ivf_flat.jl:138-167),
 * one query at a time.  nprobe is clamped to [1, nlist] (:146); the probed
 * lists are the nprobe smallest routing distances (:149-150, ties to the lower
 * centroid), empty lists included; k_eff = min(k, total probed size) (:156);
 * every candidate is pushed, in probe order and ascending id inside a list,
 * into a bounded max-heap of capacity k_eff that replaces its root only on a
 * strictly smaller distance (BoundedMaxHeap push!, src/utils/neighbor_heaps.jl
 * :84-100, sift-up :156-168, sift-down :170-190); the survivors are sorted by
 * (dist, id) (to_sorted_vector :131-134).  offsets has nlist + 1 entries.
 * Output k slots per query padded with id 0 / +inf, count[q] = k_eff. */
typedef struct { int64_t id; float dist; } heap_nb_t;

static void heap_sift_up(heap_nb_t* h, int64_t i) { /* 1-based positions */
  while (i > 1) {
    int64_t p = i >> 1;
    if (h[i - 1].dist > h[p - 1].dist) {
      heap_nb_t t = h[i - 1]; h[i - 1] = h[p - 1]; h[p - 1] = t;
      i = p;
    } else {
      break;
    }
  }
}

static void heap_sift_down(heap_nb_t* h, int64_t len, int64_t i) {
  for (;;) {
    int64_t l = i << 1, r = l + 1, big = i;
    if (l <= len && h[l - 1].dist > h[big - 1].dist) big = l;
    if (r <= len && h[r - 1].dist > h[big - 1].dist) big = r;
    if (big == i) break;
    heap_nb_t t = h[i - 1]; h[i - 1] = h[big - 1]; h[big - 1] = t;
    i = big;
  }
}

static int cmp_heap_nb(const void* a, const void* b) {
  const heap_nb_t* x = (const heap_nb_t*)a;
  const heap_nb_t* y = (const heap_nb_t*)b;
  if (x->dist != y->dist) return x->dist < y->dist ? -1 : 1;
  return (x->id > y->id) - (x->id < y->id);
}

int orc_ivf_flat_query(const float* X, int64_t d, int64_t ld, const int64_t* offsets,
                       const int64_t* ids, int64_t nlist, const float* Q, int64_t nq,
                       int64_t ldq, const float* route, int64_t k, int64_t nprobe, int metric,
                       int64_t* out_ids, float* out_dist, int64_t* out_count) {
  if (nlist <= 0) return -2;
  int64_t np = nprobe < 1 ? 1 : (nprobe > nlist ? nlist : nprobe);
  int64_t* sel = (int64_t*)malloc(sizeof(int64_t) * (size_t)np);
  for (int64_t qi = 0; qi < nq; ++qi) {
    int64_t ns = orc_select_topk(route + qi * nlist, nlist, nlist, np, sel);
    int64_t tot = 0;
    for (int64_t s = 0; s < ns; ++s) tot += offsets[sel[s]] - offsets[sel[s] - 1];
    int64_t keff = k < tot ? k : tot;
    for (int64_t r = 0; r < k; ++r) {
      out_ids[qi * k + r] = 0;
      out_dist[qi * k + r] = INFINITY;
    }
    out_count[qi] = keff > 0 ? keff : 0;
    if (keff <= 0) continue;
    heap_nb_t* h = (heap_nb_t*)malloc(sizeof(heap_nb_t) * (size_t)keff);
    int64_t len = 0;
    const float* q = Q + qi * ldq;
    for (int64_t s = 0; s < ns; ++s) {
      int64_t b = sel[s] - 1;
      for (int64_t t = offsets[b]; t < offsets[b + 1]; ++t) {
        heap_nb_t c = {ids[t], orc_metric(metric, X + (ids[t] - 1) * ld, q, d)};
        if (len < keff) {
          h[len++] = c;
          heap_sift_up(h, len);
        } else if (c.dist < h[0].dist) {
          h[0] = c;
          heap_sift_down(h, len, 1);
        }
      }
    }
    qsort(h, (size_t)len, sizeof(heap_nb_t), cmp_heap_nb);
    for (int64_t r = 0; r < len; ++r) {
      out_ids[qi * k + r] = h[r].id;
      out_dist[qi * k + r] = h[r].dist;
    }
    free(h);
  }
  free(sel);
  return 0;
}
