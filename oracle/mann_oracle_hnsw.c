/* CPU oracle, part 2: the HNSW terminals of build_ivf_hnsw_index.
 *
 * TEST INFRASTRUCTURE ONLY (like mann_oracle.c): used by tests/,
 * __graft_entry__.smoke() and bench.py's CPU baseline as the checker; the
 * product library never links or calls it.
 *
 * Restated reference (file:line relative to the reference tree):
 *   src/indices/multilevel/ivf_hnsw.jl:46-88   build_ivf_hnsw_index
 *   src/indices/hnsw/index.jl:10-67            build_index(HNSWIndex), serial path
 *   src/indices/hnsw/query.jl:63-123           single query / _query_search
 *   src/indices/hnsw/query.jl:176-238          batch query (same per-query search)
 *   src/indices/hnsw/query.jl:264-325          insert!
 *   src/indices/hnsw/query.jl:327-345          _greedy_descent
 *   src/indices/hnsw/query.jl:350-442          _search_layer(_pooled)
 *   src/indices/hnsw/query.jl:444-603          _connect_new_node!, prune, _select_into!
 *   src/indices/hnsw/neighbor_policy.jl:22-112 select_neighbors (heuristic, diversified)
 *   src/indices/hnsw/traversal.jl:33-130       BestCandidatesHeap, traversal policy
 *   src/indices/hnsw/candidate_heap.jl:15-73   NeighborMinHeap
 *   src/indices/hnsw/planner.jl:19-30          sample_layer
 *
 * Fixed choices (the reference leaves them to Julia's runtime):
 *  - Levels.  The reference's child builds draw from Random.default_rng() (the
 *    terminal builder receives no rng, multilevel/builder.jl:232-239).  Here
 *    the level of every point comes from one splitmix64 stream of the given
 *    seed, drawn in the list-contiguous row order (bucket order, ascending id
 *    inside a bucket): level = floor(-log(max(u, eps)) * lambda), lambda =
 *    1 / log(max(M, 2)) (planner.jl:19-30, index.jl:35).
 *  - Serial insertion in local id order (index.jl:58-63, threaded = false):
 *    the threaded build is not deterministic (query.jl:652-658).
 *  - Distances: Euclidean (the HNSW default; the child receives no distance,
 *    multilevel/ivf_hnsw.jl:67-72), A16 FP32 order of orc_metric.
 *  - Sorts: Julia's MergeSort is stable (query.jl:400, 439, 555); the
 *    heuristic policy's partialsort! is not (query.jl:538-541) and is restated
 *    as stable (ties in input order): parity unpinned for that tie order.
 * Ids here are 0-based local positions inside a list (Julia's 1-based local
 * ids minus one).  Graph arrays: nbr[layer][row][M + 1] and deg[layer][row]
 * over the rows of all lists (list-contiguous), a node absent from a layer
 * having degree 0.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

float orc_metric(int metric, const float* x, const float* q, int64_t d);
double orc_uniform(uint64_t* state);
int64_t orc_select_topk(const float* dist, int64_t n_dist, int64_t n_domain, int64_t nprobe,
                        int64_t* out);
int64_t orc_simple_merge(const int64_t* ids, const float* dist, int64_t total, int64_t k,
                         int64_t* out_ids, float* out_dist);
int64_t orc_disjoint_merge(const int64_t* ids, const float* dist, int64_t total, int64_t k,
                           int64_t* out_ids, float* out_dist);

#define HN_HEURISTIC 0
#define HN_DIVERSIFIED 1

typedef struct {
  float dist;
  int32_t id;
} hc_t;

/* one list's graph and the rows it indexes */
typedef struct {
  const float* X;
  int64_t d, ld;
  const int64_t* rows; /* local id -> 0-based row of X */
  int64_t n, M, cap, nl_all; /* cap = M + 1; nl_all = rows of all lists (layer stride) */
  int64_t ncap;              /* stamps allocated */
  int32_t* nbr;  /* nbr + (layer * nl_all + local) * cap, already offset to the list */
  int32_t* deg;  /* deg + layer * nl_all + local */
  int32_t entry, max_layer;
  /* visited stamps (StampVisited / BitSetVisited: the same set semantics) */
  uint32_t* stamp;
  uint32_t gen;
  hc_t* best;    /* capacity ef */
  hc_t* pend;    /* capacity n + 1: every node enters the pending heap at most once */
} hg_t;

static inline float hdist(const hg_t* g, int32_t a, const float* q) {
  return orc_metric(0, g->X + g->rows[a] * g->ld, q, g->d);
}
static inline float hdist2(const hg_t* g, int32_t a, int32_t b) {
  return orc_metric(0, g->X + g->rows[a] * g->ld, g->X + g->rows[b] * g->ld, g->d);
}
static inline int32_t* nbr_of(const hg_t* g, int layer, int32_t v) {
  return g->nbr + ((int64_t)layer * g->nl_all + v) * g->cap;
}
static inline int32_t* deg_of(const hg_t* g, int layer, int32_t v) {
  return g->deg + (int64_t)layer * g->nl_all + v;
}
static inline void new_visit(hg_t* g) {
  if (++g->gen == 0) {
    memset(g->stamp, 0, sizeof(uint32_t) * (size_t)g->ncap);
    g->gen = 1;
  }
}
static inline int visited(const hg_t* g, int32_t v) { return g->stamp[v] == g->gen; }
static inline void mark(hg_t* g, int32_t v) { g->stamp[v] = g->gen; }

/* NeighborMinHeap push / popfirst! (candidate_heap.jl:24-73), 0-based */
static void min_push(hc_t* a, int64_t* len, hc_t c) {
  int64_t i = (*len)++;
  a[i] = c;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (a[i].dist < a[p].dist) {
      hc_t t = a[i]; a[i] = a[p]; a[p] = t;
      i = p;
    } else break;
  }
}
static hc_t min_pop(hc_t* a, int64_t* len) {
  hc_t top = a[0];
  hc_t last = a[--(*len)];
  if (*len > 0) {
    a[0] = last;
    int64_t i = 0, n = *len;
    for (;;) {
      int64_t l = 2 * i + 1, r = l + 1, s = i;
      if (l < n && a[l].dist < a[s].dist) s = l;
      if (r < n && a[r].dist < a[s].dist) s = r;
      if (s == i) break;
      hc_t t = a[i]; a[i] = a[s]; a[s] = t;
      i = s;
    }
  }
  return top;
}
/* BestCandidatesHeap push! (traversal.jl:33-80): max-heap of capacity cap */
static int best_push(hc_t* a, int64_t* len, int64_t cap, hc_t c) {
  if (*len < cap) {
    int64_t i = (*len)++;
    a[i] = c;
    while (i > 0) {
      int64_t p = (i - 1) / 2;
      if (a[i].dist > a[p].dist) {
        hc_t t = a[i]; a[i] = a[p]; a[p] = t;
        i = p;
      } else break;
    }
    return 1;
  }
  if (c.dist < a[0].dist) {
    a[0] = c;
    int64_t i = 0, n = *len;
    for (;;) {
      int64_t l = 2 * i + 1, r = l + 1, g = i;
      if (l < n && a[l].dist > a[g].dist) g = l;
      if (r < n && a[r].dist > a[g].dist) g = r;
      if (g == i) break;
      hc_t t = a[i]; a[i] = a[g]; a[g] = t;
      i = g;
    }
    return 1;
  }
  return 0;
}
/* stable sort by distance (MergeSort, by = c -> c.dist) */
static void stable_sort(hc_t* a, int64_t n) {
  for (int64_t i = 1; i < n; ++i) {
    hc_t v = a[i];
    int64_t p = i;
    while (p > 0 && a[p - 1].dist > v.dist) {
      a[p] = a[p - 1];
      --p;
    }
    a[p] = v;
  }
}

/* _greedy_descent (query.jl:327-345): each pass walks the neighbour list of the
 * node current at the start of the pass, moving on every strict improvement */
static void greedy(const hg_t* g, int layer, const float* q, int32_t* cur, float* cd) {
  int32_t c = *cur;
  float best = hdist(g, c, q);
  int improved = 1;
  while (improved) {
    improved = 0;
    const int32_t center = c;
    const int32_t dg = *deg_of(g, layer, center);
    const int32_t* nb = nbr_of(g, layer, center);
    for (int32_t j = 0; j < dg; ++j) {
      float dd = hdist(g, nb[j], q);
      if (dd < best) {
        c = nb[j];
        best = dd;
        improved = 1;
      }
    }
  }
  *cur = c;
  *cd = best;
}

/* _search_layer / _search_layer_pooled (query.jl:350-442): the two differ only
 * in the early exit's `length(best) >= ef` guard, which the worst distance
 * (Inf until the heap is full) already implies.  The caller has marked what it
 * wants marked before; the entry is marked here.  Returns the best heap's
 * contents sorted stably by distance into g->best. */
static int64_t search_layer(hg_t* g, int layer, hc_t entry, const float* q, int64_t ef) {
  int64_t blen = 0, plen = 0;
  best_push(g->best, &blen, ef, entry);
  min_push(g->pend, &plen, entry);
  mark(g, entry.id);
  while (plen > 0) {
    hc_t cur = min_pop(g->pend, &plen);
    float worst = blen < ef ? INFINITY : g->best[0].dist;
    if (cur.dist > worst) break;
    const int32_t dg = *deg_of(g, layer, cur.id);
    const int32_t* nb = nbr_of(g, layer, cur.id);
    for (int32_t j = 0; j < dg; ++j) {
      int32_t v = nb[j];
      if (visited(g, v)) continue;
      mark(g, v);
      float dd = hdist(g, v, q);
      float w = blen == 0 ? INFINITY : g->best[0].dist;
      if (blen < ef || dd < w) {
        hc_t c = {dd, v};
        if (best_push(g->best, &blen, ef, c)) min_push(g->pend, &plen, c);
      }
    }
  }
  stable_sort(g->best, blen);
  return blen;
}

/* select_neighbors (neighbor_policy.jl:22-36, 59-108) for a new node: res is
 * sorted; out receives at most M candidates */
static int64_t select_new(const hg_t* g, int policy, const hc_t* res, int64_t n, hc_t* out) {
  if (n == 0) return 0;
  const int64_t cap0 = g->M;
  if (policy == HN_HEURISTIC) {
    int64_t cap = cap0 < n ? cap0 : n;
    memcpy(out, res, sizeof(hc_t) * (size_t)cap);
    return cap;
  }
  if (n <= cap0) {  /* short-circuit: returned as they are */
    memcpy(out, res, sizeof(hc_t) * (size_t)n);
    return n;
  }
  const int64_t cap = cap0;
  int64_t ns = 0;
  for (int64_t ci = 0; ci < n; ++ci) {
    hc_t c = res[ci];
    int dominated = 0;
    for (int64_t j = 0; j < ns; ++j)
      if (hdist2(g, c.id, out[j].id) < c.dist) {
        dominated = 1;
        break;
      }
    if (!dominated) out[ns++] = c;
    if (ns == cap) break;
  }
  if (ns < cap) {
    for (int64_t ci = 0; ci < n && ns < cap; ++ci) {
      int dup = 0;
      for (int64_t j = 0; j < ns; ++j)
        if (out[j].id == res[ci].id) {
          dup = 1;
          break;
        }
      if (dup) continue;
      out[ns++] = res[ci];
    }
  }
  return ns;
}

/* _prune_slot! + _select_into! (query.jl:494-603): the list of `center`
 * (degree M + 1) shrinks to M ids, written over the list in place */
static void prune(const hg_t* g, int policy, int layer, int32_t center) {
  int32_t* dg = deg_of(g, layer, center);
  int32_t* nb = nbr_of(g, layer, center);
  const int64_t n = *dg, limit = g->M;
  if (n <= limit) return;
  hc_t c[512];
  for (int64_t i = 0; i < n; ++i) {
    c[i].id = nb[i];
    c[i].dist = hdist2(g, center, nb[i]);  /* distance(center_vec, data[:, id]) */
  }
  const int64_t cap = limit < n ? limit : n;
  stable_sort(c, n);
  int64_t kept = 0;
  if (policy == HN_HEURISTIC) {
    for (int64_t i = 0; i < cap; ++i) nb[i] = c[i].id;
    *dg = (int32_t)cap;
    return;
  }
  for (int64_t ci = 0; ci < n; ++ci) {
    int dominated = 0;
    for (int64_t j = 0; j < kept; ++j)
      if (hdist2(g, c[ci].id, nb[j]) < c[ci].dist) {
        dominated = 1;
        break;
      }
    if (!dominated) {
      nb[kept++] = c[ci].id;
      if (kept == cap) break;
    }
  }
  if (kept < cap) {
    for (int64_t ci = 0; ci < n; ++ci) {
      int dup = 0;
      for (int64_t j = 0; j < kept; ++j)
        if (nb[j] == c[ci].id) {
          dup = 1;
          break;
        }
      if (dup) continue;
      nb[kept++] = c[ci].id;
      if (kept == cap) break;
    }
  }
  *dg = (int32_t)kept;
}

/* insert! (query.jl:264-325) of local node i at level L; *count = nodes so far */
static void insert(hg_t* g, int policy, int64_t ef_c, int32_t i, int L, int64_t* count) {
  if (*count == 0) {
    g->entry = i;
    g->max_layer = L;
    *count = 1;
    return;
  }
  const float* q = g->X + g->rows[i] * g->ld;
  int32_t cur = g->entry;
  float cd = hdist(g, cur, q);
  const int lo = L + 1 > 1 ? L + 1 : 1;
  for (int layer = g->max_layer; layer >= lo; --layer) greedy(g, layer, q, &cur, &cd);
  hc_t sel[512];
  for (int layer = L < g->max_layer ? L : g->max_layer; layer >= 0; --layer) {
    new_visit(g);
    hc_t e = {cd, cur};
    int64_t nr = search_layer(g, layer, e, q, ef_c);
    int64_t ns = select_new(g, policy, g->best, nr, sel);
    /* _connect_new_node! (query.jl:444-469) */
    int32_t* mine = nbr_of(g, layer, i);
    for (int64_t j = 0; j < ns; ++j) mine[j] = sel[j].id;
    *deg_of(g, layer, i) = (int32_t)ns;
    for (int64_t j = 0; j < ns; ++j) {
      int32_t center = sel[j].id;
      int32_t* dg = deg_of(g, layer, center);
      nbr_of(g, layer, center)[*dg] = i;
      *dg += 1;
      if (*dg > g->M) prune(g, policy, layer, center);
    }
    if (ns > 0) {
      cur = sel[0].id;
      cd = sel[0].dist;
    }
  }
  if (L > g->max_layer) {
    g->max_layer = L;
    g->entry = i;
  }
  *count += 1;
}

/* sample_layer (planner.jl:19-30) with lambda = 1 / log(max(M, 2)) (index.jl:35):
 * levels of n points from one splitmix64 stream */
int orc_hnsw_levels(int64_t n, int64_t M, uint64_t seed, int32_t* out_level) {
  if (M <= 0) return -2;
  const double lam = 1.0 / log((double)(M > 2 ? M : 2));
  uint64_t st = seed;
  for (int64_t i = 0; i < n; ++i) {
    double u = orc_uniform(&st);
    if (u < DBL_EPSILON) u = DBL_EPSILON;
    out_level[i] = (int32_t)floor(-log(u) * lam);
  }
  return 0;
}

static int hg_alloc(hg_t* g, int64_t n, int64_t ef) {
  g->stamp = (uint32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(uint32_t));
  g->best = (hc_t*)malloc(sizeof(hc_t) * (size_t)(ef > 0 ? ef : 1));
  g->pend = (hc_t*)malloc(sizeof(hc_t) * (size_t)(n + 1));
  g->gen = 0;
  g->ncap = n > 0 ? n : 1;
  return g->stamp && g->best && g->pend ? 0 : -5;
}
static void hg_free(hg_t* g) {
  free(g->stamp);
  free(g->best);
  free(g->pend);
}

/* The terminals of build_ivf_hnsw_index: one HNSW graph per non-empty bucket
 * over its points (ids ascending, multilevel/builder.jl:133-146), built by
 * serial insertion.  offsets/ids: the CSR of orc_partition (1-based ids);
 * level: per row of the list-contiguous order; nbr/deg sized nlayers x n rows
 * (nlayers > every level), zero-initialised by the caller; entry/max_layer per
 * bucket (-1 for an empty bucket). */
int orc_ivf_hnsw_build(const float* X, int64_t d, int64_t ld, const int64_t* offsets,
                       const int64_t* ids, int64_t nb, int64_t M, int64_t ef_c, int policy,
                       const int32_t* level, int64_t nlayers, int32_t* nbr, int32_t* deg,
                       int32_t* entry, int32_t* max_layer, int nthreads) {
  if (M <= 0 || ef_c < 1 || M + 1 > 512) return -2;
  const int64_t n_all = offsets[nb];
  for (int64_t r = 0; r < n_all; ++r)
    if (level[r] < 0 || level[r] >= nlayers) return -2;
  int err = 0;
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t lo = offsets[b], n = offsets[b + 1] - lo;
    entry[b] = -1;
    max_layer[b] = -1;
    if (n == 0) continue;
    int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) rows[i] = ids[lo + i] - 1;
    hg_t g;
    memset(&g, 0, sizeof g);
    g.X = X;
    g.d = d;
    g.ld = ld;
    g.rows = rows;
    g.n = n;
    g.M = M;
    g.cap = M + 1;
    g.nl_all = n_all;
    g.nbr = nbr + lo * (M + 1);
    g.deg = deg + lo;
    g.entry = -1;
    g.max_layer = -1;
    if (hg_alloc(&g, n, ef_c)) {
      err = 1;
    } else {
      int64_t count = 0;
      for (int64_t i = 0; i < n; ++i) insert(&g, policy, ef_c, (int32_t)i, level[lo + i], &count);
      entry[b] = g.entry;
      max_layer[b] = g.max_layer;
    }
    hg_free(&g);
    free(rows);
  }
  return err ? -5 : 0;
}

/* _query_search (query.jl:95-123) on one list's graph: the entry point is
 * marked visited, greedy descent over the upper layers, the base-layer search
 * from the node reached (marked too), the first min(k, |results|) by distance.
 * Returns the count; out_local 0-based local ids. */
static int64_t hnsw_query(hg_t* g, const float* q, int64_t k, int64_t ef_search,
                          int32_t* out_local, float* out_dist) {
  if (k <= 0 || g->n == 0) return 0;
  const int64_t ef = ef_search > k ? ef_search : k;
  new_visit(g);
  int32_t cur = g->entry;
  float cd = hdist(g, cur, q);
  mark(g, cur);
  for (int layer = g->max_layer; layer >= 1; --layer) greedy(g, layer, q, &cur, &cd);
  hc_t e = {cd, cur};
  int64_t nr = search_layer(g, 0, e, q, ef);
  int64_t lim = k < nr ? k : nr;
  for (int64_t i = 0; i < lim; ++i) {
    out_local[i] = g->best[i].id;
    out_dist[i] = g->best[i].dist;
  }
  return lim;
}

/* query(::MultiLevelIndex{KMeans, TopK, HNSW}, data, Q, k) (multilevel/query.jl:
 * 279-512; single-query form :25-106 with probe_order = 1): TopK over the
 * routing distances, empty buckets dropped, each probed list's HNSW query with
 * k and ef_search (ef = max(ef_search, k)), per-list results concatenated in
 * bucket order (batched) or probe order (single query), then the merge
 * (merge_kind 0 SimpleMerge, 1 DisjointMerge, merge.jl:79-169).  Output k
 * slots per query, padded with id 0 / +inf. */
int orc_ivf_hnsw_search(const float* X, int64_t d, int64_t ld, const int64_t* offsets,
                        const int64_t* ids, int64_t nb, int64_t M, const int32_t* nbr,
                        const int32_t* deg, const int32_t* entry, const int32_t* max_layer,
                        const float* Q, int64_t nq, int64_t ldq, const float* route,
                        int64_t nlist, int64_t k, int64_t nprobe, int64_t ef_search,
                        int merge_kind, int probe_order, int64_t* out_ids, float* out_dist,
                        int nthreads) {
  if (k <= 0 || nq == 0) return 0;
  if (nprobe <= 0 || ef_search < 1) return -2;
  const int64_t n_all = offsets[nb];
  int64_t max_n = 1;
  for (int64_t b = 0; b < nb; ++b)
    if (offsets[b + 1] - offsets[b] > max_n) max_n = offsets[b + 1] - offsets[b];
  const int64_t ef_max = ef_search > k ? ef_search : k;
  int err = 0;
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
  {
    int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)max_n);
    int64_t* sel = (int64_t*)malloc(sizeof(int64_t) * (size_t)nprobe);
    int64_t* cid = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nprobe * k));
    float* cdv = (float*)malloc(sizeof(float) * (size_t)(nprobe * k));
    int32_t* li = (int32_t*)malloc(sizeof(int32_t) * (size_t)k);
    float* ld2 = (float*)malloc(sizeof(float) * (size_t)k);
    hg_t g;
    memset(&g, 0, sizeof g);
    if (hg_alloc(&g, max_n, ef_max)) err = 1;
#pragma omp for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < nq; ++qi) {
      if (err) continue;
      int64_t ns = orc_select_topk(route + qi * nlist, nlist, nb, nprobe, sel);
      if (ns < 0) {
        err = 2;
        ns = 0;
      }
      int64_t m = 0;
      for (int64_t s = 0; s < ns; ++s)
        if (offsets[sel[s]] > offsets[sel[s] - 1]) sel[m++] = sel[s];
      if (!probe_order) /* ascending bucket */
        for (int64_t i = 1; i < m; ++i) {
          int64_t v = sel[i], p = i;
          while (p > 0 && sel[p - 1] > v) {
            sel[p] = sel[p - 1];
            --p;
          }
          sel[p] = v;
        }
      int64_t tot = 0;
      const float* q = Q + qi * ldq;
      for (int64_t s = 0; s < m; ++s) {
        const int64_t b = sel[s] - 1, lo = offsets[b], n = offsets[b + 1] - lo;
        for (int64_t i = 0; i < n; ++i) rows[i] = ids[lo + i] - 1;
        g.X = X;
        g.d = d;
        g.ld = ld;
        g.rows = rows;
        g.n = n;
        g.M = M;
        g.cap = M + 1;
        g.nl_all = n_all;
        g.nbr = (int32_t*)nbr + lo * (M + 1);
        g.deg = (int32_t*)deg + lo;
        g.entry = entry[b];
        g.max_layer = max_layer[b];
        /* the stamps are indexed by local id: a fresh generation per search */
        int64_t got = hnsw_query(&g, q, k, ef_search, li, ld2);
        for (int64_t r = 0; r < got; ++r) {
          cid[tot] = ids[lo + li[r]];
          cdv[tot] = ld2[r];
          ++tot;
        }
      }
      int64_t* oi = out_ids + qi * k;
      float* od = out_dist + qi * k;
      int64_t got = merge_kind == 0 ? orc_simple_merge(cid, cdv, tot, k, oi, od)
                                    : orc_disjoint_merge(cid, cdv, tot, k, oi, od);
      for (int64_t r = got; r < k; ++r) {
        oi[r] = 0;
        od[r] = INFINITY;
      }
    }
    hg_free(&g);
    free(rows);
    free(sel);
    free(cid);
    free(cdv);
    free(li);
    free(ld2);
  }
  return err == 1 ? -5 : err ? -2 : 0;
}
