// IVF-HNSW terminals on gfx950: the serial HNSW insertion of every list, the
// per-(query, list) HNSW search and the merge of a query's per-list results.
//
// Reference (file:line relative to the reference tree):
//   src/indices/hnsw/query.jl:264-325   insert!
//   src/indices/hnsw/query.jl:327-345   _greedy_descent
//   src/indices/hnsw/query.jl:350-442   _search_layer(_pooled)
//   src/indices/hnsw/query.jl:444-603   _connect_new_node!, _prune_slot!, _select_into!
//   src/indices/hnsw/query.jl:95-123    _query_search
//   src/indices/hnsw/neighbor_policy.jl:22-112, traversal.jl:33-130,
//   candidate_heap.jl:15-73             selection policies and the two heaps
//   src/indices/multilevel/merge.jl:79-169  SimpleMerge / DisjointMerge
//
// A graph search is sequential by nature: one wavefront follows one search.
// Lane 0 owns the heaps (best: max-heap of capacity ef in LDS; pending:
// min-heap in LDS, moved to a global region when it outgrows it -- every node
// enters it at most once, so |list| + 1 entries always suffice).  The wave
// works in parallel where the reference loops over independent items: the
// visited test of a node's neighbours (one lane each, atomic OR into the
// bitmap: the ids of a list are distinct, so the order of marking does not
// matter), their distances (four 16-lane groups, the oracle's A16 order), the
// stable sort of the results (a rank per element) and the dominance tests of
// the diversified selection (all distances to the kept set, then a ballot).
// The pushes themselves run in the reference's order on lane 0.
#include "bounds.h"
#include "hnsw.h"

namespace mann {

namespace {

constexpr int B_PEND = 4096;  // build: pending-heap entries in LDS
constexpr int B_VIS = 1024;   // build: visited words in LDS (lists of up to 32,768 rows)

// Memory operations of the workgroup's single wavefront before this point are
// visible to all of its lanes after it.
__device__ __forceinline__ void wg_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ __forceinline__ int bcast(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float bcastf(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// NeighborMinHeap push! / popfirst! (candidate_heap.jl:24-73): sift by strict
// '<' on the distance, written with a moving hole instead of swaps
__device__ void min_push(HnCand* a, int& len, HnCand c) {
  int i = len++;
  while (i > 0) {
    const int p = (i - 1) >> 1;
    const HnCand ap = a[p];
    if (c.d < ap.d) {
      a[i] = ap;
      i = p;
    } else {
      break;
    }
  }
  a[i] = c;
}
__device__ HnCand min_pop(HnCand* a, int& len) {
  const HnCand top = a[0];
  const HnCand last = a[--len];
  if (len > 0) {
    int i = 0;
    for (;;) {
      const int l = 2 * i + 1, r = l + 1;
      int s = i;
      float sd = last.d;
      if (l < len && a[l].d < sd) {
        s = l;
        sd = a[l].d;
      }
      if (r < len && a[r].d < sd) s = r;
      if (s == i) break;
      a[i] = a[s];
      i = s;
    }
    a[i] = last;
  }
  return top;
}
// BestCandidatesHeap push! (traversal.jl:33-80): max-heap of capacity cap; a
// full heap takes a candidate only if it is strictly closer than its root
__device__ bool best_push(HnCand* a, int& len, int cap, HnCand c) {
  if (len < cap) {
    int i = len++;
    while (i > 0) {
      const int p = (i - 1) >> 1;
      const HnCand ap = a[p];
      if (c.d > ap.d) {
        a[i] = ap;
        i = p;
      } else {
        break;
      }
    }
    a[i] = c;
    return true;
  }
  if (c.d < a[0].d) {
    int i = 0;
    for (;;) {
      const int l = 2 * i + 1, r = l + 1;
      int g = i;
      float gd = c.d;
      if (l < len && a[l].d > gd) {
        g = l;
        gd = a[l].d;
      }
      if (r < len && a[r].d > gd) g = r;
      if (g == i) break;
      a[i] = a[g];
      i = g;
    }
    a[i] = c;
    return true;
  }
  return false;
}

__device__ __forceinline__ int64_t list_slot(const HnGraph& G, int layer, int64_t row) {
  return layer == 0 ? row : (int64_t)(layer - 1) * G.n1 + G.upslot[row];
}
__device__ __forceinline__ int* list_of(const HnGraph& G, int layer, int64_t slot) {
  return (layer == 0 ? G.nbr0 : G.nbru) + slot * G.cap;
}
__device__ __forceinline__ int* deg_of(const HnGraph& G, int layer, int64_t slot) {
  return (layer == 0 ? G.deg0 : G.degu) + slot;
}

// LDS scratch of one wavefront: the ids a step evaluates and their distances
struct Work {
  int id[64];
  float d[64];
};

// Euclidean distances (the A16 order of the oracle) of rows row0 + ids[i],
// i < cnt <= 64, to the vector qv (d16 floats, zero padded): out[i].  Every
// lane takes part (the group tree reads all lanes of its 16-lane row).  CH:
// chunks whose loads are issued together (the build, one wavefront per SIMD,
// affords 16: one memory round trip per pass at d <= 256; the search, which
// needs occupancy, 8).  One pass: 4 U rows from ids[b], U per 16-lane group.
template <int CH, int U>
__device__ __forceinline__ void dist_rows(const float* __restrict__ x32, int64_t row0, int d16,
                                          const float* __restrict__ qv, const int* ids, int cnt,
                                          int b, float* out, int lane) {
  const int g = lane >> 4, j = lane & 15;
  {
    const float* xr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = b + g + 4 * u;
      xr[u] = x32 + (row0 + ids[i < cnt ? i : b]) * (int64_t)d16;
    }
    float acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = 0.0f;
    // the loads of CH chunks are issued before their multiply-adds (the rows
    // are scattered: each wait is a full memory latency); the adds keep the
    // chunk order of the A16 sum
    const int nch = d16 >> 4;
    int c0 = 0;
    for (; c0 + CH <= nch; c0 += CH) {
      float xv[CH][U], qq[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int t = ((c0 + c) << 4) + j;
        qq[c] = qv[t];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[c][u] = xr[u][t];
      }
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const float df = xv[c][u] - qq[c];
          acc[u] = __builtin_fmaf(df, df, acc[u]);
        }
    }
    if (CH > 8)
      for (; c0 + 8 <= nch; c0 += 8) {
        float xv[8][U], qq[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int t = ((c0 + c) << 4) + j;
          qq[c] = qv[t];
#pragma unroll
          for (int u = 0; u < U; ++u) xv[c][u] = xr[u][t];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const float df = xv[c][u] - qq[c];
            acc[u] = __builtin_fmaf(df, df, acc[u]);
          }
      }
    for (; c0 < nch; ++c0) {
      const int t = (c0 << 4) + j;
      const float qq = qv[t];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float df = xr[u][t] - qq;
        acc[u] = __builtin_fmaf(df, df, acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const float s = group16_tree(acc[u]);
      const int i = b + g + 4 * u;
      if (j == 0 && i < cnt) out[i] = __builtin_sqrtf(s);
    }
  }
}

template <int CH>
__device__ void dist_batch(const float* __restrict__ x32, int64_t row0, int d16,
                           const float* __restrict__ qv, const int* ids, int cnt, float* out,
                           int lane) {
  // U rows per 16-lane group and pass, from what is left: a pass costs one
  // memory latency whatever its width, and a wider one than the rows left
  // only issues loads of a repeated row (the diversified rule's tests against
  // a kept set of 1-16 rows, the late steps of a walk with few unvisited
  // neighbours).  Passes of 8 rows only with whole-row load batches (the
  // build); the search, held at 96 registers, takes 16 or 4.  (Passes of 32
  // rows, 128 loads per lane, slowed the 1M x 256 build from 10.4 to 12.3 s:
  // a wavefront tracks at most 63 loads in flight.)
  for (int b = 0; b < cnt;) {
    const int rem = cnt - b;
    if (rem > 8) {
      dist_rows<CH, 4>(x32, row0, d16, qv, ids, cnt, b, out, lane);
      b += 16;
    } else if (CH >= 16 && rem > 4) {
      dist_rows<CH, 2>(x32, row0, d16, qv, ids, cnt, b, out, lane);
      b += 8;
    } else {
      dist_rows<CH, 1>(x32, row0, d16, qv, ids, cnt, b, out, lane);
      b += 4;
    }
  }
}

// _greedy_descent (query.jl:327-345): each pass scores the neighbour list of
// the node current at its start and moves on every strict improvement
template <int CH>
__device__ HnCand greedy(const HnGraph& G, int layer, int64_t row0, const float* x32, int d16,
                         const float* qv, HnCand cur, Work& w, int lane, int* ndist = nullptr) {
  int c = cur.id;
  float bd = cur.d;
  for (;;) {
    const int64_t s = list_slot(G, layer, row0 + c);
    // the whole slot (cap ids) is read together with the degree
    const int* lst = list_of(G, layer, s);
    const int v = lane < G.cap ? lst[lane] : 0;
    const int deg = *deg_of(G, layer, s);
    if (lane < deg) w.id[lane] = v;
    wg_sync();
    dist_batch<CH>(x32, row0, d16, qv, w.id, deg, w.d, lane);
    if (ndist) *ndist += deg;
    wg_sync();
    int nc = c, imp = 0;
    float nd = bd;
    if (lane == 0)
      for (int t = 0; t < deg; ++t)
        if (w.d[t] < nd) {
          nd = w.d[t];
          nc = w.id[t];
          imp = 1;
        }
    c = bcast(nc);
    bd = bcastf(nd);
    wg_sync();
    if (!bcast(imp)) break;
  }
  return HnCand{bd, c};
}

// _search_layer (query.jl:350-442) from `entry` (marked here; the caller has
// cleared the bitmap or marked what it must): best-first expansion until the
// popped candidate is farther than the worst of a full best heap.  Returns
// the best heap's size; its entries are best[0 .. size).
template <int CH>
__device__ int search_layer(const HnGraph& G, int layer, int64_t row0, const float* x32, int d16,
                            const float* qv, HnCand entry, int ef, HnCand* best, HnCand* pl,
                            int plcap, HnCand* pg, unsigned* vis, Work& w, int lane,
                            int* ndist = nullptr) {
  int blen = 0, plen = 0;
  HnCand* pend = pl;
  if (lane == 0) {
    best_push(best, blen, ef, entry);
    min_push(pend, plen, entry);
    vis[entry.id >> 5] |= 1u << (entry.id & 31);
  }
  wg_sync();
  for (;;) {
    int stop = 0, cid = 0;
    if (lane == 0) {
      if (plen == 0) {
        stop = 1;
      } else {
        const HnCand cur = min_pop(pend, plen);
        const float worst = blen < ef ? INFINITY : best[0].d;
        if (cur.d > worst) stop = 1;
        cid = cur.id;
      }
    }
    if (bcast(stop)) break;
    cid = bcast(cid);
    // the unvisited neighbours of cid, in list order, marked visited
    const int64_t s = list_slot(G, layer, row0 + cid);
    // the whole slot (cap ids) is read together with the degree
    const int* lst = list_of(G, layer, s);
    const int v = lane < G.cap ? lst[lane] : 0;
    const int deg = *deg_of(G, layer, s);
    bool fresh = false;
    if (lane < deg) {
      const unsigned bit = 1u << (v & 31);
      fresh = (atomicOr(vis + (v >> 5), bit) & bit) == 0u;
    }
    const uint64_t m = __ballot(fresh);
    const int cnt = __popcll(m);
    if (fresh) w.id[__popcll(m & ((1ull << lane) - 1ull))] = v;
    wg_sync();
    dist_batch<CH>(x32, row0, d16, qv, w.id, cnt, w.d, lane);
    if (ndist) *ndist += cnt;
    wg_sync();
    if (lane == 0) {
      for (int t = 0; t < cnt; ++t) {
        const float dd = w.d[t];
        const float wst = blen == 0 ? INFINITY : best[0].d;
        if (blen < ef || dd < wst) {
          const HnCand c{dd, w.id[t]};
          if (best_push(best, blen, ef, c)) {
            if (pend == pl && plen == plcap) {  // outgrown LDS: continue in global memory
              for (int u = 0; u < plen; ++u) pg[u] = pl[u];
              pend = pg;
            }
            min_push(pend, plen, c);
          }
        }
      }
    }
    wg_sync();
  }
  return bcast(blen);
}

// dst[rank(i)] = src[i] for the stable order by distance (MergeSort,
// query.jl:400/439); only ranks below lim are written
__device__ void stable_sort_into(const HnCand* src, int n, HnCand* dst, int lim, int lane) {
  for (int i = lane; i < n; i += 64) {
    const HnCand c = src[i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const float dj = src[j].d;
      r += (dj < c.d || (dj == c.d && j < i)) ? 1 : 0;
    }
    if (r < lim) dst[r] = c;
  }
}

// The diversified rule (neighbor_policy.jl:59-108, _select_into! query.jl:550-603)
// over cands[0 .. n) (sorted; ids and distances to the centre): keep a
// candidate unless a kept one is strictly closer to it than the centre, stop at
// cap, then refill from the rejected in order.  kept/keptd receive the result
// (keptd may be null); returns its size.
template <int CH>
__device__ int diversify(const float* x32, int64_t row0, int d16, const int* cid, const float* cd,
                         int n, int cap, int* kept, float* keptd, Work& w, int lane) {
  int nk = 0;
  for (int ci = 0; ci < n && nk < cap; ++ci) {
    dist_batch<CH>(x32, row0, d16, x32 + (row0 + cid[ci]) * (int64_t)d16, kept, nk, w.d, lane);
    wg_sync();
    const bool dom = lane < nk && w.d[lane] < cd[ci];
    if (__ballot(dom) == 0) {
      if (lane == 0) {
        kept[nk] = cid[ci];
        if (keptd) keptd[nk] = cd[ci];
      }
      ++nk;
    }
    wg_sync();
  }
  if (nk < cap) {  // keepPrunedConnections: the rejected, in distance order
    int m = nk;
    if (lane == 0) {
      for (int ci = 0; ci < n && m < cap; ++ci) {
        bool dup = false;
        for (int j = 0; j < m; ++j) dup |= kept[j] == cid[ci];
        if (dup) continue;
        kept[m] = cid[ci];
        if (keptd) keptd[m] = cd[ci];
        ++m;
      }
    }
    nk = bcast(m);
    wg_sync();
  }
  return nk;
}

__global__ void __launch_bounds__(64) hnsw_build_kernel(HnswBuildArgs a) {
  __shared__ HnCand best[HN_EF_MAX];
  __shared__ HnCand pl[B_PEND];
  __shared__ unsigned visl[B_VIS];
  __shared__ int rid[HN_EF_MAX];       // sorted results: ids, distances
  __shared__ float rd[HN_EF_MAX];
  __shared__ int sel[64];              // selected neighbours of the new node
  __shared__ float seld[64];
  __shared__ int pid[64];              // a list being pruned, then sorted
  __shared__ float pdist[64];
  __shared__ int sid[64];
  __shared__ float sdist[64];
  __shared__ int kept[64];
  __shared__ Work w;
  const int lane = threadIdx.x;
  const int b = a.lorder[blockIdx.x];
  const int64_t row0 = a.list_off[b];
  const int n = a.list_off[b + 1] - (int)row0;
  if (n <= 0) {
    if (lane == 0) {
      a.entry[b] = -1;
      a.maxl[b] = -1;
    }
    return;
  }
  const HnGraph& G = a.g;
  const int nw = (n + 31) >> 5;
  unsigned* vis = nw <= B_VIS ? visl : a.vis_g + (row0 >> 5) + b;
  HnCand* pg = a.pend_g + row0 + b;
  const int M = a.M, d16 = a.d16;
  const float* x32 = a.x32;
  int ep = 0, maxL = a.level[row0];  // the first node: entry point, nothing to link
  for (int i = 1; i < n; ++i) {
    const int64_t ri = row0 + i;
    const int L = a.level[ri];
    const float* qv = x32 + ri * d16;
    if (lane == 0) w.id[0] = ep;
    wg_sync();
    dist_batch<16>(x32, row0, d16, qv, w.id, 1, w.d, lane);
    wg_sync();
    HnCand cur{w.d[0], ep};
    wg_sync();
    for (int layer = maxL; layer >= (L + 1 > 1 ? L + 1 : 1); --layer)
      cur = greedy<16>(G, layer, row0, x32, d16, qv, cur, w, lane);
    for (int layer = L < maxL ? L : maxL; layer >= 0; --layer) {
      for (int t = lane; t < nw; t += 64) vis[t] = 0u;
      wg_sync();
      const int nr = search_layer<16>(G, layer, row0, x32, d16, qv, cur, a.ef, best, pl, B_PEND,
                                      pg, vis, w, lane);
      for (int t = lane; t < nr; t += 64) {
        const HnCand c = best[t];
        int r = 0;
        for (int j = 0; j < nr; ++j) {
          const float dj = best[j].d;
          r += (dj < c.d || (dj == c.d && j < t)) ? 1 : 0;
        }
        rid[r] = c.id;
        rd[r] = c.d;
      }
      wg_sync();
      // select_neighbors (neighbor_policy.jl:22-108): the first M for the
      // heuristic policy or a short list, else the diversified rule
      int ns;
      if (a.policy == 0 || nr <= M) {
        ns = nr < M ? nr : M;
        if (lane < ns) {
          sel[lane] = rid[lane];
          seld[lane] = rd[lane];
        }
        wg_sync();
      } else {
        ns = diversify<16>(x32, row0, d16, rid, rd, nr, M, sel, seld, w, lane);
      }
      // _connect_new_node! (query.jl:444-469): the node's own list, then the
      // reverse edges in selection order, each pruned when over M
      {
        const int64_t s = list_slot(G, layer, ri);
        int* mine = list_of(G, layer, s);
        if (lane < ns) mine[lane] = sel[lane];
        if (lane == 0) *deg_of(G, layer, s) = ns;
        wg_sync();
      }
      for (int jn = 0; jn < ns; ++jn) {
        const int center = sel[jn];
        const int64_t cs = list_slot(G, layer, row0 + center);
        int* clst = list_of(G, layer, cs);
        int* cdeg = deg_of(G, layer, cs);
        const int dg = *cdeg + 1;
        if (lane == 0) {
          clst[dg - 1] = i;
          *cdeg = dg;
        }
        wg_sync();
        if (dg <= M) continue;
        // _prune_slot! (query.jl:494-529): distances from the centre, stable
        // sort, selection written over the list
        if (lane < dg) pid[lane] = clst[lane];
        wg_sync();
        dist_batch<16>(x32, row0, d16, x32 + (row0 + center) * (int64_t)d16, pid, dg, pdist,
                       lane);
        wg_sync();
        if (lane < dg) {
          const float dl = pdist[lane];
          int r = 0;
          for (int j = 0; j < dg; ++j) {
            const float dj = pdist[j];
            r += (dj < dl || (dj == dl && j < lane)) ? 1 : 0;
          }
          sid[r] = pid[lane];
          sdist[r] = dl;
        }
        wg_sync();
        const int cap = M < dg ? M : dg;
        int nk;
        if (a.policy == 0) {
          nk = cap;
          if (lane < nk) kept[lane] = sid[lane];
          wg_sync();
        } else {
          nk = diversify<16>(x32, row0, d16, sid, sdist, dg, cap, kept, nullptr, w, lane);
        }
        if (lane < nk) clst[lane] = kept[lane];
        if (lane == 0) *cdeg = nk;
        wg_sync();
      }
      if (ns > 0) cur = HnCand{seld[0], sel[0]};
      wg_sync();
    }
    if (L > maxL) {
      maxL = L;
      ep = i;
    }
  }
  if (lane == 0) {
    a.entry[b] = ep;
    a.maxl[b] = maxL;
  }
}

// _query_search (query.jl:95-123) of query q in list b for every (q, probe)
// pair: the entry point is marked visited, the upper layers are descended
// greedily, the base layer is searched from the node reached (ef), and the
// first min(k, results) in distance order are written.
// (at most 96 VGPRs: 5 wavefronts per SIMD, no spills)
__global__ void __launch_bounds__(64, 5) hnsw_search_kernel(HnswSearchArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  HnCand* best = (HnCand*)smem;      // ef
  HnCand* res = best + a.ef;          // ef
  HnCand* pl = res + a.ef;            // pend_lds
  Work& w = *(Work*)(pl + a.pend_lds);
  unsigned* visl = (unsigned*)(&w + 1);  // vis_words
  const int lane = threadIdx.x;
  HnCand* pg = a.pend_g + (int64_t)blockIdx.x * a.pend_stride;
  unsigned* visg = a.vis_g + (int64_t)blockIdx.x * a.vis_stride;
  const HnGraph& G = a.g;
  const int64_t np = (int64_t)a.nq * a.nprobe;
  for (int64_t p = blockIdx.x; p < np; p += gridDim.x) {
    const int q = (int)(p / a.nprobe), r = (int)(p % a.nprobe);
    if (r >= a.pcount[q]) {
      if (lane == 0) a.out_cnt[p] = 0;
      continue;
    }
    const int b = a.probes[p];
    const int64_t row0 = a.list_off[b];
    const int n = a.list_off[b + 1] - (int)row0;
    const int nw = (n + 31) >> 5;
    unsigned* vis = nw <= a.vis_words ? visl : visg;
    for (int t = lane; t < nw; t += 64) vis[t] = 0u;
    const float* qv = a.q32 + (int64_t)q * a.d16;
    const int ep = a.entry[b];
    wg_sync();
    if (lane == 0) {
      w.id[0] = ep;
      vis[ep >> 5] |= 1u << (ep & 31);
    }
    wg_sync();
    dist_batch<8>(a.x32, row0, a.d16, qv, w.id, 1, w.d, lane);
    wg_sync();
    HnCand cur{w.d[0], ep};
    wg_sync();
    int nd = 1;
    for (int layer = a.maxl[b]; layer >= 1; --layer)
      cur = greedy<8>(G, layer, row0, a.x32, a.d16, qv, cur, w, lane, &nd);
    const int nr = search_layer<8>(G, 0, row0, a.x32, a.d16, qv, cur, a.ef, best, pl, a.pend_lds,
                                   pg, vis, w, lane, &nd);
    if (lane == 0 && a.ndist) atomicAdd(a.ndist, (unsigned long long)nd);
    const int lim = a.k < nr ? a.k : nr;
    stable_sort_into(best, nr, res, lim, lane);
    wg_sync();
    int* orow = a.out_row + p * a.k;
    float* od = a.out_d + p * a.k;
    for (int t = lane; t < lim; t += 64) {
      orow[t] = (int)row0 + res[t].id;
      od[t] = res[t].d;
    }
    if (lane == 0) a.out_cnt[p] = lim;
    wg_sync();
  }
}

// merge_results over a query's per-list results (merge.jl:79-169), one thread
// per query: the lists in bucket order (the batched child order) or probe
// order (the single-query path); DisjointMerge pushes every entry through the
// bounded max-heap of neighbor_heaps.jl:84-100 and sorts the survivors by
// (dist, id); SimpleMerge is the stable order by distance of the
// concatenation (a merge of the sorted lists preferring the earlier list).
__global__ void hnsw_merge_kernel(HnswMergeArgs a) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.nq) return;
  const int P = a.pcount[q], k = a.k, np = a.nprobe;
  const int* pr = a.probes + (int64_t)q * np;
  const int* cnt = a.out_cnt + (int64_t)q * np;
  int* ord = a.scratch + (int64_t)q * 2 * np;  // list order (probe slots)
  int* cur = ord + np;                       // SimpleMerge cursors
  int last = -1;
  for (int s = 0; s < P; ++s) {
    if (a.probe_order) {
      ord[s] = s;
      continue;
    }
    int bs = -1;
    for (int r = 0; r < P; ++r)
      if (pr[r] > last && (bs < 0 || pr[r] < pr[bs])) bs = r;
    ord[s] = bs;
    last = pr[bs];
  }
  int64_t* oi = a.ids + (int64_t)q * k;
  float* od = a.dist + (int64_t)q * k;
  const int64_t base = (int64_t)q * np * k;
  int got = 0;
  if (a.merge == 1) {
    int tot = 0;
    for (int s = 0; s < P; ++s) tot += cnt[ord[s]];
    const int cap = k < tot ? k : tot;
    HnCand* h = a.heap + (int64_t)q * k;
    int len = 0;
    for (int s = 0; s < P; ++s) {
      const int r = ord[s];
      for (int t = 0; t < cnt[r]; ++t) {
        const int64_t e = base + (int64_t)r * k + t;
        const HnCand c{a.out_d[e], a.perm[a.out_row[e]]};
        if (len < cap) {
          int x = len++;
          while (x > 0) {
            const int p = (x - 1) >> 1;
            if (c.d > h[p].d) {
              h[x] = h[p];
              x = p;
            } else {
              break;
            }
          }
          h[x] = c;
        } else if (cap > 0 && c.d < h[0].d) {
          int x = 0;
          for (;;) {
            const int l = 2 * x + 1, rr = l + 1;
            int g = x;
            float gd = c.d;
            if (l < len && h[l].d > gd) {
              g = l;
              gd = h[l].d;
            }
            if (rr < len && h[rr].d > gd) g = rr;
            if (g == x) break;
            h[x] = h[g];
            x = g;
          }
          h[x] = c;
        }
      }
    }
    // to_sorted_vector: by (dist, id)
    for (int i = 1; i < len; ++i) {
      const HnCand v = h[i];
      int p = i;
      while (p > 0 && (h[p - 1].d > v.d || (h[p - 1].d == v.d && h[p - 1].id > v.id))) {
        h[p] = h[p - 1];
        --p;
      }
      h[p] = v;
    }
    for (int i = 0; i < len; ++i) {
      oi[i] = (int64_t)h[i].id + 1;
      od[i] = h[i].d;
    }
    got = len;
  } else {
    for (int s = 0; s < P; ++s) cur[ord[s]] = 0;
    while (got < k) {
      int br = -1;
      float bd = 0.0f;
      for (int s = 0; s < P; ++s) {
        const int r = ord[s];
        if (cur[r] >= cnt[r]) continue;
        const float dv = a.out_d[base + (int64_t)r * k + cur[r]];
        if (br < 0 || dv < bd) {
          br = r;
          bd = dv;
        }
      }
      if (br < 0) break;
      const int64_t e = base + (int64_t)br * k + cur[br];
      oi[got] = (int64_t)a.perm[a.out_row[e]] + 1;
      od[got] = bd;
      ++cur[br];
      ++got;
    }
  }
  for (int i = got; i < k; ++i) {
    oi[i] = 0;
    od[i] = INFINITY;
  }
}

}  // namespace

void launch_hnsw_build(const HnswBuildArgs& a, int grid, hipStream_t st) {
  if (grid <= 0) return;
  hipLaunchKernelGGL(hnsw_build_kernel, dim3(grid), dim3(64), 0, st, a);
}

size_t hnsw_search_lds(int ef, int pend_lds, int vis_words) {
  return (size_t)2 * ef * sizeof(HnCand) + (size_t)pend_lds * sizeof(HnCand) + sizeof(Work) +
         (size_t)vis_words * sizeof(unsigned);
}

void launch_hnsw_search(const HnswSearchArgs& a, int grid, hipStream_t st) {
  if (grid <= 0) return;
  hipLaunchKernelGGL(hnsw_search_kernel, dim3(grid), dim3(64),
                     hnsw_search_lds(a.ef, a.pend_lds, a.vis_words), st, a);
}

void launch_hnsw_merge(const HnswMergeArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  hipLaunchKernelGGL(hnsw_merge_kernel, dim3((unsigned)ceil_div(a.nq, 64)), dim3(64), 0, st, a);
}

}  // namespace mann
