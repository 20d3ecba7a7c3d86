// IVF-HNSW: the HNSW terminals of build_ivf_hnsw_index
// (src/indices/multilevel/ivf_hnsw.jl:46-88 -> TerminalConfig(HNSWIndex, (M,
// ef_construction, ef_search, neighbor_policy)), multilevel/builder.jl:232-239)
// built and searched on the GPU over the list-contiguous rows of an IVF handle.
#include <cfloat>
#include <cmath>

#include "core.h"
#include "hnsw.h"
#include "util.h"

namespace mann {

namespace {

uint64_t splitmix_next(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

HnGraph graph_of(const mann_ivf& iv) {
  HnGraph g{};
  g.nbr0 = iv.hn_nbr0;
  g.deg0 = iv.hn_deg0;
  g.upslot = iv.hn_upslot;
  g.nbru = iv.hn_nbru;
  g.degu = iv.hn_degu;
  g.n1 = iv.hn_n1;
  g.cap = iv.hn_M + 1;
  return g;
}

template <class T>
T* dev_alloc_zero(size_t count, hipStream_t st) {
  T* p = nullptr;
  MANN_HIP(hipMalloc(&p, sizeof(T) * (count > 0 ? count : 1)));
  MANN_HIP(hipMemsetAsync(p, 0, sizeof(T) * (count > 0 ? count : 1), st));
  return p;
}

}  // namespace

void ivf_hnsw_build(mann_ivf& iv, int M, int ef_c, int policy, uint64_t seed) {
  MANN_REQUIRE(!iv.flat, ERR_ARGUMENT,
               "HNSW terminals belong to a MultiLevelIndex, not an IVFFlatIndex handle");
  MANN_REQUIRE(!iv.normalized, ERR_ARGUMENT,
               "HNSW terminals search with the Euclidean distance (the HNSWIndex default)");
  MANN_REQUIRE(M > 0, ERR_ARGUMENT, "M must be positive");
  MANN_REQUIRE(M <= HN_M_MAX, ERR_ARGUMENT,
               "M = " + std::to_string(M) + " exceeds the GPU build's maximum of " +
                   std::to_string(HN_M_MAX));
  MANN_REQUIRE(ef_c >= 1, ERR_ARGUMENT, "ef_construction must be at least 1");
  MANN_REQUIRE(ef_c <= HN_EF_MAX, ERR_ARGUMENT,
               "ef_construction = " + std::to_string(ef_c) +
                   " exceeds the GPU build's maximum of " + std::to_string(HN_EF_MAX));
  MANN_REQUIRE(policy == 0 || policy == 1, ERR_ARGUMENT, "Unknown neighbor_policy");
  MANN_REQUIRE(iv.n < (int64_t)1 << 31, ERR_ARGUMENT, "at most 2^31 - 1 points");
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int64_t n = iv.n;
  // sample_layer (planner.jl:19-30) with lambda = 1 / log(max(M, 2))
  // (index.jl:35), one splitmix64 stream over the list-contiguous rows
  std::vector<int> lev((size_t)std::max<int64_t>(n, 1)), up((size_t)std::max<int64_t>(n, 1), -1);
  const double lam = 1.0 / std::log((double)std::max(M, 2));
  uint64_t s = seed;
  int top = 0;
  int64_t n1 = 0;
  for (int64_t i = 0; i < n; ++i) {
    double u = (double)(splitmix_next(s) >> 11) * (1.0 / 9007199254740992.0);
    if (u < DBL_EPSILON) u = DBL_EPSILON;
    lev[i] = (int)std::floor(-std::log(u) * lam);
    top = std::max(top, lev[i]);
    if (lev[i] >= 1) up[i] = (int)n1++;
  }
  iv.free_hnsw();
  const int cap = M + 1, layers = top + 1;
  iv.hn_M = M;
  iv.hn_layers = layers;
  iv.hn_policy = policy;
  iv.hn_ef_c = ef_c;
  iv.hn_n1 = n1;
  iv.hn_nbr0 = dev_alloc_zero<int>((size_t)n * cap, st);
  iv.hn_deg0 = dev_alloc_zero<int>((size_t)n, st);
  iv.hn_nbru = dev_alloc_zero<int>((size_t)(layers - 1) * n1 * cap, st);
  iv.hn_degu = dev_alloc_zero<int>((size_t)(layers - 1) * n1, st);
  iv.hn_entry = dev_alloc_zero<int>((size_t)iv.nb, st);
  iv.hn_maxl = dev_alloc_zero<int>((size_t)iv.nb, st);
  MANN_HIP(hipMalloc(&iv.hn_level, sizeof(int) * (size_t)std::max<int64_t>(n, 1)));
  MANN_HIP(hipMalloc(&iv.hn_upslot, sizeof(int) * (size_t)std::max<int64_t>(n, 1)));
  MANN_HIP(hipMemcpyAsync(iv.hn_level, lev.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  MANN_HIP(hipMemcpyAsync(iv.hn_upslot, up.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
  // per list: |list| + 1 pending entries, ceil(|list| / 32) visited words
  HnCand* pend = c.buf("hn_bpend").as<HnCand>((size_t)(n + iv.nb));
  unsigned* vis = c.buf("hn_bvis").as<unsigned>((size_t)((n >> 5) + iv.nb + 1));
  HnswBuildArgs a{};
  a.x32 = iv.x32;
  a.d16 = (int)iv.d16;
  a.list_off = iv.list_off;
  a.lorder = iv.lorder;
  a.nb = iv.nb;
  a.level = iv.hn_level;
  a.g = graph_of(iv);
  a.M = M;
  a.ef = ef_c;
  a.policy = policy;
  a.pend_g = pend;
  a.vis_g = vis;
  a.entry = iv.hn_entry;
  a.maxl = iv.hn_maxl;
  {
    ScopedTime tm("hnsw_build", st);
    launch_hnsw_build(a, iv.nb, st);
  }
  MANN_HIP(hipGetLastError());
  MANN_HIP(hipStreamSynchronize(st));
  timer_flush();
}

void ivf_hnsw_set_graph(mann_ivf& iv, int layers, int M, const int* nbr, const int* deg,
                        const int* entry, const int* maxl, const int* level_in) {
  MANN_REQUIRE(!iv.flat, ERR_ARGUMENT,
               "HNSW terminals belong to a MultiLevelIndex, not an IVFFlatIndex handle");
  MANN_REQUIRE(!iv.normalized, ERR_ARGUMENT,
               "HNSW terminals search with the Euclidean distance (the HNSWIndex default)");
  MANN_REQUIRE(M > 0 && M <= HN_M_MAX, ERR_ARGUMENT,
               "M must lie in [1, " + std::to_string(HN_M_MAX) + "]");
  MANN_REQUIRE(layers >= 1 && layers <= 64, ERR_ARGUMENT, "layers must lie in [1, 64]");
  const int64_t n = iv.n, cap = M + 1;
  // levels: given, or the highest layer where a row has neighbours (the entry
  // point: its bucket's top layer)
  std::vector<int> lev((size_t)std::max<int64_t>(n, 1), 0);
  for (int64_t r = 0; r < n; ++r) {
    if (level_in) {
      lev[r] = level_in[r];
      MANN_REQUIRE(lev[r] >= 0 && lev[r] < layers, ERR_ARGUMENT, "level out of range");
    } else {
      for (int l = layers - 1; l >= 1; --l)
        if (deg[(int64_t)l * n + r] > 0) {
          lev[r] = l;
          break;
        }
    }
  }
  for (int b = 0; b < iv.nb; ++b) {
    const int64_t lo = iv.h_off[b], sz = iv.h_off[b + 1] - lo;
    if (sz == 0) {
      MANN_REQUIRE(entry[b] == -1 || entry[b] == 0, ERR_ARGUMENT, "entry point of an empty list");
      continue;
    }
    MANN_REQUIRE(entry[b] >= 0 && entry[b] < sz, ERR_ARGUMENT, "entry point out of range");
    MANN_REQUIRE(maxl[b] >= 0 && maxl[b] < layers, ERR_ARGUMENT, "top layer out of range");
    int& le = lev[lo + entry[b]];
    if (!level_in) le = std::max(le, maxl[b]);
    MANN_REQUIRE(le >= maxl[b], ERR_ARGUMENT, "the entry point is not on its top layer");
    // every list within its bucket, its degree at most M, every listed node on
    // the layer (the kernels index the upper layers by that)
    for (int64_t i = 0; i < sz; ++i)
      for (int l = 0; l < layers; ++l) {
        const int dg = deg[(int64_t)l * n + lo + i];
        MANN_REQUIRE(dg >= 0 && dg <= M, ERR_ARGUMENT, "degree out of range");
        MANN_REQUIRE(dg == 0 || lev[lo + i] >= l, ERR_ARGUMENT,
                     "a row has neighbours above its level");
        const int* lst = nbr + ((int64_t)l * n + lo + i) * cap;
        for (int j = 0; j < dg; ++j) {
          MANN_REQUIRE(lst[j] >= 0 && lst[j] < sz, ERR_ARGUMENT, "neighbour id out of range");
          MANN_REQUIRE(lev[lo + lst[j]] >= l, ERR_ARGUMENT, "a neighbour is above its level");
        }
      }
  }
  std::vector<int> up((size_t)std::max<int64_t>(n, 1), -1);
  int64_t n1 = 0;
  for (int64_t r = 0; r < n; ++r)
    if (lev[r] >= 1) up[r] = (int)n1++;
  std::vector<int> nbu((size_t)std::max<int64_t>((layers - 1) * n1 * cap, 1), 0),
      dgu((size_t)std::max<int64_t>((layers - 1) * n1, 1), 0);
  for (int l = 1; l < layers; ++l)
    for (int64_t r = 0; r < n; ++r) {
      if (up[r] < 0) continue;
      const int64_t s = (int64_t)(l - 1) * n1 + up[r];
      dgu[s] = deg[(int64_t)l * n + r];
      for (int64_t j = 0; j < cap; ++j) nbu[s * cap + j] = nbr[((int64_t)l * n + r) * cap + j];
    }
  std::vector<int> ent(iv.nb), mx(iv.nb);
  for (int b = 0; b < iv.nb; ++b) {
    const bool empty = iv.h_off[b + 1] == iv.h_off[b];
    ent[b] = empty ? -1 : entry[b];
    mx[b] = empty ? -1 : maxl[b];
  }
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  iv.free_hnsw();
  iv.hn_M = M;
  iv.hn_layers = layers;
  iv.hn_n1 = n1;
  auto upload = [&](int*& dst, const int* src, int64_t count) {
    MANN_HIP(hipMalloc(&dst, sizeof(int) * (size_t)std::max<int64_t>(count, 1)));
    if (count > 0)
      MANN_HIP(hipMemcpyAsync(dst, src, sizeof(int) * count, hipMemcpyHostToDevice, st));
  };
  upload(iv.hn_nbr0, nbr, n * cap);
  upload(iv.hn_deg0, deg, n);
  upload(iv.hn_nbru, nbu.data(), (layers - 1) * n1 * cap);
  upload(iv.hn_degu, dgu.data(), (layers - 1) * n1);
  upload(iv.hn_upslot, up.data(), n);
  upload(iv.hn_level, lev.data(), n);
  upload(iv.hn_entry, ent.data(), iv.nb);
  upload(iv.hn_maxl, mx.data(), iv.nb);
  MANN_HIP(hipStreamSynchronize(st));
}

void ivf_hnsw_search(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int ef_search,
                     const float* route_in, int merge, int probe_order, int64_t* out_ids,
                     float* out_dist) {
  MANN_REQUIRE(iv.hn_nbr0 != nullptr, ERR_ARGUMENT,
               "the index has no HNSW terminals (build them with mann_ivf_hnsw_build)");
  MANN_REQUIRE(ef_search >= 1, ERR_ARGUMENT, "ef_search must be at least 1");
  const int ef = std::max(ef_search, k);
  MANN_REQUIRE(ef <= HN_EF_MAX, ERR_ARGUMENT,
               "max(ef_search, k) = " + std::to_string(ef) +
                   " exceeds the GPU search's maximum of " + std::to_string(HN_EF_MAX));
  const int nq = (int)Q.nq;
  if (nq == 0 || k <= 0) return;
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int avail = std::min(iv.nb, iv.nlist);
  nprobe = std::min(nprobe, avail);
  const float* R = route_in;
  if (!R) {
    float* rb = c.buf("hn_R").as<float>((size_t)nq * iv.nlist);
    ivf_route(iv, Q, rb);
    R = rb;
  }
  int* probes = c.buf("hn_probes").as<int>((size_t)nq * nprobe);
  int* pcount = c.buf("hn_pcount").as<int>((size_t)nq);
  int* prows = c.buf("hn_prows").as<int>((size_t)nq);
  ivf_probe(iv, R, nq, nprobe, probes, pcount, prows);
  const int64_t np = (int64_t)nq * nprobe;
  int* orow = c.buf("hn_orow").as<int>((size_t)np * k);
  float* od = c.buf("hn_od").as<float>((size_t)np * k);
  int* ocnt = c.buf("hn_ocnt").as<int>((size_t)np);
  // pending heap in LDS up to 4 ef entries (at least 512), the visited bitmap
  // in LDS when the largest list fits 1,024 words; the global spill regions
  // hold |largest list| + 1 entries and its words per workgroup.  The searches
  // wait on memory latency: as many one-wavefront workgroups as the LDS of
  // 256 CUs holds, within 1 GiB of spill regions
  const int64_t max_list = std::max<int64_t>(iv.max_list, 1);
  const int pend_lds = (int)std::min<int64_t>(std::max(512, 4 * ef), max_list + 1);
  const int64_t words = (max_list + 31) >> 5;
  const int vis_words = words <= 1024 ? (int)words : 0;
  const int64_t pstride = max_list + 1 > pend_lds ? max_list + 1 : 0;
  const int64_t vstride = vis_words ? 0 : words;
  const size_t lds = hnsw_search_lds(ef, pend_lds, vis_words);
  const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(32, (160 * 1024) / (int64_t)lds));
  const int64_t spill = pstride * (int64_t)sizeof(HnCand) + vstride * 4;
  int64_t g = std::min<int64_t>(np, 256 * per_cu);
  if (spill > 0) g = std::min<int64_t>(g, std::max<int64_t>(256, ((int64_t)1 << 30) / spill));
  const int grid = (int)std::max<int64_t>(g, 1);
  HnCand* pg = c.buf("hn_spill").as<HnCand>((size_t)std::max<int64_t>(pstride * grid, 1));
  unsigned* vg = c.buf("hn_svis").as<unsigned>((size_t)std::max<int64_t>(vstride * grid, 1));
  HnswSearchArgs a{};
  a.x32 = iv.x32;
  a.q32 = Q.q32;
  a.d16 = (int)iv.d16;
  a.list_off = iv.list_off;
  a.probes = probes;
  a.pcount = pcount;
  a.nq = nq;
  a.nprobe = nprobe;
  a.entry = iv.hn_entry;
  a.maxl = iv.hn_maxl;
  a.g = graph_of(iv);
  a.k = k;
  a.ef = ef;
  a.pend_lds = pend_lds;
  a.vis_words = vis_words;
  a.pend_g = pg;
  a.pend_stride = pstride;
  a.vis_g = vg;
  a.vis_stride = vstride;
  a.out_row = orow;
  a.out_d = od;
  a.out_cnt = ocnt;
  // the walks' distance count (two ints cleared by a fill kernel)
  int* nd = c.buf("hn_ndist").as<int>(2);
  fill_i32(nd, 2, 0, st);
  a.ndist = (unsigned long long*)nd;
  {
    ScopedTime tm("hnsw_search", st);
    launch_hnsw_search(a, grid, st);
  }
  HnswMergeArgs m{};
  m.out_row = orow;
  m.out_d = od;
  m.out_cnt = ocnt;
  m.probes = probes;
  m.pcount = pcount;
  m.nq = nq;
  m.nprobe = nprobe;
  m.k = k;
  m.merge = merge;
  m.probe_order = probe_order;
  m.perm = iv.perm;
  m.heap = c.buf("hn_heap").as<HnCand>((size_t)nq * k);
  m.scratch = c.buf("hn_mscr").as<int>((size_t)nq * 2 * nprobe);
  m.ids = out_ids;
  m.dist = out_dist;
  {
    ScopedTime tm("hnsw_merge", st);
    launch_hnsw_merge(m, st);
  }
  MANN_HIP(hipGetLastError());
  unsigned long long ndh = 0;
  MANN_HIP(hipMemcpyAsync(&ndh, nd, sizeof(ndh), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  SearchStats s{};
  s.probes = np;
  s.hnsw_distances = (long long)ndh;
  last_stats() = s;
}

}  // namespace mann
