// MultiLevelIndex{KMeansTransform, TopKRouting, BruteForceIndex} on the GPU:
// build (partition_by_transform + list-contiguous reordered base) and the
// batched query (route -> invert -> one persistent scan over all (list, query
// tile, row chunk) items -> certified exact re-rank in (dist, bucket, id)
// order, which is the reference's batched SimpleMerge order).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "core.h"
#include "ivf.h"
#include "kmeans.h"

namespace mann {

static constexpr int ROWS_ITEM = 2048;

mann_ivf* ivf_build(const mann_dataset& ds, const float* C, int nlist, const int64_t* assign_host) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int64_t n = ds.n;
  int nb = 0;
  for (int64_t j = 0; j < n; ++j) {
    MANN_REQUIRE(assign_host[j] >= 1, ERR_ARGUMENT,
                 "Assignment out of bounds: " + std::to_string(assign_host[j]));
    nb = std::max<int>(nb, (int)assign_host[j]);
  }
  MANN_REQUIRE(nb > 0, ERR_ARGUMENT, "Transform produced invalid bucket assignments");
  auto* iv = new mann_ivf();
  try {
    iv->device = c.device;
    iv->d = ds.d;
    iv->n = n;
    iv->d16 = ds.d16;
    iv->kp = ds.kp;
    iv->n_pad = ds.n_pad;
    iv->nlist = nlist;
    iv->nb = nb;
    iv->normalized = ds.normalized;
    iv->ex = ds.ex;
    iv->rx = ds.rx;
    // stable counting sort of the assignment (transforms/utils.jl:142-174)
    std::vector<int> a32(n);
    for (int64_t j = 0; j < n; ++j) a32[j] = (int)assign_host[j] - 1;
    int* da = c.buf("ivb_assign").as<int>(n);
    MANN_HIP(hipMemcpyAsync(da, a32.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    const int nblk = (int)ceil_div(n, sort_block());
    int* table = c.buf("ivb_table").as<int>((size_t)nb * nblk);
    MANN_HIP(hipMalloc(&iv->list_off, sizeof(int) * (nb + 1)));
    MANN_HIP(hipMalloc(&iv->perm, sizeof(int) * n));
    launch_counting_sort(da, n, nb, table, iv->list_off, iv->perm, st);
    iv->h_off.resize(nb + 1);
    MANN_HIP(hipMemcpyAsync(iv->h_off.data(), iv->list_off, sizeof(int) * (nb + 1),
                            hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    iv->h_off[nb] = (int)n;
    MANN_HIP(hipMemcpyAsync(iv->list_off + nb, &iv->h_off[nb], sizeof(int),
                            hipMemcpyHostToDevice, st));
    std::vector<int> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return iv->h_off[x + 1] - iv->h_off[x] > iv->h_off[y + 1] - iv->h_off[y];
    });
    MANN_HIP(hipMalloc(&iv->lorder, sizeof(int) * nb));
    MANN_HIP(hipMemcpyAsync(iv->lorder, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
    // list-contiguous copy of the base
    MANN_HIP(hipMalloc(&iv->x32, sizeof(float) * iv->n_pad * iv->d16));
    MANN_HIP(hipMalloc(&iv->xb, sizeof(bf16_t) * iv->n_pad * iv->kp));
    MANN_HIP(hipMalloc(&iv->xn, sizeof(float) * iv->n_pad));
    MANN_HIP(hipMalloc(&iv->xn2, sizeof(float) * iv->n_pad));
    MANN_HIP(hipMemsetAsync(iv->x32, 0, sizeof(float) * iv->n_pad * iv->d16, st));
    gather_rows_f32(ds.x32, iv->perm, (int)n, (int)iv->d16, iv->x32, st);
    gather_rows_bf16(ds.xb, ds.xn, iv->perm, (int)n, (int)iv->n_pad, (int)iv->kp, iv->xb, iv->xn,
                     st);
    gather_f32(ds.xn2, iv->perm, (int)n, iv->xn2, st);
    // routing operands: centroids and their A16 norms (KMeansTransform.jl:157-158)
    MANN_HIP(hipMalloc(&iv->c32, sizeof(float) * (size_t)nlist * iv->d16));
    MANN_HIP(hipMalloc(&iv->cn2, sizeof(float) * (size_t)nlist));
    MANN_HIP(hipMemcpyAsync(iv->c32, C, sizeof(float) * (size_t)nlist * iv->d16,
                            hipMemcpyDeviceToDevice, st));
    launch_a16_norms(iv->c32, nlist, (int)iv->d16, iv->cn2, st);
    MANN_HIP(hipStreamSynchronize(st));
    MANN_HIP(hipGetLastError());
  } catch (...) {
    delete iv;
    throw;
  }
  return iv;
}

// Routing distances nq x nlist (FP32 norms form).
void ivf_route(const mann_ivf& iv, const QueryOps& Q, float* R) {
  hipStream_t st = ctx().stream;
  ScopedTime tm("ivf_route", st);
  launch_route_dist(Q.q32, Q.qn2, (int)Q.nq, iv.c32, iv.cn2, iv.nlist, (int)iv.d16, R, st);
}

// Probe lists of each query: probes [nq][nprobe] (buckets, 0-based, -1 pad).
void ivf_probe(const mann_ivf& iv, const float* R, int nq, int nprobe, int* probes, int* pcount,
               int* prows) {
  hipStream_t st = ctx().stream;
  const int avail = std::min(iv.nb, iv.nlist);
  MANN_REQUIRE(avail > 0, ERR_ARGUMENT, "Transform produced no routing distances for TopKRouting");
  MANN_REQUIRE(avail <= route_max(), ERR_UNSUPPORTED, "more than 4096 routing buckets");
  ScopedTime tm("ivf_probe", st);
  launch_probe_select(R, nq, iv.nlist, avail, nprobe, iv.list_off, probes, pcount, prows, st);
}

void ivf_search(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int metric,
                const float* route_in, int64_t* out_ids, float* out_dist) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  SearchStats& S = last_stats();
  S = SearchStats();
  const int nq = (int)Q.nq;
  if (nq == 0 || k <= 0) return;
  const int cap = select_cap_max();
  MANN_REQUIRE(k <= select_exmax() / 2, ERR_UNSUPPORTED,
               "k too large for the GPU selection buffer (max 256)");
  const int BN = scan_tile_queries();
  const float* R = route_in;
  if (!R) {
    float* Rb = c.buf("iv_R").as<float>((size_t)nq * iv.nlist);
    ivf_route(iv, Q, Rb);
    R = Rb;
  }
  int* probes = c.buf("iv_probes").as<int>((size_t)nq * nprobe);
  int* pcount = c.buf("iv_pcount").as<int>(nq);
  int* prows = c.buf("iv_prows").as<int>(nq);
  ivf_probe(iv, R, nq, nprobe, probes, pcount, prows);

  // emission thresholds (certified by an exact row sample)
  float* tau = c.buf("iv_tau").as<float>(nq);
  int* keff = c.buf("iv_keff").as<int>(nq);
  IvfTauArgs ta{};
  ta.nq = nq;
  ta.nprobe = nprobe;
  ta.k = k;
  ta.d = (int)iv.d;
  ta.d16 = (int)iv.d16;
  ta.metric = metric;
  ta.emit_all_below = cap;
  ta.probes = probes;
  ta.pcount = pcount;
  ta.prows = prows;
  ta.list_off = iv.list_off;
  ta.x32 = iv.x32;
  ta.xn2 = iv.xn2;
  ta.q32 = Q.q32;
  ta.qn2 = Q.qn2;
  ta.qn = Q.qn;
  ta.qe = Q.qe;
  ta.qr = Q.qr;
  ta.ex = iv.ex;
  ta.rx = iv.rx;
  ta.tau = tau;
  ta.k_eff_q = keff;
  {
    ScopedTime tm("ivf_tau", st);
    launch_ivf_tau(ta, st);
  }

  // invert probe lists into work items
  const int nb = iv.nb;
  int* qcount = c.buf("iv_qcount").as<int>(nb);
  int* qfill = c.buf("iv_qfill").as<int>(nb);
  int* qoff = c.buf("iv_qoff").as<int>(nb);
  int* ioff = c.buf("iv_ioff").as<int>(nb);
  int* qlist = c.buf("iv_qlist").as<int>((size_t)nq * nprobe);
  int64_t max_items = 0;
  for (int b = 0; b < nb; ++b) {
    int64_t sz = iv.h_off[b + 1] - iv.h_off[b];
    max_items += ceil_div(sz, ROWS_ITEM) * (int64_t)ceil_div(nq, BN);
  }
  max_items = std::min<int64_t>(max_items, (int64_t)nq * nprobe * ceil_div(iv.n, ROWS_ITEM));
  ScanItem* items = c.buf("iv_items").as<ScanItem>((size_t)std::max<int64_t>(max_items, 1));
  int* ctr = c.buf("iv_ctr").as<int>(4);  // [0] n_items, [1] item counter, [2] flags, [3] stats
  int* stats = c.buf("iv_stats").as<int>(4);
  fill_i32(qcount, nb, 0, st);
  fill_i32(qfill, nb, 0, st);
  fill_i32(ctr, 4, 0, st);
  fill_i32(stats, 4, 0, st);
  {
    ScopedTime tm("ivf_plan", st);
    launch_ivf_plan(probes, pcount, nq, nprobe, nb, iv.list_off, iv.lorder, BN, ROWS_ITEM, qcount,
                    qfill, qoff, ioff, qlist, items, ctr, st);
  }

  unsigned* cand = c.buf("iv_cand").as<unsigned>((size_t)nq * cap * 2);
  int* cnt = c.buf("iv_cnt").as<int>(nq);
  fill_i32(cnt, nq, 0, st);
  ScanArgs a{};
  a.xb = iv.xb;
  a.xn = iv.xn;
  a.qb = Q.qb;
  a.kp = (int)iv.kp;
  a.d = (int)iv.d;
  a.qidx = qlist;
  a.items = items;
  a.n_items_dev = ctr;
  a.item_counter = ctr + 1;
  a.tau = tau;
  a.cand = cand;
  a.cnt = cnt;
  a.cap = cap;
  a.parts = 1;
  int dev_cus = 256;
  (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, c.device);
  {
    ScopedTime tm("ivf_scan", st);
    launch_scan(a, SCAN_GATHER, 2 * dev_cus, SCAN_KIND_ITEMS, st);
  }

  int* flist = c.buf("iv_flist").as<int>(nq);
  float* tau2 = c.buf("iv_tau2").as<float>(nq);
  SelectArgs s{};
  s.nq = nq;
  s.cand = cand;
  s.cnt = cnt;
  s.cap = cap;
  s.parts = 1;
  s.tau = tau;
  s.x32 = iv.x32;
  s.q32 = Q.q32;
  s.xn2 = iv.xn2;
  s.qn2 = Q.qn2;
  s.d = (int)iv.d;
  s.d16 = (int)iv.d16;
  s.metric = metric;
  s.perm = iv.perm;
  s.qn = Q.qn;
  s.qe = Q.qe;
  s.qr = Q.qr;
  s.ex = iv.ex;
  s.rx = iv.rx;
  s.k = k;
  s.k_req = k;
  s.k_eff = k;
  s.k_eff_q = keff;
  s.shortlist = std::min(select_exmax() / 2, std::max(tuning().shortlist_mult * k, 32));
  s.out_ids = out_ids;
  s.out_dist = out_dist;
  s.flag_count = ctr + 2;
  s.flag_list = flist;
  s.tau_new = tau2;
  s.stats = stats;
  {
    ScopedTime tm("ivf_select", st);
    launch_select_rerank(s, nq, st);
  }
  // overflowed queries: exact scan of their probed lists
  int2* ranges = c.buf("iv_ranges").as<int2>((size_t)nq * nprobe);
  launch_probe_ranges(probes, pcount, nq, nprobe, iv.list_off, ranges, st);
  ExactArgs e{};
  e.qlist = flist;
  e.nq_dev = ctr + 2;
  e.x32 = iv.x32;
  e.q32 = Q.q32;
  e.xn2 = iv.xn2;
  e.qn2 = Q.qn2;
  e.d = (int)iv.d;
  e.d16 = (int)iv.d16;
  e.metric = metric;
  e.perm = iv.perm;
  e.range_cnt = pcount;
  e.ranges = ranges;
  e.range_stride = nprobe;
  e.k = k;
  e.k_req = k;
  e.out_ids = out_ids;
  e.out_dist = out_dist;
  {
    ScopedTime tm("ivf_exact_fallback", st);
    launch_exact_topk(e, nq, st);
  }
  int h[4], hs[4];
  MANN_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  MANN_HIP(hipGetLastError());
  S.exact = h[2];
  S.reranked = hs[0];
  S.overflow = hs[1];
}

}  // namespace mann
