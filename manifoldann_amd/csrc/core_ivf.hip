// MultiLevelIndex{KMeansTransform, TopKRouting, BruteForceIndex} on the GPU:
// build (partition_by_transform + list-contiguous reordered base) and the
// batched query (route -> invert -> one persistent scan over all (list, query
// tile, row chunk) items -> certified exact re-rank in (dist, bucket, id)
// order, which is the reference's batched SimpleMerge order).
//
// Residual encoding: the bf16 operands of list l hold x - c_l and q - c_l
// (c_l its centroid).  Distances are translation invariant, so the first pass
// approximates the same |q - x|; the residuals are far shorter than the
// vectors, which narrows the certification band of every bound by the same
// factor.  The exact FP32 re-rank always reads the original rows.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "core.h"
#include "ivf.h"
#include "kmeans.h"

namespace mann {


// The probe lists a search used (for the merges that replay them).
struct ProbeState {
  const int* probes = nullptr;  // [nq][nprobe]
  const int* pcount = nullptr;
  const int* prows = nullptr;
  int nprobe = 0;
};

static void ivf_search_sel(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int metric,
                           const float* route_in, int64_t* out_ids, float* out_dist, float mass,
                           ProbeState* ps = nullptr);

ShapeKey ivf_shape_key(const mann_ivf& iv, int64_t nq, int k, int nprobe, int metric) {
  ShapeKey key;
  key.path = 1;
  key.device = iv.device;
  key.metric = metric;
  key.k = k;
  key.n = iv.n;
  key.d = iv.d;
  key.nprobe = nprobe;
  key.lists = iv.nlist;
  int64_t b = 1;
  while (b < nq) b <<= 1;
  key.nq_bucket = b;
  return key;
}

void ivf_search(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int metric,
                const float* route_in, int64_t* out_ids, float* out_dist, float mass, int merge,
                int probe_order) {
  // TopKRouting takes min(nprobe, available buckets) lists (routing.jl:108-112):
  // a larger nprobe changes nothing but the buffers sized by it
  if (mass <= 0.0f) nprobe = std::max(1, std::min(nprobe, std::min(iv.nb, iv.nlist)));
  // knobs the autotuner chose for this shape, for this call only
  TuningScope tscope;
  if (!shape_tuning_disabled() && mass <= 0.0f) {
    Tuning t = tuning();
    if (shape_tuning_apply(ivf_shape_key(iv, Q.nq, k, nprobe, metric), t)) tscope.set(t);
  }
  const int kind = iv.flat ? MERGE_HEAP_ROWS
                           : (merge == 1 ? MERGE_HEAP_LISTS
                                         : (probe_order ? MERGE_PROBE : MERGE_BATCHED));
  if (kind == MERGE_BATCHED || Q.nq == 0 || k <= 0) {
    ivf_search_sel(iv, Q, k, nprobe, metric, route_in, out_ids, out_dist, mass);
    return;
  }
  MANN_REQUIRE(!ctx().capture, ERR_ARGUMENT,
               "a captured IVF search merges in the batched SimpleMerge order only");
  // the (dist, bucket, id) top-(k+1), then the kind's merge: untied queries
  // re-sorted, queries tied at the k-th place replayed (kernels_merge.hip)
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  MANN_REQUIRE(k + 1 <= select_exmax() / 2, ERR_ARGUMENT,
               "k = " + std::to_string(k) + " exceeds this merge's maximum of " +
                   std::to_string(select_exmax() / 2 - 1) + " neighbours per query");
  const int nq = (int)Q.nq;
  int64_t* ti = c.buf("ivm_ti").as<int64_t>((size_t)nq * (k + 1));
  float* td = c.buf("ivm_td").as<float>((size_t)nq * (k + 1));
  ProbeState P;
  ivf_search_sel(iv, Q, k + 1, nprobe, metric, route_in, ti, td, mass, &P);
  SearchStats keep = last_stats();
  int* tied = c.buf("ivm_tied").as<int>(nq);
  int* ntied = c.buf("ivm_ntied").as<int>(1);
  fill_i32(ntied, 1, 0, st);
  MergeArgs m{};
  m.ti = ti;
  m.td = td;
  m.nq = nq;
  m.k = k;
  m.kind = kind;
  m.probes = P.probes;
  m.pcount = P.pcount;
  m.nprobe = P.nprobe;
  m.bucket_of = iv.bucket_of;
  m.perm = iv.perm;
  m.list_off = iv.list_off;
  m.out_ids = out_ids;
  m.out_dist = out_dist;
  m.tied_list = tied;
  m.tied_count = ntied;
  {
    ScopedTime tm("ivf_merge", st);
    launch_merge_fix(m, st);
  }
  int nt = 0;
  MANN_HIP(hipMemcpyAsync(&nt, ntied, sizeof(int), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  if (nt > 0) {
    // replay the tied queries in batches whose probed rows fit the distance
    // buffer (each tied query's rows are computed once, in replay order)
    std::vector<int> ht(nt), hrows(nq);
    MANN_HIP(hipMemcpyAsync(ht.data(), tied, sizeof(int) * nt, hipMemcpyDeviceToHost, st));
    MANN_HIP(hipMemcpyAsync(hrows.data(), P.prows, sizeof(int) * nq, hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    const int64_t budget = (int64_t)1 << 24;  // distances per batch (64 MiB)
    int* seq = c.buf("ivm_seq").as<int>((size_t)nt * P.nprobe);
    int* seqoff = c.buf("ivm_seqoff").as<int>((size_t)nt * (P.nprobe + 1));
    int64_t* dbase = c.buf("ivm_dbase").as<int64_t>(nt);
    std::vector<int64_t> hb(nt);
    ReplayArgs r{};
    r.tied = tied;
    r.nprobe = P.nprobe;
    r.bucket_order = kind == MERGE_HEAP_LISTS && !probe_order;
    r.probes = P.probes;
    r.pcount = P.pcount;
    r.list_off = iv.list_off;
    r.x32 = iv.x32;
    r.q32 = Q.q32;
    r.xn2 = iv.xn2;
    r.qn2 = Q.qn2;
    r.d16 = (int)iv.d16;
    r.metric = metric;
    ScopedTime tm("ivf_merge_replay", st);
    for (int s0 = 0; s0 < nt;) {
      // slots [s0, s1): at least one, more while their rows fit the budget
      int64_t tot = 0;
      int s1 = s0, maxr = 0;
      while (s1 < nt && (s1 == s0 || tot + hrows[ht[s1]] <= budget)) {
        hb[s1 - s0] = tot;
        tot += hrows[ht[s1]];
        maxr = std::max(maxr, hrows[ht[s1]]);
        ++s1;
      }
      float* dist = c.buf("ivm_dist").as<float>((size_t)std::max<int64_t>(tot, 1));
      MANN_HIP(hipMemcpyAsync(dbase, hb.data(), sizeof(int64_t) * (s1 - s0),
                              hipMemcpyHostToDevice, st));
      r.slot0 = s0;
      r.seq = seq;
      r.seqoff = seqoff;
      r.dist = dist;
      r.dbase = dbase;
      launch_replay(r, m, s1 - s0, maxr, st);
      // dbase (a host copy source) and dist are reused by the next batch
      MANN_HIP(hipStreamSynchronize(st));
      s0 = s1;
    }
  }
  MANN_HIP(hipGetLastError());
  keep.heap_replays = nt;
  last_stats() = keep;
}

mann_ivf* ivf_build(const mann_dataset& ds, const float* C, int nlist, const int64_t* assign_host,
                    bool all_lists, int route_metric) {
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
  // IVF-Flat keeps one list per centroid, empty trailing ones included, so that
  // they take probe slots as in ivf_flat.jl:68-77,150
  if (all_lists) nb = std::max(nb, nlist);
  const int flat = all_lists ? 1 : 0;
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
    iv->flat = flat;
    iv->normalized = ds.normalized;
    iv->route_metric = route_metric;
    // stable counting sort of the assignment (transforms/utils.jl:142-174)
    std::vector<int> a32(n);
    for (int64_t j = 0; j < n; ++j) a32[j] = (int)assign_host[j] - 1;
    int* da = c.buf("ivb_assign").as<int>(n);
    MANN_HIP(hipMemcpyAsync(da, a32.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    const int nblk = (int)ceil_div(n, sort_block());
    int* table = c.buf("ivb_table").as<int>((size_t)counting_sort_table_ints(n, nb));
    MANN_HIP(hipMalloc(&iv->list_off, sizeof(int) * (nb + 1)));
    MANN_HIP(hipMalloc(&iv->perm, sizeof(int) * n));
    launch_counting_sort(da, n, nb, table, iv->list_off, iv->perm, st);
    MANN_HIP(hipMalloc(&iv->bucket_of, sizeof(int) * n));
    MANN_HIP(hipMemcpyAsync(iv->bucket_of, da, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    iv->h_off.resize(nb + 1);
    MANN_HIP(hipMemcpyAsync(iv->h_off.data(), iv->list_off, sizeof(int) * (nb + 1),
                            hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    iv->h_off[nb] = (int)n;
    for (int b = 0; b < nb; ++b)
      iv->max_list = std::max<int64_t>(iv->max_list, iv->h_off[b + 1] - iv->h_off[b]);
    MANN_HIP(hipMemcpyAsync(iv->list_off + nb, &iv->h_off[nb], sizeof(int),
                            hipMemcpyHostToDevice, st));
    std::vector<int> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return iv->h_off[x + 1] - iv->h_off[x] > iv->h_off[y + 1] - iv->h_off[y];
    });
    MANN_HIP(hipMalloc(&iv->lorder, sizeof(int) * nb));
    MANN_HIP(hipMemcpyAsync(iv->lorder, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
    // routing operands: centroids and their A16 norms (KMeansTransform.jl:157-158).
    // Buckets beyond the centroids (an external assignment) get a zero centre.
    const int nc = std::max(nb, nlist);
    MANN_HIP(hipMalloc(&iv->c32, sizeof(float) * (size_t)nc * iv->d16));
    MANN_HIP(hipMalloc(&iv->cn2, sizeof(float) * (size_t)nlist));
    MANN_HIP(hipMemsetAsync(iv->c32, 0, sizeof(float) * (size_t)nc * iv->d16, st));
    MANN_HIP(hipMemcpyAsync(iv->c32, C, sizeof(float) * (size_t)nlist * iv->d16,
                            hipMemcpyDeviceToDevice, st));
    launch_a16_norms(iv->c32, nlist, (int)iv->d16, iv->cn2, st);
    // bf16 operands of the certified TopKRouting (norms form, rows not unit)
    const int avail = std::min(nb, nlist);
    if (!iv->normalized && (route_metric == EUCLIDEAN || route_metric == SQEUCLIDEAN) &&
        avail <= route_select_cap()) {
      iv->r_pad = round_up(avail, 256);
      MANN_HIP(hipMalloc(&iv->rb, sizeof(bf16_t) * (size_t)iv->r_pad * iv->kp));
      MANN_HIP(hipMalloc(&iv->rn, sizeof(float) * (size_t)iv->r_pad));
      float* t32 = c.buf("ivb_rt32").as<float>((size_t)iv->r_pad * iv->d16);
      unsigned* rm = c.buf("ivb_rmax").as<unsigned>(3);
      MANN_HIP(hipMemsetAsync(rm, 0, sizeof(unsigned) * 3, st));
      launch_prep_rows(iv->c32, iv->d16, avail, iv->r_pad, (int)iv->d, (int)iv->d16, (int)iv->kp,
                       0, t32, iv->rb, iv->rn, nullptr, nullptr, rm, rm + 1, st);
      launch_norm_bound_max(iv->cn2, avail, (int)iv->d, rm + 2, st);
      unsigned hr[3];
      MANN_HIP(hipMemcpyAsync(hr, rm, sizeof(hr), hipMemcpyDeviceToHost, st));
      MANN_HIP(hipStreamSynchronize(st));
      std::memcpy(&iv->rex, &hr[0], 4);
      std::memcpy(&iv->rrx, &hr[1], 4);
      std::memcpy(&iv->rra, &hr[2], 4);
    }
    // residual centres: unit rows residualise against the unit centroids
    // (any centre serves, distances being translation invariant; the nearer
    // one keeps the residuals, and so the certification band, small)
    if (iv->normalized) {
      MANN_HIP(hipMalloc(&iv->cres, sizeof(float) * (size_t)nc * iv->d16));
      unit_rows(iv->c32, iv->cres, iv->d16, (int)iv->d, nc, st);
    } else {
      iv->cres = iv->c32;
    }
    // list-contiguous copy of the base; bf16 residuals against the list centre
    MANN_HIP(hipMalloc(&iv->x32, sizeof(float) * iv->n_pad * iv->d16));
    MANN_HIP(hipMalloc(&iv->xb, sizeof(bf16_t) * iv->n_pad * iv->kp));
    MANN_HIP(hipMalloc(&iv->xn, sizeof(float) * iv->n_pad));
    MANN_HIP(hipMalloc(&iv->xn2, sizeof(float) * iv->n_pad));
    MANN_HIP(hipMemsetAsync(iv->x32, 0, sizeof(float) * iv->n_pad * iv->d16, st));
    gather_rows_f32(ds.x32, iv->perm, (int)n, (int)iv->d16, iv->x32, st);
    gather_f32(ds.xn2, iv->perm, (int)n, iv->xn2, st);
    int* rowb = c.buf("ivb_rowb").as<int>(n);
    gather_rows_i32(da, iv->perm, (int)n, 1, rowb, st);
    unsigned* mx = c.buf("ivb_max").as<unsigned>(2);
    MANN_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned) * 2, st));
    launch_prep_rows(iv->x32, iv->d16, n, iv->n_pad, (int)iv->d, (int)iv->d16, (int)iv->kp,
                     iv->normalized, iv->x32, iv->xb, iv->xn, nullptr, nullptr, mx, mx + 1, st,
                     iv->cres, rowb, iv->d16);
    unsigned hm[2];
    MANN_HIP(hipMemcpyAsync(hm, mx, sizeof(hm), hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    std::memcpy(&iv->ex, &hm[0], 4);
    std::memcpy(&iv->rx, &hm[1], 4);
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
  launch_route_dist(Q.q32, Q.qn2, (int)Q.nq, iv.c32, iv.cn2, iv.nlist, (int)iv.d16, R, st,
                    iv.route_metric);
}

// Probe lists of each query: probes [nq][nprobe] (buckets, 0-based, -1 pad).
void ivf_probe(const mann_ivf& iv, const float* R, int nq, int nprobe, int* probes, int* pcount,
               int* prows, float mass) {
  hipStream_t st = ctx().stream;
  const int avail = std::min(iv.nb, iv.nlist);
  MANN_REQUIRE(avail > 0, ERR_ARGUMENT, "Transform produced no routing distances for TopKRouting");
  MANN_REQUIRE(avail <= route_max(), ERR_UNSUPPORTED, "more than 4096 routing buckets");
  if (mass > 0.0f)
    MANN_REQUIRE(nprobe <= 64 && avail <= 1024, ERR_UNSUPPORTED,
                 "probability-mass probing supports nprobe <= 64 and at most 1024 lists");
  ScopedTime tm("ivf_probe", st);
  launch_probe_select(R, nq, iv.nlist, avail, nprobe, mass, iv.list_off, probes, pcount, prows,
                      st);
}

// TopKRouting (routing.jl:94-118) from a bf16 first pass over the centroids:
// a DENSE scan of S^ for every (query, centroid), then route_select_kernel
// re-ranks in FP32 only the centroids the norms-form bound cannot exclude.
// The probes equal ivf_route + ivf_probe's (same keys, same tie order).
static bool route_certified_applies(const mann_ivf& iv, int nprobe, float mass) {
  return iv.rb != nullptr && mass <= 0.0f && nprobe <= 64 && tuning().ivf_route_cert;
}

static void ivf_route_certified(const mann_ivf& iv, const QueryOps& Q, int nprobe, int* probes,
                                int* pcount, int* prows) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int nq = (int)Q.nq;
  const int avail = std::min(iv.nb, iv.nlist);
  const int cfg = std::min(std::max(tuning().ivf_route_cfg, 0), scan_num_cfg() - 1);
  const int BM = scan_tile_rows(cfg), BN = scan_tile_queries(cfg);
  const int64_t ldd = round_up(avail, 4);
  // query chunks of at most 2^26 values
  const int qc = (int)std::min<int64_t>(
      nq, std::max<int64_t>(BN, ((int64_t)1 << 26) / ldd / BN * BN));
  float* dense = c.buf("iv_rdense").as<float>((size_t)qc * ldd);
  for (int q0 = 0; q0 < nq; q0 += qc) {
    const int nqc = std::min(qc, nq - q0);
    ScanArgs a{};
    a.cfg = cfg;
    a.xb = iv.rb;
    a.xn = iv.rn;
    a.qb = Q.qb + (int64_t)q0 * iv.kp;
    a.qnorm = Q.qn + q0;
    a.kp = (int)iv.kp;
    a.d = (int)iv.d;
    a.nq = nqc;
    // row partitions when the chunk has too few query tiles to fill the chip
    const int qt = (int)ceil_div(nqc, BN);
    const int pd = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(avail, BM),
                                                                ceil_div(512, qt)));
    a.rows_per_part = (int)round_up(ceil_div(avail, pd), BM);
    a.parts = (int)ceil_div(avail, a.rows_per_part);
    a.qtiles = qt;
    a.n_rows = avail;
    a.dense = dense;
    a.ldd = ldd;
    a.dvec = 1;  // ldd and the partitions' first rows are multiples of 4
    {
      ScopedTime tm("ivf_route_scan", st);
      launch_scan(a, SCAN_DENSE, qt * a.parts, SCAN_KIND_GRID, st);
    }
    RouteSelArgs r{};
    r.dense = dense;
    r.ldd = ldd;
    r.avail = avail;
    r.nprobe = nprobe;
    r.q0 = q0;
    r.nq = nqc;
    r.c32 = iv.c32;
    r.cn2 = iv.cn2;
    r.q32 = Q.q32;
    r.qn2 = Q.qn2;
    r.d16 = (int)iv.d16;
    r.d = (int)iv.d;
    r.sq = iv.route_metric == SQEUCLIDEAN ? 1 : 0;
    r.qe = Q.qe;
    r.qr = Q.qr;
    r.ex = iv.rex;
    r.rx = iv.rrx;
    r.ra = iv.rra;
    r.list_off = iv.list_off;
    r.probes = probes;
    r.pcount = pcount;
    r.prows = prows;
    ScopedTime tm("ivf_route_select", st);
    launch_route_select(r, st);
  }
}

static void ivf_search_sel(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int metric,
                           const float* route_in, int64_t* out_ids, float* out_dist, float mass,
                           ProbeState* ps) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  SearchStats& S = last_stats();
  S = SearchStats();
  const int nq = (int)Q.nq;
  if (nq == 0 || k <= 0) return;
  const int cap = select_cap_max();
  MANN_REQUIRE(k <= select_exmax() / 2, ERR_ARGUMENT,
               "k = " + std::to_string(k) + " exceeds the IVF search's maximum of " +
                   std::to_string(select_exmax() / 2) + " neighbours per query");
  const int cfg = tuning().ivf_cfg;
  const int BN = scan_tile_queries(cfg);
  const int ROWS_ITEM = tuning().ivf_rows_item;  // rows per work item
  int* probes = c.buf("iv_probes").as<int>((size_t)nq * nprobe);
  int* pcount = c.buf("iv_pcount").as<int>(nq);
  int* prows = c.buf("iv_prows").as<int>(nq);
  if (!route_in && route_certified_applies(iv, nprobe, mass)) {
    ivf_route_certified(iv, Q, nprobe, probes, pcount, prows);
  } else {
    const float* R = route_in;
    if (!R) {
      float* Rb = c.buf("iv_R").as<float>((size_t)nq * iv.nlist);
      ivf_route(iv, Q, Rb);
      R = Rb;
    }
    ivf_probe(iv, R, nq, nprobe, probes, pcount, prows, mass);
  }
  if (ps) {
    ps->probes = probes;
    ps->pcount = pcount;
    ps->prows = prows;
    ps->nprobe = nprobe;
  }

  // invert probe lists into work items
  const int nb = iv.nb;
  int* qcount = c.buf("iv_qcount").as<int>(nb);
  int* qfill = c.buf("iv_qfill").as<int>(nb);
  int* qoff = c.buf("iv_qoff").as<int>(nb);
  int* ioff = c.buf("iv_ioff").as<int>(nb);
  const int64_t max_entries = (int64_t)nq * nprobe;
  int* qlist = c.buf("iv_qlist").as<int>((size_t)max_entries);
  int* qlist_list = c.buf("iv_qlist_list").as<int>((size_t)max_entries);
  int* slot_of = c.buf("iv_slot_of").as<int>((size_t)max_entries);
  int64_t max_items = 0;
  for (int b = 0; b < nb; ++b) {
    int64_t sz = iv.h_off[b + 1] - iv.h_off[b];
    max_items += ceil_div(sz, ROWS_ITEM) * (int64_t)ceil_div(nq, BN);
  }
  max_items = std::min<int64_t>(max_items, (int64_t)nq * nprobe * ceil_div(iv.n, ROWS_ITEM));
  ScanItem* items = c.buf("iv_items").as<ScanItem>((size_t)std::max<int64_t>(max_items, 1));
  // threshold from a bf16 scan of each query's first list (dense S^ values),
  // unless the dense block would be too large
  const bool dense_tau = tuning().ivf_tau_mode == 1 &&
                         (int64_t)nq * iv.max_list <= ((int64_t)1 << 29);
  int* first_cnt = dense_tau ? c.buf("iv_first_cnt").as<int>(nb) : nullptr;
  // [0] n_items, [1] item counter, [2] flagged after pass 1, [3] flagged after
  // pass 2, [4] entries, [5] threshold-scan / refinement items, [6] their item
  // counter (16 entries: a captured search's launcher reads 8 from [2])
  int* ctr = c.buf("iv_ctr").as<int>(16);
  int* stats = c.buf("iv_stats").as<int>(kStatSlots * 4);
  fill_i32(qcount, nb, 0, st);
  fill_i32(qfill, nb, 0, st);
  fill_i32(ctr, 16, 0, st);
  fill_i32(stats, kStatSlots * 4, 0, st);
  {
    ScopedTime tm("ivf_plan", st);
    launch_ivf_plan(probes, pcount, nq, nprobe, nb, iv.list_off, iv.lorder, BN, ROWS_ITEM, qcount,
                    qfill, qoff, ioff, qlist, qlist_list, slot_of, items, ctr, ctr + 4, first_cnt,
                    st);
  }
  // residual query rows per (list, query) entry and per-query bounds
  bf16_t* qres = c.buf("iv_qres").as<bf16_t>((size_t)max_entries * iv.kp);
  float* rnorm = c.buf("iv_rnorm").as<float>((size_t)max_entries);
  unsigned* qem = c.buf("iv_qem").as<unsigned>((size_t)nq * 2);
  fill_i32((int*)qem, (int64_t)nq * 2, 0, st);
  {
    ScopedTime tm("ivf_residual", st);
    launch_residual_queries(Q.q32, iv.cres, (int)iv.d, (int)iv.d16, (int)iv.kp, iv.normalized,
                            qlist, qlist_list,
                            ctr + 4, max_entries, qres, rnorm, qem, qem + nq, st);
  }
  const float* qe = (const float*)qem;
  const float* qr = (const float*)(qem + nq);

  // emission thresholds
  float* tau = c.buf("iv_tau").as<float>(nq);
  int* keff = c.buf("iv_keff").as<int>(nq);
  if (dense_tau) {
    // bf16 residual scan of each query's first list; the k-th smallest S^
    // bounds the k-th exact value from above (certified by construction)
    // a tile config of its own: a list is the first probe of few queries
    const int tcfg = tuning().ivf_tau_cfg;
    const int BNt = scan_tile_queries(tcfg);
    // each query has one first list: sum over lists of ceil(first_cnt / BNt) <=
    // nb + nq / BNt query tiles, each over at most ceil(max_list / ROWS_ITEM) chunks
    const int64_t max_fitems =
        (nb + ceil_div(nq, BNt)) * std::max<int64_t>(1, ceil_div(iv.max_list, ROWS_ITEM));
    ScanItem* fitems = c.buf("iv_fitems").as<ScanItem>((size_t)max_fitems);
    int* fioff = c.buf("iv_fioff").as<int>(nb);
    // rows of max_list values, padded to a multiple of 4 for 16-byte stores
    // (4-row minima take 8 slots per 32-row block: fewer)
    const int64_t ldd = round_up(std::max<int64_t>(iv.max_list, 8 * ceil_div(iv.max_list, 32)), 4);
    // lists of at least grp_min rows store 4-row minima: at least k of them
    // (work items hold at least min(list, rows_item) rows, so grp_min <= rows_item)
    const int grp_min = tuning().ivf_tau_grp && std::max(64, 4 * k) <= ROWS_ITEM
                            ? std::max(64, 4 * k)
                            : 0;
    float* dense = c.buf("iv_dense").as<float>((size_t)nq * ldd);
    {
      ScopedTime tm("ivf_tau_scan", st);
      launch_first_items(first_cnt, qoff, nb, iv.list_off, iv.lorder, BNt, ROWS_ITEM,
                         tuning().ivf_first_rows, fioff, fitems, ctr + 5, st);
      fill_i32(ctr + 6, 1, 0, st);
      ScanArgs a{};
      a.cfg = tcfg;
      a.xb = iv.xb;
      a.xn = iv.xn;
      a.qb = qres;
      a.col_norm = rnorm;
      a.b_by_slot = 1;
      a.kp = (int)iv.kp;
      a.d = (int)iv.d;
      a.qidx = qlist;
      a.items = fitems;
      a.n_items_dev = ctr + 5;
      a.item_counter = ctr + 6;
      a.dense = dense;
      a.ldd = ldd;
      a.dense_rel = 1;
      a.dvec = 1;  // items start at their list's first row plus multiples of 256
      a.parts = 1;
      a.dense_grp = grp_min;
      int cus = 256;
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device);
      launch_scan(a, SCAN_DENSE, 2 * cus, SCAN_KIND_ITEMS, st);
    }
    IvfDenseTauArgs td{};
    td.nq = nq;
    td.nprobe = nprobe;
    td.k = k;
    td.d = (int)iv.d;
    td.metric = metric;
    td.emit_all_below = cap;
    td.probes = probes;
    td.pcount = pcount;
    td.prows = prows;
    td.list_off = iv.list_off;
    td.dense = dense;
    td.ldd = ldd;
    td.qe = qe;
    td.qr = qr;
    td.ex = iv.ex;
    td.rx = iv.rx;
    td.tau = tau;
    td.k_eff_q = keff;
    td.max_rows = tuning().ivf_first_rows;
    td.grp_min = grp_min;
    ScopedTime tm("ivf_tau", st);
    launch_ivf_tau_dense(td, st);
  } else {
    // exact FP32 sample of the first rows of the probed lists
    IvfTauArgs ta{};
    ta.nq = nq;
    ta.nprobe = nprobe;
    ta.k = k;
    ta.d = (int)iv.d;
    ta.d16 = (int)iv.d16;
    ta.metric = metric;
    ta.emit_all_below = cap;
    ta.sample_rows = std::max(1, std::min(tuning().tau_rows, ivf_tau_rows_max()));
    ta.probes = probes;
    ta.pcount = pcount;
    ta.prows = prows;
    ta.list_off = iv.list_off;
    ta.x32 = iv.x32;
    ta.xn2 = iv.xn2;
    ta.q32 = Q.q32;
    ta.qn2 = Q.qn2;
    ta.qe = qe;
    ta.qr = qr;
    ta.ex = iv.ex;
    ta.rx = iv.rx;
    ta.tau = tau;
    ta.k_eff_q = keff;
    ta.qlist = qlist;
    ta.qlist_list = qlist_list;
    ta.n_entries = ctr + 4;
    ta.max_entries = max_entries;
    ScopedTime tm("ivf_tau", st);
    launch_ivf_tau(ta, st);
  }

  unsigned* cand = c.buf("iv_cand").as<unsigned>((size_t)nq * cap * 2);
  int* cnt = c.buf("iv_cnt").as<int>(nq);
  fill_i32(cnt, nq, 0, st);
  ScanArgs a{};
  a.cfg = cfg;
  a.xb = iv.xb;
  a.xn = iv.xn;
  a.qb = qres;
  a.col_norm = rnorm;
  a.b_by_slot = 1;
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
  s.qe = qe;
  s.qr = qr;
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
  if (c.capture) {
    // a captured search reads nothing back: the refinement pass and the
    // second selection run for every query the device flags (launch shapes
    // sized for all of them), the exact fallback for the first
    // fallback_slots; the graph's launcher runs the rest of the fallback
    const int64_t rmax = (int64_t)nq * (ceil_div(iv.n, ROWS_ITEM) + nprobe);
    ScanItem* ritems = c.buf("iv_ritems").as<ScanItem>((size_t)rmax);
    int* flist2 = c.buf("iv_flist2").as<int>(nq);
    {
      ScopedTime tm("ivf_refine_scan", st);
      fill_i32(ctr + 5, 2, 0, st);
      launch_refine_items(flist, ctr + 2, nq, nprobe, probes, pcount, slot_of, iv.list_off,
                          ROWS_ITEM, ritems, ctr + 5, st);
      reset_listed(cnt, flist, ctr + 2, nq, st);
      a.items = ritems;
      a.n_items_dev = ctr + 5;
      a.item_counter = ctr + 6;
      a.tau = tau2;
      launch_scan(a, SCAN_GATHER, 2 * dev_cus, SCAN_KIND_ITEMS, st);
    }
    s.qlist = flist;
    s.nq_dev = ctr + 2;
    s.tau = tau2;
    s.flag_count = ctr + 3;
    s.flag_list = flist2;
    s.tau_new = tau;
    {
      ScopedTime tm("ivf_select_pass2", st);
      launch_select_rerank(s, nq, st);
    }
    int2* ranges = c.buf("iv_ranges").as<int2>((size_t)nq * nprobe);
    launch_probe_ranges(probes, pcount, nq, nprobe, iv.list_off, ranges, st);
    ExactArgs e{};
    e.qlist = flist2;
    e.nq_dev = ctr + 3;
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
    CaptureInfo& ci = *c.capture;
    ci.e = run_exact_fallback(e, std::min(nq, ci.fallback_slots), iv.n, nprobe);
    ci.flags = ctr + 2;  // [0] refined, [1] left for the exact fallback
    ci.stats = stats;
    ci.filled = true;
    return;
  }
  int h[8], hs[kStatSlots * 4];
  MANN_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  S.probes = h[4];
  const int nflag = h[2];
  int nflag2 = 0;
  if (nflag > 0) {
    // refinement: the flagged queries re-scan their lists with the threshold
    // of the candidates they already hold (single-query items)
    const int64_t rmax = (int64_t)nflag * (ceil_div(iv.n, ROWS_ITEM) + nprobe);
    ScanItem* ritems = c.buf("iv_ritems").as<ScanItem>((size_t)rmax);
    int* flist2 = c.buf("iv_flist2").as<int>(nq);
    {
      ScopedTime tm("ivf_refine_scan", st);
      fill_i32(ctr + 5, 2, 0, st);
      launch_refine_items(flist, ctr + 2, nflag, nprobe, probes, pcount, slot_of, iv.list_off,
                          ROWS_ITEM, ritems, ctr + 5, st);
      reset_listed(cnt, flist, ctr + 2, nflag, st);
      a.items = ritems;
      a.n_items_dev = ctr + 5;
      a.item_counter = ctr + 6;
      a.tau = tau2;
      launch_scan(a, SCAN_GATHER, 2 * dev_cus, SCAN_KIND_ITEMS, st);
    }
    s.qlist = flist;
    s.nq_dev = ctr + 2;
    s.tau = tau2;
    s.flag_count = ctr + 3;
    s.flag_list = flist2;
    s.tau_new = tau;  // not read afterwards
    {
      ScopedTime tm("ivf_select_pass2", st);
      launch_select_rerank(s, nflag, st);
    }
    MANN_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    nflag2 = h[3];
    if (nflag2 > 0) {
      // still uncertified: exact scan of their probed lists
      int2* ranges = c.buf("iv_ranges").as<int2>((size_t)nq * nprobe);
      launch_probe_ranges(probes, pcount, nq, nprobe, iv.list_off, ranges, st);
      ExactArgs e{};
      e.qlist = flist2;
      e.nq_dev = ctr + 3;
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
      ScopedTime tm("ivf_exact_fallback", st);
      run_exact_fallback(e, nflag2, iv.n, nprobe);
    }
  }
  MANN_HIP(hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  MANN_HIP(hipGetLastError());
  S.pass2 = nflag;
  S.exact = nflag2;
  add_stats(hs, S);
}

// ---- autotuner -----------------------------------------------------------------
// Coordinate descent over the IVF search's knobs on the caller's own batch: the
// list-scan tile configuration, the rows of a work item (the work partition of
// the persistent scan), the first-list threshold tile and the shortlist.  Each
// setting runs the search twice and keeps the second time (HIP events on the
// calling thread's stream); the best is stored for the shape.
AutotuneResult autotune_ivf(const mann_ivf& iv, const QueryOps& Q, int k, int nprobe, int metric,
                            const float* route, int64_t* ids, float* dist) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  AutotuneResult R;
  MANN_REQUIRE(Q.nq > 0 && k > 0, ERR_ARGUMENT, "autotuning needs queries and k > 0");
  nprobe = std::max(1, std::min(nprobe, std::min(iv.nb, iv.nlist)));
  struct Guard {  // the thread's knobs and shape lookup come back whatever happens
    Tuning saved = tuning();
    ~Guard() {
      tuning() = saved;
      shape_tuning_disabled() = false;
    }
  } guard;
  shape_tuning_disabled() = true;
  hipEvent_t e0, e1;
  MANN_HIP(hipEventCreate(&e0));
  MANN_HIP(hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } ev{e0, e1};
  auto time_once = [&](const Tuning& t) {
    tuning() = t;
    ivf_search(iv, Q, k, nprobe, metric, route, ids, dist, 0.0f, 0, 0);  // (warm)
    MANN_HIP(hipEventRecord(e0, st));
    ivf_search(iv, Q, k, nprobe, metric, route, ids, dist, 0.0f, 0, 0);
    MANN_HIP(hipEventRecord(e1, st));
    MANN_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    MANN_HIP(hipEventElapsedTime(&ms, e0, e1));
    ++R.trials;
    return (double)ms;
  };
  Tuning best = guard.saved;
  double best_ms = time_once(best);
  auto sweep = [&](auto setter, std::initializer_list<int> values) {
    for (int v : values) {
      Tuning t = best;
      setter(t, v);
      const double ms = time_once(t);
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
  };
  // (the 4-stage configurations exceed the LDS with the gather buffers)
  sweep([](Tuning& t, int v) { t.ivf_cfg = v; }, {5, 0, 2, 1});
  sweep([](Tuning& t, int v) { t.ivf_rows_item = v; }, {1024, 2048, 4096});
  sweep([](Tuning& t, int v) { t.ivf_tau_cfg = v; }, {9, 8, 0});
  sweep([](Tuning& t, int v) { t.shortlist_mult = v; }, {1, 2});
  R.best = best;
  R.ms = best_ms;
  shape_tuning_put(ivf_shape_key(iv, Q.nq, k, nprobe, metric), best);
  return R;
}

}  // namespace mann
