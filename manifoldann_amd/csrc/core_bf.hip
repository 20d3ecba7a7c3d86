// Runtime plumbing, dataset/query preparation and the exact-kNN pipeline
// (BruteForceIndex query and build_knn_graph).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "core.h"
#include "fp8.h"

namespace mann {

// ------------------------------------------------------------------ runtime --

ThreadCtx& ctx() {
  thread_local ThreadCtx c;
  int dev = 0;
  MANN_HIP(hipGetDevice(&dev));
  if (c.device != dev || c.stream == nullptr) {
    // buffers of another device are dropped (not freed: they belong to it)
    c.bufs.clear();
    c.device = dev;
    MANN_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  }
  return c;
}

Timer& timer() {
  static Timer t;
  return t;
}

int timer_begin(const char* name, hipStream_t st) {
  Timer& t = timer();
  // (event creation allocates; a search being captured records no events)
  if (!t.enabled || ctx().arena || ctx().capture) return -1;
  std::lock_guard<std::mutex> g(t.mu);
  Timer::Pending p;
  p.name = name;
  MANN_HIP(hipEventCreate(&p.a));
  MANN_HIP(hipEventCreate(&p.b));
  MANN_HIP(hipEventRecord(p.a, st));
  t.pending.push_back(p);
  return (int)t.pending.size() - 1;
}

void timer_end(int token, hipStream_t st) {
  if (token < 0) return;
  Timer& t = timer();
  std::lock_guard<std::mutex> g(t.mu);
  if (token < (int)t.pending.size()) (void)hipEventRecord(t.pending[token].b, st);
}

void timer_flush() {
  Timer& t = timer();
  std::lock_guard<std::mutex> g(t.mu);
  for (auto& p : t.pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto& e = t.totals[p.name];
      e.first += 1;
      e.second += ms;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  t.pending.clear();
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// The tuning knobs: one process-wide table written under a mutex by
// mann_tune_set, and per host thread a snapshot that every API call refreshes
// on entry, so a call reads one consistent set of knobs however other threads
// retune meanwhile.
namespace {
struct TuningStore {
  std::mutex mu;
  Tuning table;
  uint64_t version = 0;
};
TuningStore& tuning_store() {
  static TuningStore s;
  return s;
}
}  // namespace

Tuning& tuning() {
  thread_local Tuning t;
  return t;
}

void tuning_refresh() {
  thread_local uint64_t seen = ~0ull;
  TuningStore& s = tuning_store();
  std::lock_guard<std::mutex> g(s.mu);
  if (seen != s.version) {
    tuning() = s.table;
    seen = s.version;
  }
}

Tuning tuning_shared() {
  TuningStore& s = tuning_store();
  std::lock_guard<std::mutex> g(s.mu);
  return s.table;
}

void tuning_publish(const Tuning& t) {
  TuningStore& s = tuning_store();
  {
    std::lock_guard<std::mutex> g(s.mu);
    s.table = t;
    ++s.version;
  }
  tuning_refresh();
}

// ---- per-shape knobs chosen by the autotuner -----------------------------------

bool ShapeKey::operator<(const ShapeKey& o) const {
  const int64_t a[] = {path, device, metric, first_pass, self_join, k, n, d, nq_bucket, nprobe,
                       lists};
  const int64_t b[] = {o.path, o.device, o.metric, o.first_pass, o.self_join, o.k, o.n, o.d,
                       o.nq_bucket, o.nprobe, o.lists};
  return std::lexicographical_compare(a, a + 11, b, b + 11);
}

namespace {
// the knobs a shape entry carries: the IVF search's, or the exact search's
void copy_search_knobs(const ShapeKey& key, const Tuning& from, Tuning& to) {
  if (key.path == 1) {
    to.ivf_cfg = from.ivf_cfg;
    to.ivf_rows_item = from.ivf_rows_item;
    to.ivf_tau_cfg = from.ivf_tau_cfg;
    to.shortlist_mult = from.shortlist_mult;
    return;
  }
  to.scan_cfg = from.scan_cfg;
  to.emit_per_k = from.emit_per_k;
  to.shortlist_mult = from.shortlist_mult;
  to.scan8_cfg = from.scan8_cfg;
  to.fp8_emit_per_k = from.fp8_emit_per_k;
}
std::map<ShapeKey, Tuning>& shape_table() {
  static std::map<ShapeKey, Tuning> m;  // guarded by the tuning store's mutex
  return m;
}
}  // namespace

void shape_tuning_put(const ShapeKey& key, const Tuning& t) {
  TuningStore& s = tuning_store();
  std::lock_guard<std::mutex> g(s.mu);
  shape_table()[key] = t;
}

bool shape_tuning_apply(const ShapeKey& key, Tuning& t) {
  TuningStore& s = tuning_store();
  std::lock_guard<std::mutex> g(s.mu);
  auto it = shape_table().find(key);
  if (it == shape_table().end()) return false;
  copy_search_knobs(key, it->second, t);
  return true;
}

void shape_tuning_clear() {
  TuningStore& s = tuning_store();
  std::lock_guard<std::mutex> g(s.mu);
  shape_table().clear();
}

bool& shape_tuning_disabled() {
  thread_local bool off = false;
  return off;
}

ShapeKey bf_shape_key(const mann_dataset& ds, int64_t nq, int k_req, int metric, bool self_join) {
  ShapeKey key;
  key.path = 0;
  key.device = ds.device;
  key.metric = metric;
  key.first_pass = ds.first_pass;
  key.self_join = self_join ? 1 : 0;
  key.k = k_req;
  key.n = ds.n;
  key.d = ds.d;
  int64_t b = 1;
  while (b < nq) b <<= 1;
  key.nq_bucket = b;
  return key;
}

SearchStats& last_stats() {
  thread_local SearchStats s;
  return s;
}

void add_stats(const int* slots, SearchStats& s) {
  for (int i = 0; i < kStatSlots; ++i) {
    s.reranked += slots[4 * i];
    s.overflow += slots[4 * i + 1];
    s.emitted += slots[4 * i + 2];
  }
}

// ------------------------------------------------------------------ dataset --

}  // namespace mann

mann_dataset::~mann_dataset() {
  (void)hipSetDevice(device);
  void* ps[] = {x32, xb, xn, err, nrm, xn2, sb, sxn, center, qfold, sbq, x8, xn8, err8, nrm8};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

namespace mann {


mann_dataset* create_dataset(const float* data, int64_t d, int64_t n, int64_t ld, int metric) {
  MANN_REQUIRE(d > 0, ERR_ARGUMENT, "Dataset must have at least one dimension");
  MANN_REQUIRE(n > 0, ERR_ARGUMENT, "Dataset must contain at least one point");
  MANN_REQUIRE(ld >= d, ERR_ARGUMENT, "leading dimension smaller than d");
  MANN_REQUIRE(n < (int64_t)1 << 31, ERR_UNSUPPORTED, "more than 2^31 points");
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  auto* ds = new mann_dataset();
  ds->device = c.device;
  ds->d = d;
  ds->n = n;
  ds->d16 = round_up(d, 16);
  ds->kp = round_up(d, 32);  // multiple of the scan K chunk
  ds->n_pad = round_up(n, 256) + 256;
  ds->normalized = metric == COSINE;
  try {
    const float* src = data;
    float* staged = nullptr;
    if (!is_device_ptr(data)) {
      MANN_HIP(hipMalloc(&staged, sizeof(float) * (size_t)(ld * (n - 1) + d)));
      MANN_HIP(hipMemcpy(staged, data, sizeof(float) * (size_t)(ld * (n - 1) + d),
                         hipMemcpyHostToDevice));
      src = staged;
    }
    MANN_HIP(hipMalloc(&ds->x32, sizeof(float) * ds->n_pad * ds->d16));
    MANN_HIP(hipMalloc(&ds->xb, sizeof(bf16_t) * ds->n_pad * ds->kp));
    MANN_HIP(hipMalloc(&ds->xn, sizeof(float) * ds->n_pad));
    MANN_HIP(hipMalloc(&ds->err, sizeof(float) * ds->n_pad));
    MANN_HIP(hipMalloc(&ds->nrm, sizeof(float) * ds->n_pad));
    MANN_HIP(hipMalloc(&ds->xn2, sizeof(float) * ds->n_pad));
    unsigned* mx = c.buf("ds_max").as<unsigned>(2);
    fill_i32((int*)mx, 2, 0, st);
    if (!ds->normalized) {
      // first pass on centred vectors: |x - mean| is smaller than |x|, so the
      // bf16 rounding bounds (and the certified band) shrink
      MANN_HIP(hipMalloc(&ds->center, sizeof(float) * d));
      double* tmp = c.buf("ds_colsum").as<double>(d);
      launch_mean_rows(src, ld, n, (int)d, tmp, ds->center, st);
    }
    launch_prep_rows(src, ld, n, ds->n_pad, (int)d, (int)ds->d16, (int)ds->kp, ds->normalized,
                     ds->x32, ds->xb, ds->xn, ds->err, ds->nrm, mx, mx + 1, st, ds->center);
    launch_a16_norms(ds->x32, n, (int)ds->d16, ds->xn2, st);
    if (tuning().fold_norms && ds->kp - d >= 6) {
      MANN_HIP(hipMalloc(&ds->qfold, sizeof(bf16_t) * ds->n_pad * ds->kp));
      launch_fold_queries(ds->xb, ds->xn, n, ds->n_pad, (int)ds->kp, (int)d, ds->qfold, st);
      launch_fold_base(ds->xb, ds->xn, n, (int)ds->kp, (int)d, st);
      ds->fold = 1;
    }
    // stride sample: rows floor(i * n / m)
    // at least n / 40 rows, so that the sample's order statistics stay close to
    // the k-th neighbour distance for large n
    ds->m = (int)std::min<int64_t>(n, std::max<int64_t>(tuning().sample_rows, n / 40));
    ds->m_pad = (int)round_up(ds->m, 256) + 256;
    std::vector<int> sidx(ds->m);
    for (int i = 0; i < ds->m; ++i) sidx[i] = (int)((int64_t)i * n / ds->m);
    int* didx = c.buf("ds_sidx").as<int>(ds->m);
    MANN_HIP(hipMemcpyAsync(didx, sidx.data(), sizeof(int) * ds->m, hipMemcpyHostToDevice, st));
    MANN_HIP(hipMalloc(&ds->sb, sizeof(bf16_t) * (size_t)ds->m_pad * ds->kp));
    MANN_HIP(hipMalloc(&ds->sxn, sizeof(float) * (size_t)ds->m_pad));
    gather_rows_bf16(ds->xb, ds->xn, didx, ds->m, ds->m_pad, (int)ds->kp, ds->sb, ds->sxn, st);
    if (ds->fold) {  // (reads the first d columns of sb only)
      MANN_HIP(hipMalloc(&ds->sbq, sizeof(bf16_t) * (size_t)ds->m_pad * ds->kp));
      launch_fold_queries(ds->sb, ds->sxn, ds->m, ds->m_pad, (int)ds->kp, (int)d, ds->sbq, st);
    }
    unsigned hm[2];
    MANN_HIP(hipMemcpyAsync(hm, mx, sizeof(hm), hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    MANN_HIP(hipGetLastError());
    if (staged) (void)hipFree(staged);
    memcpy(&ds->ex, &hm[0], 4);
    memcpy(&ds->rx, &hm[1], 4);
  } catch (...) {
    delete ds;
    throw;
  }
  return ds;
}

void dataset_set_first_pass(mann_dataset& ds, int kind) {
  MANN_REQUIRE(kind == 0 || kind == 1, ERR_ARGUMENT, "first pass must be 0 (bf16) or 1 (FP8)");
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  if (kind == 0) {
    ds.first_pass = 0;
    MANN_HIP(hipStreamSynchronize(st));
    void* ps[] = {ds.x8, ds.xn8, ds.err8, ds.nrm8};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    ds.x8 = nullptr;
    ds.xn8 = ds.err8 = ds.nrm8 = nullptr;
    return;
  }
  if (ds.x8) {
    ds.first_pass = 1;
    return;
  }
  const int64_t kp8 = round_up(ds.d, 128);  // even number of 64-byte chunks
  Fp8PrepArgs p{};
  p.x32 = ds.x32;
  p.d16 = (int)ds.d16;
  p.d = (int)ds.d;
  p.kp8 = (int)kp8;
  p.n = ds.n;
  p.n_pad = ds.n_pad;
  p.normalize = ds.normalized;
  p.center = ds.center;
  unsigned* mx = c.buf("ds8_max").as<unsigned>(3);
  fill_i32((int*)mx, 3, 0, st);
  launch_fp8_maxabs(p, mx + 2, st);
  unsigned hm[3];
  MANN_HIP(hipMemcpyAsync(hm, mx, sizeof(hm), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  float maxabs;
  memcpy(&maxabs, &hm[2], 4);
  unsigned char* x8 = nullptr;
  float *xn8 = nullptr, *err8 = nullptr, *nrm8 = nullptr;
  try {
    MANN_HIP(hipMalloc(&x8, (size_t)ds.n_pad * kp8));
    MANN_HIP(hipMalloc(&xn8, sizeof(float) * ds.n_pad));
    MANN_HIP(hipMalloc(&err8, sizeof(float) * ds.n_pad));
    MANN_HIP(hipMalloc(&nrm8, sizeof(float) * ds.n_pad));
  } catch (...) {
    void* ps[] = {x8, xn8, err8, nrm8};
    for (void* q : ps)
      if (q) (void)hipFree(q);
    throw;
  }
  p.scale_exp = fp8_scale_exp(maxabs);
  p.x8 = x8;
  p.xn8 = xn8;
  p.err8 = err8;
  p.nrm8 = nrm8;
  p.max_err = mx;
  p.max_nrm = mx + 1;
  launch_fp8_prep(p, st);
  MANN_HIP(hipMemcpyAsync(hm, mx, sizeof(hm), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  MANN_HIP(hipGetLastError());
  ds.kp8 = kp8;
  ds.sexp8 = p.scale_exp;
  ds.x8 = x8;
  ds.xn8 = xn8;
  ds.err8 = err8;
  ds.nrm8 = nrm8;
  memcpy(&ds.ex8, &hm[0], 4);
  memcpy(&ds.rx8, &hm[1], 4);
  ds.first_pass = 1;
}

QueryOps prepare_queries(const float* q, int64_t d, int64_t nq, int64_t ldq, int normalize,
                         int d16, int kp, const char* tag, const float* center) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  std::string t(tag);
  QueryOps Q;
  Q.nq = nq;
  if (nq == 0) return Q;
  int64_t nq_pad = round_up(nq, 256);
  const float* src = q;
  if (!is_device_ptr(q)) {
    float* stg = c.buf(t + "_stage").as<float>((size_t)(ldq * (nq - 1) + d));
    MANN_HIP(hipMemcpyAsync(stg, q, sizeof(float) * (size_t)(ldq * (nq - 1) + d),
                            hipMemcpyHostToDevice, st));
    src = stg;
  }
  float* q32 = c.buf(t + "_q32").as<float>(nq_pad * d16);
  bf16_t* qb = c.buf(t + "_qb").as<bf16_t>(nq_pad * kp);
  float* qn = c.buf(t + "_qn").as<float>(nq_pad);
  float* qe = c.buf(t + "_qe").as<float>(nq_pad);
  float* qr = c.buf(t + "_qr").as<float>(nq_pad);
  float* qn2 = c.buf(t + "_qn2").as<float>(nq_pad);
  {
    ScopedTime tm("prep_queries", st);
    launch_prep_rows(src, ldq, nq, nq_pad, (int)d, d16, kp, normalize, q32, qb, qn, qe, qr,
                     nullptr, nullptr, st, center);
    launch_a16_norms(q32, nq, d16, qn2, st);
  }
  Q.q32 = q32;
  Q.qb = qb;
  Q.qn = qn;
  Q.qe = qe;
  Q.qr = qr;
  Q.qn2 = qn2;
  return Q;
}

QueryOps dataset_as_queries(const mann_dataset& ds) {
  QueryOps Q;
  Q.nq = ds.n;
  Q.q32 = ds.x32;
  Q.qb = ds.xb;
  Q.qn = ds.xn;
  Q.qe = ds.err;
  Q.qr = ds.nrm;
  Q.qn2 = ds.xn2;
  return Q;
}

// ------------------------------------------------------------ exact search --

ExactArgs run_exact_fallback(ExactArgs e, int nflag, int64_t total_rows, int nranges) {
  if (nflag <= 0) return e;
  ThreadCtx& c = ctx();
  const int kr = std::min(e.k_req, exact_kmax());
  // segment rows: few queries get short segments, so that the launch still
  // spreads over about 1,024 workgroups, but not so short that the keys below
  // the segments' common bound (about total_rows * kr / seg of them) overflow
  // half the merge's sort; allocation-free calls keep the longest, which their
  // scratch sizing assumes
  int seg = exact_seg_rows();
  if (!c.arena) {
    const int64_t par = ceil_div(total_rows, std::max<int64_t>(1, 1024 / nflag));
    const int64_t mrg = ceil_div(2 * total_rows * kr, (int64_t)exact_merge_cap());
    seg = (int)std::min<int64_t>(seg, std::max<int64_t>(256, round_up(std::max(par, mrg), 256)));
  }
  const int max_segments = (int)ceil_div(total_rows, seg) + nranges;
  int CH = 2048;
  if (c.arena) {  // as many queries per launch as the scratch holds partials for
    // (a carve takes an eighth more than asked, plus alignment, per buffer)
    const size_t per = (size_t)max_segments * kr * sizeof(uint64_t) + 8;
    const size_t left = c.arena->left();
    CH = (int)std::max<size_t>(
        1, std::min<size_t>(2048, (left > 1024 ? left - 1024 : 0) / (per + per / 8 + 16)));
  }
  e.nseg = max_segments;
  e.seg_rows = seg;
  e.partial = c.buf("ex_partial").as<uint64_t>((size_t)std::min(nflag, CH) * max_segments * kr);
  e.bound = c.buf("ex_bound").as<unsigned long long>((size_t)std::min(nflag, CH));
  for (int base = 0; base < nflag; base += CH) {
    e.slot_base = base;
    launch_exact_topk(e, std::min(CH, nflag - base), c.stream);
  }
  e.slot_base = 0;
  return e;
}

// Emission threshold per query from the dataset's stride sample: the rank-th
// smallest bf16 first-pass value over the sample rows, rank chosen so that
// about `target` rows of the whole base lie below it.
// target2 > 0: a second threshold (tau2) for target2 rows from the same sample.
static void sample_tau(const mann_dataset& ds, const QueryOps& Q, const bf16_t* qop, int nq,
                       int64_t target, int cfg, float* tau, int64_t target2 = 0,
                       float* tau2 = nullptr) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int n = (int)ds.n;
  const int BM = scan_tile_rows(cfg), BN = scan_tile_queries(cfg);
  auto rank_of = [&](int64_t t) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(ds.m, (t * ds.m + n / 2) / n));
  };
  const int rank = rank_of(target);
  const int rank2 = tau2 ? rank_of(target2) : 0;
  // a long sample (rank far below the number of 32-row blocks): only the
  // block minima are stored; the rank-th smallest of them equals the
  // sample's rank-th smallest unless two of its rank smallest values share
  // a block, and is never below it (a threshold estimate either way)
  const bool bmin = ds.m / 32 >= 64 * std::max(rank, rank2);
  // otherwise the minimum of every 4 consecutive sample rows (8 per 32-row
  // block): the rank-th smallest of those minima is never below the sample's
  // rank-th smallest value and equals it unless two of its rank smallest share
  // a run of 4 rows; a quarter of the values are written and ranked
  const bool grp = !bmin && tuning().sample_grp && 8 * ceil_div(ds.m, 32) >= std::max(rank, rank2);
  const int mcols = bmin ? (int)ceil_div(ds.m, 32) : (grp ? (int)(8 * ceil_div(ds.m, 32)) : ds.m);
  // the dense sample block is processed in query chunks of at most 2^28 values
  int64_t qc_max = std::max<int64_t>(BN, ((int64_t)1 << 28) / mcols / BN * BN);
  if (c.arena)  // the sample block in query chunks the scratch holds
    qc_max = std::max<int64_t>(
        BN, std::min<int64_t>(qc_max, (int64_t)(c.arena->left() / 2 / sizeof(float)) / mcols /
                                          BN * BN));
  const int qc = (int)std::min<int64_t>(nq, qc_max);
  float* dense = c.buf("bf_dense").as<float>((size_t)qc * mcols);
  // self-join: the rows are the queries and the sample rows the query
  // columns, so each store instruction writes whole 128-byte row segments of
  // the [query][sample] block (S^ is the same quantity either way round)
  const bool swap = Q.qb == ds.xb && !bmin && !grp;
  for (int q0 = 0; q0 < nq; q0 += qc) {
    const int nqc = std::min(qc, nq - q0);
    if (swap) {
      ScanArgs a{};
      a.cfg = cfg;
      a.xb = ds.xb + (int64_t)q0 * ds.kp;
      a.xn = ds.xn + q0;
      a.qb = ds.fold ? ds.sbq : ds.sb;
      a.qnorm = ds.sxn;
      a.fold = ds.fold;
      a.kp = (int)ds.kp;
      a.d = (int)ds.d;
      a.nq = ds.m;
      const int qt = (int)ceil_div(ds.m, BN);
      const int pd = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(nqc, BM),
                                                                  ceil_div(256, qt)));
      a.rows_per_part = (int)round_up(ceil_div(nqc, pd), BM);
      a.parts = (int)ceil_div(nqc, a.rows_per_part);
      a.qtiles = qt;
      a.n_rows = nqc;
      a.dense = dense;
      a.ldd = ds.m;
      a.dense_t = 1;
      {
        ScopedTime tm("scan_dense", st);
        launch_scan(a, SCAN_DENSE, qt * a.parts, SCAN_KIND_GRID, st);
      }
      TauArgs ta{dense, (int64_t)ds.m, ds.m, rank, tau + q0, nqc};
      ScopedTime tm("tau_sample", st);
      launch_tau_from_sample(ta, st);
      if (tau2) {
        TauArgs tb{dense, (int64_t)ds.m, ds.m, rank2, tau2 + q0, nqc};
        launch_tau_from_sample(tb, st);
      }
      continue;
    }
    ScanArgs a{};
    a.cfg = cfg;
    a.xb = ds.sb;
    a.xn = ds.sxn;
    a.qb = qop + (int64_t)q0 * ds.kp;
    a.qnorm = Q.qn + q0;
    a.fold = ds.fold;
    a.kp = (int)ds.kp;
    a.d = (int)ds.d;
    a.nq = nqc;
    // row partitions only when the query chunk has too few tiles to fill the
    // chip (a large sample; each partition re-reads the chunk's queries)
    const int qt = (int)ceil_div(nqc, BN);
    const int pd = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(ds.m, BM),
                                                                ceil_div(256, qt)));
    a.rows_per_part = (int)round_up(ceil_div(ds.m, pd), BM);
    a.parts = (int)ceil_div(ds.m, a.rows_per_part);
    a.qtiles = qt;
    a.n_rows = ds.m;
    a.dense = dense;
    a.ldd = mcols;
    a.dense_bmin = bmin;
    a.dense_grp = grp ? 1 : 0;
    a.dvec = mcols % 4 == 0;  // row partitions start at multiples of BM
    {
      ScopedTime tm("scan_dense", st);
      launch_scan(a, SCAN_DENSE, qt * a.parts, SCAN_KIND_GRID, st);
    }
    TauArgs ta{dense, (int64_t)mcols, mcols, rank, tau + q0, nqc};
    ScopedTime tm("tau_sample", st);
    launch_tau_from_sample(ta, st);
    if (tau2) {
      TauArgs tb{dense, (int64_t)mcols, mcols, rank2, tau2 + q0, nqc};
      launch_tau_from_sample(tb, st);
    }
  }
}

// FP8 operands of a query batch (scratch of the calling thread).
struct QueryOps8 {
  const unsigned char* q8 = nullptr;
  const float* qn8 = nullptr;
  const float* qe8 = nullptr;
  const float* qr8 = nullptr;
  const float* qmul = nullptr;
};

static QueryOps8 prepare_queries_fp8(const mann_dataset& ds, const QueryOps& Q) {
  ThreadCtx& c = ctx();
  const int64_t nq_pad = round_up(Q.nq, 256);
  QueryOps8 R;
  unsigned char* q8 = c.buf("bf_q8").as<unsigned char>((size_t)nq_pad * ds.kp8);
  float* qn8 = c.buf("bf_qn8").as<float>(nq_pad);
  float* qe8 = c.buf("bf_qe8").as<float>(nq_pad);
  float* qr8 = c.buf("bf_qr8").as<float>(nq_pad);
  float* qmul = c.buf("bf_qmul").as<float>(nq_pad);
  Fp8PrepArgs p{};
  p.x32 = Q.q32;
  p.d16 = (int)ds.d16;
  p.d = (int)ds.d;
  p.kp8 = (int)ds.kp8;
  p.n = Q.nq;
  p.n_pad = nq_pad;
  p.normalize = ds.normalized;
  p.center = ds.center;
  p.scale_exp = -1001;  // a scale per query
  p.other_exp = ds.sexp8;
  p.x8 = q8;
  p.xn8 = qn8;
  p.err8 = qe8;
  p.nrm8 = qr8;
  p.rowmul = qmul;
  {
    ScopedTime tm("prep_queries_fp8", c.stream);
    launch_fp8_prep(p, c.stream);
  }
  R.q8 = q8;
  R.qn8 = qn8;
  R.qe8 = qe8;
  R.qr8 = qr8;
  R.qmul = qmul;
  return R;
}

void bf_search(const mann_dataset& ds, const QueryOps& Q, int k, int metric, int exclude_self,
               int64_t* out_ids, float* out_dist, int* out_count) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  SearchStats& S = last_stats();
  S = SearchStats();
  const int nq = (int)Q.nq;
  if (nq == 0 || k <= 0) return;
  // k beyond the certified selection's capacity: the exact all-distances path
  if (k + (exclude_self ? 1 : 0) > select_exmax() / 2) {
    bf_search_all(ds, Q, k, metric, exclude_self, out_ids, out_dist, out_count);
    return;
  }
  // knobs the autotuner chose for this shape, for this call only
  TuningScope tscope;
  if (!shape_tuning_disabled()) {
    Tuning t = tuning();
    if (shape_tuning_apply(bf_shape_key(ds, nq, k + (exclude_self ? 1 : 0), metric,
                                        Q.qb == ds.xb),
                           t))
      tscope.set(t);
  }
  // FP8 first pass (selected on the dataset; not in allocation-free calls)
  const bool fp8 = ds.first_pass == 1 && ds.x8 != nullptr && c.arena == nullptr;
  // folded-norm query operand: -2 q~ and the parts of |q~|^2 (see mann_dataset)
  const bf16_t* qop = Q.qb;
  if (ds.fold && Q.qb == ds.xb && ds.qfold) {
    qop = ds.qfold;
  } else if (ds.fold) {
    const int64_t nq_pad = round_up(Q.nq, 256);
    bf16_t* qf = c.buf("bf_qfold").as<bf16_t>((size_t)nq_pad * ds.kp);
    launch_fold_queries(Q.qb, Q.qn, Q.nq, nq_pad, (int)ds.kp, (int)ds.d, qf, st);
    qop = qf;
  }
  const int n = (int)ds.n;
  const int k_req = k + (exclude_self ? 1 : 0);
  const int k_eff = std::min(k_req, n);
  const int cfg = tuning().scan_cfg;
  // first-pass tiles (the bf16 sample pass keeps the bf16 configuration)
  const int fcfg = fp8 ? std::min(std::max(tuning().scan8_cfg, 0), scan8_num_cfg() - 1) : cfg;
  const int BM = fp8 ? scan8_tile_rows(fcfg) : scan_tile_rows(cfg);
  const int BN = fp8 ? scan8_tile_queries(fcfg) : scan_tile_queries(cfg);
  // candidates per query over all regions: what the selection holds, or for
  // FP8 what the bf16 re-score holds (it passes the smallest to the selection)
  const int cap_total = fp8 ? rescore_cap() : select_cap_max();
  // row partitions: enough workgroups to fill the chip, at most 32 regions
  const int qtiles = (int)ceil_div(nq, BN);
  const int row_tiles = (int)ceil_div(n, BM);
  int parts =
      std::max(1, std::min(std::min(row_tiles, 32), (int)ceil_div(tuning().target_groups, qtiles)));
  if (tuning().xcd_map == 3)  // whole bands of partitions (see the scan kernel's grid order)
    parts = std::min(std::min(row_tiles, 32), (int)round_up(parts, tuning().xcd_band));
  int rpp = (int)round_up(ceil_div(n, parts), BM);
  parts = (int)ceil_div(n, rpp);
  // kNN graph with folded norms: one triangle of the self-join, each block
  // emitting both directions into twice the regions
  const bool sym = !fp8 && Q.qb == ds.xb && ds.fold && tuning().sym_join && BM == BN &&
                   2 * parts <= 64;
  const int rstride = sym ? 2 * parts : parts;
  const int cap = cap_total / rstride;
  // the refinement pass only sees the flagged queries (few query tiles): more,
  // smaller row partitions keep the chip busy (up to 64 regions per query)
  int rpp2 = (int)round_up(ceil_div(n, std::min(row_tiles, 64)), BM);
  const int parts2 = (int)ceil_div(n, rpp2);
  const int cap2 = cap_total / parts2;

  unsigned* cand = c.buf("bf_cand").as<unsigned>((size_t)nq * cap_total * 2);
  int* cnt = c.buf("bf_cnt").as<int>((size_t)nq * std::max(std::max(rstride, parts2), 64));
  // (the symmetric join stages a whole tile of rows' thresholds: tiles reach
  // up to 255 rows past the last one)
  float* tau = c.buf("bf_tau").as<float>((size_t)round_up(nq, 256) + 256);
  float* tau2 = c.buf("bf_tau2").as<float>(nq);
  int* flags = c.buf("bf_flags").as<int>(8);   // [0] pass1 count, [1] pass2 count
  int* stats = c.buf("bf_stats").as<int>(kStatSlots * 4);
  int* list1 = c.buf("bf_list1").as<int>(nq);
  int* list2 = c.buf("bf_list2").as<int>(nq);
  fill_i32(flags, 8, 0, st);
  fill_i32(stats, kStatSlots * 4, 0, st);

  QueryOps8 Q8;
  float* tau16w = nullptr;  // FP8: bf16 sample threshold for the bf16 second pass
  if (fp8) Q8 = prepare_queries_fp8(ds, Q);
  // ---- emission threshold per query (every row fits its region: emit all)
  if (rpp <= cap) {
    fill_f32(tau, nq, INFINITY, st);
  } else if (fp8) {
    // a bf16 threshold near the (fp8_emit_per_k k_req)-th row, carried into
    // the FP8 domain by both bands
    const int64_t target =
        std::min<int64_t>((int64_t)tuning().fp8_emit_per_k * k_req, cap_total / 4);
    // and the bf16 path's own threshold for the bf16 second pass (used when
    // tighter than the one from the FP8 pass's k-th distance)
    const int64_t target16 = std::min<int64_t>(
        (int64_t)tuning().sample_rank * tuning().emit_per_k * k_req, select_cap_max() / 2);
    float* tau16 = c.buf("bf_tau16").as<float>(nq);
    tau16w = c.buf("bf_tau16w").as<float>(nq);
    sample_tau(ds, Q, qop, nq, target, cfg, tau16, target16, tau16w);
    Fp8TauArgs ta{};
    ta.nq = nq;
    ta.d = (int)ds.d + (ds.fold ? 3 : 0);
    ta.tau16 = tau16;
    ta.qe16 = Q.qe;
    ta.qr16 = Q.qr;
    ta.ex16 = ds.ex;
    ta.rx16 = ds.rx;
    ta.qe8 = Q8.qe8;
    ta.qr8 = Q8.qr8;
    ta.ex8 = ds.ex8;
    ta.rx8 = ds.rx8;
    ta.tau8 = tau;
    launch_fp8_tau(ta, st);
  } else {
    // threshold rank in the sample: the sample's r-th smallest value sits near
    // the (r n / m)-th smallest row; aim at about 15 k_req emissions per query,
    // bounded by half the candidate capacity
    const int64_t target = std::min<int64_t>(
        (int64_t)tuning().sample_rank * tuning().emit_per_k * k_req, cap_total / 2);
    sample_tau(ds, Q, qop, nq, target, cfg, tau);
  }

  // ---- filtered first pass over all rows
  ScanArgs a{};
  if (fp8) {
    a.cfg = fcfg;
    a.x8 = ds.x8;
    a.q8 = Q8.q8;
    a.qmul = Q8.qmul;
    a.xn = ds.xn8;
    a.qnorm = Q8.qn8;
    a.kp = (int)ds.kp8;
  } else {
    a.cfg = cfg;
    a.xb = ds.xb;
    a.xn = ds.xn;
    a.qb = qop;
    a.qnorm = Q.qn;
    a.fold = ds.fold;
    a.kp = (int)ds.kp;
  }
  a.xcd_map = tuning().xcd_map;
  a.xcd_band = tuning().xcd_band;
  a.cpol = tuning().scan_cpol;
  a.qtiles = qtiles;
  a.d = (int)ds.d;
  a.nq = nq;
  a.parts = parts;
  a.rows_per_part = rpp;
  a.n_rows = n;
  a.tau = tau;
  a.cand = cand;
  a.cnt = cnt;
  a.cap = cap;
  auto first_pass = [&](const char* name, int grid, int kind) {
    ScopedTime tm(name, st);
    if (fp8) launch_scan8(a, grid, kind, st);
    else launch_scan(a, SCAN_FILTER, grid, kind, st);
  };
  if (sym) {
    // items (query tile, partition) on and below the diagonal; every region
    // starts empty (transposed emissions append, and a query tile has no
    // direct region in the partitions above it)
    int grid = 0;
    for (int p = 0; p < parts; ++p)
      grid += (int)std::min<int64_t>(qtiles, ceil_div((int64_t)(p + 1) * rpp, BN));
    fill_i32(cnt, (int64_t)nq * rstride, 0, st);
    a.rstride = rstride;
    ScopedTime tm("scan_filter", st);
    launch_scan(a, SCAN_SYM, grid, SCAN_KIND_GRID, st);
    a.rstride = 0;
  } else {
    first_pass("scan_filter", qtiles * parts, SCAN_KIND_GRID);
  }

  // FP8: re-score the survivors on bf16 into one region per query
  const int keep = select_cap_max();
  unsigned* cand_sel = cand;
  int* cnt_sel = cnt;
  float* dropmin = nullptr;
  RescoreArgs ra{};
  if (fp8) {
    cand_sel = c.buf("bf_cand_sel").as<unsigned>((size_t)nq * keep * 2);
    cnt_sel = c.buf("bf_cnt_sel").as<int>(nq);
    dropmin = c.buf("bf_dropmin").as<float>(nq);
    ra.nq = nq;
    ra.cand8 = cand;
    ra.cnt8 = cnt;
    ra.cap8 = cap;
    ra.parts8 = parts;
    ra.xb = ds.xb;
    ra.xn = ds.xn;
    ra.qb = Q.qb;
    ra.qn = Q.qn;
    ra.kp = (int)ds.kp;
    ra.d = (int)ds.d;
    ra.cand = cand_sel;
    ra.cnt = cnt_sel;
    ra.keep = keep;
    ra.dropmin = dropmin;
    ra.stats = stats;
    ScopedTime tm("rescore_bf16", st);
    launch_rescore(ra, nq, st);
  }

  SelectArgs s{};
  s.nq = nq;
  s.cand = cand_sel;
  s.cnt = cnt_sel;
  s.cap = fp8 ? keep : cap;
  s.parts = fp8 ? 1 : rstride;
  s.tau = tau;
  s.x32 = ds.x32;
  s.q32 = Q.q32;
  s.xn2 = ds.xn2;
  s.qn2 = Q.qn2;
  // certification band: with folded norms the MFMA sums d + 6 terms (the
  // bound of d + 3 covers them and the FP32 rounding of the two norms); the
  // bf16 re-score of the FP8 path sums the d products only
  s.d = (int)ds.d + (ds.fold && !fp8 ? 3 : 0);
  s.d16 = (int)ds.d16;
  s.metric = metric;
  s.qe = Q.qe;
  s.qr = Q.qr;
  s.ex = ds.ex;
  s.rx = ds.rx;
  if (fp8) {
    s.cert_qe = Q8.qe8;
    s.cert_qr = Q8.qr8;
    s.cert_ex = ds.ex8;
    s.cert_rx = ds.rx8;
    s.cert_g = kFp8G;
    s.dropmin = dropmin;
    s.no_emit_stat = 1;
    s.tau_new_bf16 = 1;  // the second pass runs on bf16
    s.tau_alt = tau16w;
  }
  s.k = k;
  s.k_req = k_req;
  s.exclude_self = exclude_self;
  s.k_eff = k_eff;
  s.shortlist = std::min(select_exmax() / 2, std::max(tuning().shortlist_mult * k_req, 32));
  s.out_ids = out_ids;
  s.out_dist = out_dist;
  s.out_count = out_count;
  s.flag_count = flags;
  s.flag_list = list1;
  s.tau_new = tau2;
  s.stats = stats;
  {
    ScopedTime tm("select_rerank", st);
    launch_select_rerank(s, nq, st);
  }

  // ---- second pass for queries that failed certification (device-side
  // count), always on bf16: after an FP8 first pass its threshold is the bf16
  // exclusion threshold of the k-th exact distance the FP8 pass found (the
  // much narrower bf16 band resolves what the FP8 band could not hold)
  if (fp8) {
    const int BM16 = scan_tile_rows(cfg), BN16 = scan_tile_queries(cfg);
    const int rt16 = (int)ceil_div(n, BM16);
    // at most 48 regions: the threshold may be the sample's (hundreds of rows
    // per query), and each region holds 1/parts of the candidates
    const int rp2 = (int)round_up(ceil_div(n, std::min(rt16, 48)), BM16);
    const int p2 = (int)ceil_div(n, rp2);
    a = ScanArgs{};
    a.cfg = cfg;
    a.xb = ds.xb;
    a.xn = ds.xn;
    a.qb = qop;
    a.qnorm = Q.qn;
    a.fold = ds.fold;
    a.kp = (int)ds.kp;
    a.xcd_map = tuning().xcd_map;
    a.xcd_band = tuning().xcd_band;
    a.cpol = tuning().scan_cpol;
    a.qtiles = (int)ceil_div(nq, BN16);
    a.d = (int)ds.d;
    a.nq = nq;
    a.n_rows = n;
    a.cand = cand;
    a.cnt = cnt;
    a.qidx = list1;
    a.nq_dev = flags;
    a.tau = tau2;
    a.parts = p2;
    a.rows_per_part = rp2;
    a.cap = select_cap_max() / p2;
    {
      ScopedTime tm("scan_filter_pass2", st);
      launch_scan(a, SCAN_FILTER, a.qtiles * p2, SCAN_KIND_REFINE, st);
    }
    s.cand = cand;
    s.cnt = cnt;
    s.parts = p2;
    s.cap = a.cap;
    s.d = (int)ds.d + (ds.fold ? 3 : 0);
    s.cert_qe = nullptr;
    s.cert_qr = nullptr;
    s.dropmin = nullptr;
    s.no_emit_stat = 0;
  } else {
    a.qidx = list1;
    a.nq_dev = flags;
    a.tau = tau2;
    a.parts = parts2;
    a.rows_per_part = rpp2;
    a.cap = cap2;
    first_pass("scan_filter_pass2", qtiles * parts2, SCAN_KIND_REFINE);
    s.parts = parts2;
    s.cap = cap2;
  }
  s.tau_new_bf16 = 0;
  s.tau_alt = nullptr;
  s.qlist = list1;
  s.nq_dev = flags;
  s.tau = tau2;
  s.flag_count = flags + 1;
  s.flag_list = list2;
  s.tau_new = tau;  // unused afterwards
  {
    ScopedTime tm("select_rerank_pass2", st);
    launch_select_rerank(s, nq, st);
  }
  // ---- exact fallback for what is still uncertified
  ExactArgs e{};
  e.qlist = list2;
  e.nq_dev = flags + 1;
  e.x32 = ds.x32;
  e.q32 = Q.q32;
  e.xn2 = ds.xn2;
  e.qn2 = Q.qn2;
  e.d = (int)ds.d;
  e.d16 = (int)ds.d16;
  e.metric = metric;
  e.n_rows = n;
  e.k = k;
  e.k_req = k_req;
  e.exclude_self = exclude_self;
  e.out_ids = out_ids;
  e.out_dist = out_dist;
  e.out_count = out_count;
  MANN_REQUIRE(k_req <= exact_kmax(), ERR_UNSUPPORTED, "k too large for the exact fallback");
  if (c.capture) {
    // under graph capture: no read-back; the fallback covers the first
    // fallback_slots flagged queries (the kernels skip slots past the device
    // count), the graph's launcher runs the rest if there are more
    CaptureInfo& ci = *c.capture;
    ci.e = run_exact_fallback(e, std::min(nq, ci.fallback_slots), n, 0);
    ci.flags = flags;
    ci.stats = stats;
    ci.filled = true;
    return;
  }
  int hf[8], hs[kStatSlots * 4];
  MANN_HIP(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipMemcpyAsync(hs, stats, sizeof(hs), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  {
    ScopedTime tm("exact_fallback", st);
    run_exact_fallback(e, hf[1], n, 0);
  }
  MANN_HIP(hipStreamSynchronize(st));
  MANN_HIP(hipGetLastError());
  S.pass2 = hf[0];
  S.exact = hf[1];
  add_stats(hs, S);
}

// ---- autotuner -----------------------------------------------------------------
// Coordinate descent over the exact search's knobs on the caller's own batch:
// the tile configuration, then the emission target, then the shortlist (for
// an FP8 dataset: its tile configuration and emission target too).  Each
// setting runs the search twice and keeps the second time (HIP events on the
// calling thread's stream).  The best setting is stored for the shape.
AutotuneResult autotune_bf(const mann_dataset& ds, const QueryOps& Q, int k, int metric,
                           int exclude_self, int64_t* ids, float* dist) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  AutotuneResult R;
  const int nq = (int)Q.nq;
  MANN_REQUIRE(nq > 0 && k > 0, ERR_ARGUMENT, "autotuning needs queries and k > 0");
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
    bf_search(ds, Q, k, metric, exclude_self, ids, dist, nullptr);  // (warm: buffers sized)
    MANN_HIP(hipEventRecord(e0, st));
    bf_search(ds, Q, k, metric, exclude_self, ids, dist, nullptr);
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
  if (ds.first_pass == 1 && ds.x8) {
    sweep([](Tuning& t, int v) { t.scan8_cfg = v; }, {0, 1, 4});
    sweep([](Tuning& t, int v) { t.fp8_emit_per_k = v; }, {2, 3});
  } else {
    sweep([](Tuning& t, int v) { t.scan_cfg = v; }, {5, 6, 1, 3, 0});
    sweep([](Tuning& t, int v) { t.emit_per_k = v; }, {8, 15, 24});
  }
  sweep([](Tuning& t, int v) { t.shortlist_mult = v; }, {1, 2});
  R.best = best;
  R.ms = best_ms;
  shape_tuning_put(bf_shape_key(ds, nq, k + (exclude_self ? 1 : 0), metric, Q.qb == ds.xb), best);
  return R;
}

}  // namespace mann
