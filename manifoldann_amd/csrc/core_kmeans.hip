// k-means on the GPU: KMeansTransform.fit! (KMeansTransform.jl:106-162) over
// lloyd! (kmeans/lloyd.jl:36-134) with kmeans++ / random / given initial
// centroids (kmeans/init.jl:22-128).
//
// E-step: the bf16 MFMA scan in TOP1 mode (points are the queries, centroids
// the rows) emits only centroids that can still be the minimum; the FP32
// norms-form re-rank (A16 order) picks the argmin with ties to the lower index.
// M-step: stable counting sort of the assignment, FP64 sums over fixed
// 256-member chunks, chunk partials summed in order, mean rounded once to FP32.
#include <algorithm>
#include <cstring>
#include <vector>

#include "kmeans.h"
#include "core.h"

namespace mann {

namespace {

constexpr int TOP1_CAP = 64;     // emitted centroids kept per point
constexpr int MSTEP_CHUNK = 256;

// Prepared centroid rows (the scan's base operand) in thread scratch.
struct CentroidOps {
  int k = 0, k_pad = 0;
  float* c32 = nullptr;     // k_pad x d16
  bf16_t* cb = nullptr;     // k_pad x kp
  float* cn = nullptr;      // |c~|^2, +inf on padding
  float* cn2 = nullptr;     // A16 |c|^2
  float ex = 0, rx = 0;
  float ra = 0;  // upper bound of the largest uncentred centroid norm
};

CentroidOps prepare_centroids(const float* C, int k, const mann_dataset& ds, const char* tag) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  std::string t(tag);
  CentroidOps P;
  P.k = k;
  P.k_pad = (int)round_up(k, 256) + 256;
  P.c32 = c.buf(t + "_c32").as<float>((size_t)P.k_pad * ds.d16);
  P.cb = c.buf(t + "_cb").as<bf16_t>((size_t)P.k_pad * ds.kp);
  P.cn = c.buf(t + "_cn").as<float>(P.k_pad);
  P.cn2 = c.buf(t + "_cn2").as<float>(P.k_pad);
  float* ce = c.buf(t + "_ce").as<float>(P.k_pad);
  float* cr = c.buf(t + "_cr").as<float>(P.k_pad);
  unsigned* mx = c.buf(t + "_mx").as<unsigned>(3);
  fill_i32((int*)mx, 3, 0, st);
  launch_prep_rows(C, ds.d16, k, P.k_pad, (int)ds.d, (int)ds.d16, (int)ds.kp, 0, P.c32, P.cb, P.cn,
                   ce, cr, mx, mx + 1, st, ds.center);
  launch_a16_norms(P.c32, k, (int)ds.d16, P.cn2, st);
  launch_norm_bound_max(P.cn2, k, (int)ds.d, mx + 2, st);
  unsigned h[3];
  MANN_HIP(hipMemcpyAsync(h, mx, sizeof(h), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  memcpy(&P.ex, &h[0], 4);
  memcpy(&P.rx, &h[1], 4);
  memcpy(&P.ra, &h[2], 4);
  return P;
}

}  // namespace

// E-step over the dataset: assign (0-based) and mindist; `primed` uses the
// current content of `assign` to set each point's emission threshold.
void kmeans_estep(const mann_dataset& ds, const float* C, int k, int* assign, float* mindist,
                  bool primed, long long* overflow_out) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int64_t n = ds.n;
  CentroidOps P = prepare_centroids(C, k, ds, "km");
  float* tau = c.buf("km_tau").as<float>(n);
  unsigned* cand = c.buf("km_cand").as<unsigned>((size_t)n * TOP1_CAP * 2);
  int* cnt = c.buf("km_cnt").as<int>(n);
  int* ovf = c.buf("km_ovf").as<int>(1);
  fill_i32(ovf, 1, 0, st);
  Top1Args t{};
  t.n = n;
  t.k = k;
  t.d16 = (int)ds.d16;
  t.x32 = ds.x32;
  t.xn2 = ds.xn2;
  t.c32 = P.c32;
  t.cn2 = P.cn2;
  t.cand = cand;
  t.cnt = cnt;
  t.cap = TOP1_CAP;
  t.assign = assign;
  t.mindist = mindist;
  t.overflow = ovf;
  if (primed) {
    ScopedTime tm("kmeans_prime", st);
    launch_prime_threshold(t, ds.err, ds.nrm, P.ex, P.rx, P.ra, (int)ds.d, tau, st);
  } else {
    fill_f32(tau, n, INFINITY, st);
  }
  ScanArgs a{};
  a.cfg = tuning().kmeans_cfg;
  a.xb = P.cb;
  a.xn = P.cn;
  a.qb = ds.xb;
  a.kp = (int)ds.kp;
  a.d = (int)ds.d;
  a.nq = (int)n;
  a.parts = 1;
  a.rows_per_part = (int)round_up(k, scan_tile_rows(a.cfg));
  a.n_rows = k;
  a.tau = tau;
  a.cand = cand;
  a.cnt = cnt;
  a.cap = TOP1_CAP;
  a.qnorm = ds.xn;
  a.t1_qe = ds.err;
  a.t1_qr = ds.nrm;
  a.t1_qn2 = ds.xn2;
  a.t1_ex = P.ex;
  a.t1_rx = P.rx;
  a.t1_ra = P.ra;
  {
    ScopedTime tm("kmeans_assign_scan", st);
    launch_scan(a, SCAN_TOP1, (int)ceil_div(n, scan_tile_queries(a.cfg)), SCAN_KIND_GRID, st);
  }
  {
    ScopedTime tm("kmeans_assign_select", st);
    launch_top1_select(t, st);
  }
  if (overflow_out) {
    int h = 0;
    MANN_HIP(hipMemcpyAsync(&h, ovf, sizeof(int), hipMemcpyDeviceToHost, st));
    MANN_HIP(hipStreamSynchronize(st));
    *overflow_out += h;
  }
}

// Lloyd iterations from the centroids in C (k x d16, device, updated in place).
int64_t kmeans_lloyd(const mann_dataset& ds, float* C, int k, int64_t max_iters, double tol,
                     int* assign, float* mindist, long long* overflow) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int64_t n = ds.n;
  const int d = (int)ds.d, d16 = (int)ds.d16;
  const int nb = (int)ceil_div(n, sort_block());
  int* table = c.buf("km_table").as<int>((size_t)counting_sort_table_ints(n, k));
  int* offsets = c.buf("km_off").as<int>(k + 1);
  int* members = c.buf("km_members").as<int>(n);
  float* newc = c.buf("km_newc").as<float>((size_t)k * d16);
  float* mv = c.buf("km_mv").as<float>(k);
  unsigned long long* far = c.buf("km_far").as<unsigned long long>(1);
  std::vector<int> hoff(k + 1);
  std::vector<float> hmv(k);
  int64_t it = 0;
  bool primed = false;
  for (int64_t iter = 1; iter <= max_iters; ++iter) {
    it = iter;
    kmeans_estep(ds, C, k, assign, mindist, primed, overflow);
    primed = true;
    {
      ScopedTime tm("kmeans_sort", st);
      launch_counting_sort(assign, n, k, table, offsets, members, st);
    }
    MANN_HIP(hipMemcpyAsync(hoff.data(), offsets, sizeof(int) * (k + 1), hipMemcpyDeviceToHost,
                            st));
    MANN_HIP(hipStreamSynchronize(st));
    hoff[k] = (int)n;
    // chunk list: fixed 256-member chunks inside each cluster
    std::vector<int2> chunks;
    std::vector<int> chunk_off(k + 1), counts(k);
    for (int cl = 0; cl < k; ++cl) {
      chunk_off[cl] = (int)chunks.size();
      counts[cl] = hoff[cl + 1] - hoff[cl];
      for (int s = hoff[cl]; s < hoff[cl + 1]; s += MSTEP_CHUNK)
        chunks.push_back(make_int2(s, std::min(s + MSTEP_CHUNK, hoff[cl + 1])));
    }
    chunk_off[k] = (int)chunks.size();
    const int nch = (int)chunks.size();
    int2* dch = c.buf("km_chunks").as<int2>(std::max(nch, 1));
    int* dcoff = c.buf("km_choff").as<int>(k + 1);
    int* dcnt = c.buf("km_counts").as<int>(k);
    double* part = c.buf("km_part").as<double>((size_t)std::max(nch, 1) * d);
    MANN_HIP(hipMemcpyAsync(dch, chunks.data(), sizeof(int2) * nch, hipMemcpyHostToDevice, st));
    MANN_HIP(hipMemcpyAsync(dcoff, chunk_off.data(), sizeof(int) * (k + 1), hipMemcpyHostToDevice,
                            st));
    MANN_HIP(hipMemcpyAsync(dcnt, counts.data(), sizeof(int) * k, hipMemcpyHostToDevice, st));
    {
      ScopedTime tm("kmeans_update", st);
      launch_chunk_sums(ds.x32, d, d16, members, dch, nch, part, st);
      launch_cluster_means(part, dcoff, dcnt, k, d, d16, newc, st);
    }
    // empty clusters: farthest point (first on ties), marked -Inf (lloyd.jl:81-104)
    for (int cl = 0; cl < k; ++cl) {
      if (counts[cl] != 0) continue;
      fill_i32((int*)far, 2, 0, st);
      launch_farthest(mindist, n, far, st);
      unsigned long long hk = 0;
      MANN_HIP(hipMemcpyAsync(&hk, far, sizeof(hk), hipMemcpyDeviceToHost, st));
      MANN_HIP(hipStreamSynchronize(st));
      int64_t idx = (int64_t)(0xffffffffu - (unsigned)(hk & 0xffffffffu));
      MANN_HIP(hipMemcpyAsync(newc + (int64_t)cl * d16, ds.x32 + idx * d16, sizeof(float) * d16,
                              hipMemcpyDeviceToDevice, st));
      const float ninf = -INFINITY;
      MANN_HIP(hipMemcpyAsync(mindist + idx, &ninf, sizeof(float), hipMemcpyHostToDevice, st));
      MANN_HIP(hipStreamSynchronize(st));
    }
    launch_movement(C, newc, k, d16, mv, st);
    MANN_HIP(hipMemcpyAsync(hmv.data(), mv, sizeof(float) * k, hipMemcpyDeviceToHost, st));
    MANN_HIP(hipMemcpyAsync(C, newc, sizeof(float) * (size_t)k * d16, hipMemcpyDeviceToDevice,
                            st));
    MANN_HIP(hipStreamSynchronize(st));
    double maxmove = 0.0;
    for (int cl = 0; cl < k; ++cl) maxmove = std::max(maxmove, (double)hmv[cl]);
    if (maxmove < tol) break;
  }
  kmeans_estep(ds, C, k, assign, mindist, primed, overflow);
  return it;
}

// kmeans++ with draws u_0..u_{k-1} of the splitmix64 stream of `seed`.
void kmeans_pp(const mann_dataset& ds, int k, uint64_t seed, float* C, int* picked) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  const int64_t n = ds.n;
  std::vector<double> draws(k);
  uint64_t s = seed;
  for (int i = 0; i < k; ++i) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    draws[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
  double* dd = c.buf("pp_draws").as<double>(k);
  double* mins = c.buf("pp_mins").as<double>(n);
  double* bsum = c.buf("pp_bsum").as<double>(ceil_div(n, 1024));
  MANN_HIP(hipMemcpyAsync(dd, draws.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
  ScopedTime tm("kmeanspp", st);
  launch_pp_first(dd, n, ds.x32, (int)ds.d16, C, picked, mins, st);
  for (int r = 1; r < k; ++r)
    launch_pp_round(ds.x32, ds.xn2, n, (int)ds.d16, C, r, mins, bsum, dd, picked, st);
}

// Uniform random distinct indices (first k of a Fisher-Yates permutation with
// the same splitmix64 stream), for init = :random.
std::vector<int> random_indices(int64_t n, int k, uint64_t seed) {
  std::vector<int> p(n);
  for (int64_t i = 0; i < n; ++i) p[i] = (int)i;
  uint64_t s = seed;
  for (int64_t i = n - 1; i >= 1; --i) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    int64_t j = (int64_t)(u * (double)(i + 1));
    if (j > i) j = i;
    std::swap(p[i], p[j]);
  }
  p.resize(k);
  return p;
}

}  // namespace mann
