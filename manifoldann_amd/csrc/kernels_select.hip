// Per-query selection: threshold estimation from a row sample, shortlist
// selection + FP32 exact re-rank with certification, and the exact fallback.
//
// Exact distances follow the oracle's A16 order: a 16-lane group owns one
// candidate, lane j accumulates elements i = j (mod 16) with fmaf in increasing
// i, and the group reduces with the 8/4/2/1 xor tree.  Ordering is by
// (distance, row) with the row in the scanned layout; for the IVF layout
// (lists contiguous, ascending ids inside a list) that is (dist, bucket, id),
// the batched reference merge order.
#include "bounds.h"
#include "radix.h"

namespace mann {

namespace {

constexpr int SEL_CAP = 1024;   // max candidates kept per query (LDS sort size)
constexpr int EX_BATCH = 256;   // rows per batch in the exact fallback
constexpr int EX_KMAX = 256;    // max k_req of the exact fallback
constexpr int EX_SEG = 4096;    // rows per block of the exact fallback (at most)
constexpr int EX_MCAP = 4096;   // keys the merge holds per query
constexpr int EX_MPER = 8;      // keys each merge thread reads per round

__global__ void fill_u64_kernel(unsigned long long* p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ void bitonic_sort(uint64_t* s, int P) {
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint64_t a = s[i], b = s[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

__device__ __forceinline__ int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// rank-th smallest of each query's m sample values (one block per query)
// one workgroup per query (long rows)
__global__ __launch_bounds__(256) void tau_from_sample_block_kernel(TauArgs a) {
  __shared__ BlockRankScratch sm;
  const int q = blockIdx.x;
  if (q >= a.nq) return;
  const uint32_t key = block_rank_key(a.dense + (int64_t)q * a.ldd, a.m, a.rank, sm);
  if (threadIdx.x == 0) a.tau[q] = ord2f(key);
}

// one wavefront per query, four per workgroup
__global__ __launch_bounds__(256) void tau_from_sample_kernel(TauArgs a) {
  __shared__ uint32_t buf_all[4][256];
  __shared__ unsigned hist_all[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = blockIdx.x * 4 + w;
  if (q >= a.nq) return;
  const uint32_t key =
      wave_rank_key(a.dense + (int64_t)q * a.ldd, a.m, a.rank, buf_all[w], hist_all[w], lane);
  if (lane == 0) a.tau[q] = ord2f(key);
}

// ---- wave-per-query selection ------------------------------------------------

constexpr int KPL = SEL_CAP / 64;   // candidate keys held per lane
constexpr int EXMAX = 512;          // exactly re-ranked candidates per query
constexpr int EXR = EXMAX / 64;     // per lane, in the register sort

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bitonic sort of N (64..EXMAX, power of two) keys: element e lives in lane
// e % 64, register e / 64.
template <int R>
__device__ __forceinline__ void wave_sort(uint64_t (&v)[R], int N, int lane) {
  for (int size = 2; size <= N; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 64) {
        const int sr = stride >> 6;
#pragma unroll
        for (int jj = 0; jj < R; ++jj) {
          if (jj * 64 >= N || (jj & sr)) continue;
          const int e = jj * 64 + lane;
          const bool asc = (e & size) == 0;
          uint64_t lo = v[jj], hi = v[jj | sr];
          if ((lo > hi) == asc) {
            v[jj] = hi;
            v[jj | sr] = lo;
          }
        }
      } else {
#pragma unroll
        for (int jj = 0; jj < R; ++jj) {
          if (jj * 64 >= N) continue;
          const int e = jj * 64 + lane;
          const uint64_t o = __shfl_xor(v[jj], stride);
          const bool asc = (e & size) == 0;
          const bool low = (lane & stride) == 0;
          const uint64_t mn = v[jj] < o ? v[jj] : o, mx = v[jj] < o ? o : v[jj];
          v[jj] = (low == asc) ? mn : mx;
        }
      }
    }
  }
}

// Exact metric of up to 4 rows at once for one 16-lane group (A16 order each).
__device__ __forceinline__ void exact_metric4(const SelectArgs& a, const uint32_t* rows, int nr,
                                              int64_t q, int j, float* out) {
  const float* y = a.q32 + q * a.d16;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  const float* x[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) x[u] = a.x32 + (int64_t)rows[u < nr ? u : 0] * a.d16;
  const bool diff = a.metric == EUCLIDEAN || a.metric == SQEUCLIDEAN;
  auto step = [&](float yv, const float* xv) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (diff) {
        const float df = xv[u] - yv;
        acc[u] = __builtin_fmaf(df, df, acc[u]);
      } else {
        acc[u] = __builtin_fmaf(xv[u], yv, acc[u]);
      }
    }
  };
  // eight steps per iteration with all 40 loads issued first (the chains stay
  // in ascending t), so the row gathers overlap instead of waiting step by step
  constexpr int SB = 8;
  int t = 0;
  for (; t + 16 * SB <= a.d16; t += 16 * SB) {
    float yv[SB], xv[SB][4];
#pragma unroll
    for (int s2 = 0; s2 < SB; ++s2) {
      yv[s2] = y[t + 16 * s2 + j];
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[s2][u] = x[u][t + 16 * s2 + j];
    }
#pragma unroll
    for (int s2 = 0; s2 < SB; ++s2) step(yv[s2], xv[s2]);
  }
  for (; t < a.d16; t += 16) {
    float xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) xv[u] = x[u][t + j];
    step(y[t + j], xv);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    float v = group16_tree(acc[u]);
    const int64_t r = rows[u < nr ? u : 0];
    if (a.metric == EUCLIDEAN) v = __builtin_sqrtf(v);
    else if (a.metric == COSINE) {
      float cv = 1.0f - v / (__builtin_sqrtf(a.xn2[r]) * __builtin_sqrtf(a.qn2[q]));
      v = cv > 0.0f ? cv : 0.0f;
    } else if (a.metric == NORMS_FORM) {
      float tt = a.xn2[r] + a.qn2[q];
      float sq = tt - 2.0f * v;
      v = __builtin_sqrtf(sq > 0.0f ? sq : 0.0f);
    }
    out[u] = v;
  }
}

// Re-rank ex[from, to) (candidate keys -> exact keys) with the wave's 4 groups.
__device__ __forceinline__ void rerank_range(const SelectArgs& a, uint64_t* ex, int from, int to,
                                             int64_t q, int lane) {
  const int g = lane >> 4, j = lane & 15;
  for (int i0 = from + g * 4; i0 < to; i0 += 16) {
    const int nr = min(4, to - i0);
    uint32_t rows[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) rows[u] = u < nr ? (uint32_t)(ex[i0 + u] & 0xffffffffu) : 0u;
    float dv[4];
    exact_metric4(a, rows, nr, q, j, dv);
    if (j == 0)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (u < nr) ex[i0 + u] = pack_key(dv[u], rows[u]);
  }
  wave_sync();
}

__global__ __launch_bounds__(256) void select_rerank_kernel(SelectArgs a) {
  __shared__ unsigned hist_all[4][256];
  __shared__ uint64_t ex_all[4][EXMAX];
  __shared__ int pre_all[4][64];  // regions per query: at most 64
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nq = a.nq_dev ? *a.nq_dev : a.nq;
  const int slot = blockIdx.x * 4 + w;
  if (slot >= nq) return;
  const int q = a.qlist ? a.qlist[slot] : slot;
  unsigned* hist = hist_all[w];
  uint64_t* ex = ex_all[w];
  int* pre = pre_all[w];

  // ---- candidates of all regions, KPL keys per lane
  int cp = 0;
  bool ovf = false;
  if (lane < a.parts) {
    const int v = a.cnt[(int64_t)q * a.parts + lane];
    ovf = v > a.cap;
    cp = v < a.cap ? v : a.cap;
  }
  bool overflow = __ballot(ovf) != 0;
  int inc = cp;
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane < a.parts) pre[lane] = inc - cp;  // exclusive prefix
  const int c = __shfl(inc, 63);
  wave_sync();
  // candidate i lies in the last region p with pre[p] <= i (binary search;
  // empty regions repeat a prefix); all addresses first, then all loads
  int64_t off[KPL];
#pragma unroll
  for (int t = 0; t < KPL; ++t) {
    const int i = t * 64 + lane;
    int lo = 0, hi = t * 64 < c ? a.parts - 1 : 0;  // (no search past the candidates)
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    off[t] = i < c ? (((int64_t)q * a.parts + lo) * (int64_t)a.cap + (i - pre[lo])) * 2 : -1;
  }
  uint2 ev[KPL];
#pragma unroll
  for (int t = 0; t < KPL; ++t) ev[t] = *(const uint2*)(a.cand + (off[t] >= 0 ? off[t] : 0));
  uint64_t key[KPL];
#pragma unroll
  for (int t = 0; t < KPL; ++t) {
    uint64_t kv = ~0ull;
    if (off[t] >= 0) {
      const uint2 e = ev[t];
      kv = pack_key(__uint_as_float(e.x), e.y);
    }
    key[t] = kv;
  }

  // ---- shortlist cutoff: the s-th smallest a' (radix select on its order key)
  const int s = c < a.shortlist ? c : a.shortlist;
  uint32_t cut = 0;
  if (s > 0) {
    uint32_t prefix = 0, mask = 0;
    unsigned want = (unsigned)s;
    for (int shift = 24; shift >= 0; shift -= 8) {
#pragma unroll
      for (int b = 0; b < 4; ++b) hist[4 * lane + b] = 0;
      wave_sync();
#pragma unroll
      for (int t = 0; t < KPL; ++t) {
        const uint32_t o = (uint32_t)(key[t] >> 32);
        if (t * 64 + lane < c && (o & mask) == prefix) atomicAdd(&hist[(o >> shift) & 255u], 1u);
      }
      wave_sync();
      const unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                     h3 = hist[4 * lane + 3];
      const unsigned tot = h0 + h1 + h2 + h3;
      unsigned cum = tot;
      for (int o = 1; o < 64; o <<= 1) {
        unsigned t = __shfl_up(cum, o);
        if (lane >= o) cum += t;
      }
      const unsigned excl = cum - tot;
      const bool here = excl < want && cum >= want;
      const unsigned long long hm = __ballot(here);
      const int src = __ffsll((long long)hm) - 1;
      unsigned bin = 0, nw = 0;
      if (here) {
        unsigned cc = excl;
        if (cc + h0 >= want) { bin = 4 * lane; nw = want - cc; }
        else if (cc + h0 + h1 >= want) { bin = 4 * lane + 1; nw = want - cc - h0; }
        else if (cc + h0 + h1 + h2 >= want) { bin = 4 * lane + 2; nw = want - cc - h0 - h1; }
        else { bin = 4 * lane + 3; nw = want - cc - h0 - h1 - h2; }
      }
      bin = __shfl(bin, src);
      want = __shfl(nw, src);
      prefix |= bin << shift;
      mask |= 255u << shift;
      wave_sync();
    }
    cut = prefix;
  }

  // ---- S1 = candidates with a' <= cutoff (ties included), compacted into ex
  int n1 = 0;
  bool exovf = false;
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int t = 0; t < KPL; ++t) {
    const bool pred = s > 0 && t * 64 + lane < c && (uint32_t)(key[t] >> 32) <= cut;
    const unsigned long long m = __ballot(pred);
    const int pos = n1 + __popcll(m & lt);
    if (pred && pos < EXMAX) ex[pos] = key[t];
    n1 += __popcll(m);
  }
  if (n1 > EXMAX) {
    exovf = true;
    n1 = EXMAX;
  }
  wave_sync();
  rerank_range(a, ex, 0, n1, q, lane);

  // ---- exact order, D_k, exclusion threshold, second batch
  const int k_eff = a.k_eff_q ? a.k_eff_q[q] : a.k_eff;
  uint64_t v[EXR];
  auto sort_ex = [&](int n) {
    int N = 64;
    while (N < n) N <<= 1;
#pragma unroll
    for (int jj = 0; jj < EXR; ++jj) {
      const int e = jj * 64 + lane;
      v[jj] = e < n ? ex[e] : ~0ull;
    }
    wave_sort(v, N, lane);
#pragma unroll
    for (int jj = 0; jj < EXR; ++jj) {
      const int e = jj * 64 + lane;
      if (e < n) ex[e] = v[jj];
    }
    wave_sync();
  };
  auto threshold = [&](int n) -> double {
    double T = INFINITY;
    if (n >= k_eff && k_eff > 0) {
      const uint64_t kk = ex[k_eff - 1];
      T = exclusion_threshold(ord2f((uint32_t)(kk >> 32)), a.metric, a.qe[q], a.qr[q],
                              a.ex, a.rx, a.d);
    }
    return T;
  };
  // exclusion threshold in the first pass's own domain (cert_* bounds)
  auto cert_threshold = [&](int n) -> double {
    double T = INFINITY;
    if (n >= k_eff && k_eff > 0) {
      const uint64_t kk = ex[k_eff - 1];
      T = exclusion_threshold(ord2f((uint32_t)(kk >> 32)), a.metric, a.cert_qe[q],
                              a.cert_qr[q], a.cert_ex, a.cert_rx, a.d, a.cert_g);
    }
    return T;
  };
  sort_ex(n1);
  double T = threshold(n1);
  int n = n1;
  {
    // S2: cutoff < a' and a' <= T
#pragma unroll
    for (int t = 0; t < KPL; ++t) {
      const uint32_t o = (uint32_t)(key[t] >> 32);
      const bool pred = t * 64 + lane < c && (s == 0 || o > cut) &&
                        (double)ord2f(o) <= T;
      const unsigned long long m = __ballot(pred);
      const int pos = n + __popcll(m & lt);
      if (pred && pos < EXMAX) ex[pos] = key[t];
      n += __popcll(m);
    }
    if (n > EXMAX) {
      exovf = true;
      n = EXMAX;
    }
    wave_sync();
    if (n > n1) {
      rerank_range(a, ex, n1, n, q, lane);
      sort_ex(n);
      T = threshold(n);
    }
  }

  // ---- certification and outputs
  const float tq = a.tau[q];
  const bool bad = overflow || exovf;
  bool certified;
  double Tc = T;  // exclusion threshold in the domain of tau
  if (a.cert_qe) {
    // first pass in its own domain: its threshold covers its exclusion
    // threshold, and every re-scored value left out lies above the bf16 one
    Tc = cert_threshold(n);
    const double dm = (double)a.dropmin[q];
    certified = !bad && n >= k_eff && (tq == INFINITY || (double)tq > Tc) && dm > T;
  } else if (bad) {
    certified = false;
  } else if (tq == INFINITY) {
    certified = true;
  } else {
    certified = (n >= k_eff) && ((double)tq > T);
  }
  if (lane == 0) {
    if (!certified) {
      const int f = atomicAdd(a.flag_count, 1);
      a.flag_list[f] = q;
      // the k-th exact distance among real rows bounds the true k-th from
      // above, so its exclusion threshold is a valid emission threshold
      const double Tn = a.tau_new_bf16 ? T : Tc;
      float tn = INFINITY;
      if (n >= k_eff && Tn < 3.0e38) tn = nextafterf((float)Tn, INFINITY);
      // the sample threshold of the second pass's own arithmetic, when it is
      // tighter (an overflowed first pass can leave a loose k-th distance)
      if (a.tau_alt) tn = fminf(tn, a.tau_alt[q]);
      // no progress possible: emit everything (with a separate first pass, a
      // tighter threshold is progress too: it drops candidates)
      if (!a.cert_qe && !(tn > tq) && !overflow) tn = INFINITY;
      a.tau_new[q] = tn;
    }
    if (a.stats) {
      int* sp = a.stats + (slot % kStatSlots) * 4;
      atomicAdd(sp, n);
      if (bad) atomicAdd(sp + 1, 1);
      if (!a.no_emit_stat) atomicAdd(sp + 2, c);
    }
  }
  const int lim = n < a.k_req ? n : a.k_req;
  int64_t* oi = a.out_ids + (int64_t)q * a.k;
  float* od = a.out_dist + (int64_t)q * a.k;
  // position of the query itself among the first lim (kNN graph self removal)
  int selfpos = -1;
  for (int e0 = 0; e0 < lim; e0 += 64) {
    const int e = e0 + lane;
    bool is_self = false;
    if (a.exclude_self && e < lim) {
      const uint32_t r = (uint32_t)(ex[e] & 0xffffffffu);
      const int64_t id = a.perm ? (int64_t)a.perm[r] : (int64_t)r;
      is_self = id == q;
    }
    const unsigned long long m = __ballot(is_self);
    if (m && selfpos < 0) selfpos = e0 + __ffsll((long long)m) - 1;
  }
  const int count = (selfpos >= 0 ? lim - 1 : lim) < a.k ? (selfpos >= 0 ? lim - 1 : lim) : a.k;
  for (int e = lane; e < lim; e += 64) {
    if (e == selfpos) continue;
    const int o = (selfpos >= 0 && e > selfpos) ? e - 1 : e;
    if (o >= a.k) continue;
    const uint64_t kk = ex[e];
    const uint32_t r = (uint32_t)(kk & 0xffffffffu);
    const int64_t id = a.perm ? (int64_t)a.perm[r] : (int64_t)r;
    oi[o] = id + 1;
    od[o] = ord2f((uint32_t)(kk >> 32));
  }
  for (int o = count + lane; o < a.k; o += 64) {
    oi[o] = 0;
    od[o] = INFINITY;
  }
  if (a.out_count && lane == 0) a.out_count[q] = count;
}

// Exact fallback, stage 1: block (slot, seg) computes the exact metric of the
// rows of one segment (at most EX_SEG rows inside one range of the query) and
// keeps their kr smallest (dist, row) keys.  Rows whose key cannot enter the
// running top kr are dropped before any sort, so the sorts stop once the
// threshold has settled.
__global__ __launch_bounds__(256) void exact_partial_kernel(ExactArgs a) {
  const int slot = a.slot_base + (int)blockIdx.x;
  const int tid = threadIdx.x, g = tid >> 4, j = tid & 15;
  const int kr = a.k_req < EX_KMAX ? a.k_req : EX_KMAX;
  uint64_t* out = a.partial + ((int64_t)blockIdx.x * gridDim.y + blockIdx.y) * kr;
  __shared__ uint64_t m[EX_BATCH + EX_KMAX];
  __shared__ int s_rb, s_re, s_any;
  if (slot >= *a.nq_dev) return;
  const int q = a.qlist[slot];
  if (tid == 0) {  // locate segment blockIdx.y among the segments of the ranges
    const int nr = a.range_cnt ? a.range_cnt[q] : 1;
    int seg = (int)blockIdx.y, rb = 0, re = 0;
    for (int ri = 0; ri < nr; ++ri) {
      int lo = 0, hi = a.n_rows;
      if (a.range_cnt) {
        const int2 rg = a.ranges[(int64_t)q * a.range_stride + ri];
        lo = rg.x;
        hi = rg.y;
      }
      const int ns = (hi - lo + a.seg_rows - 1) / a.seg_rows;
      if (seg < ns) {
        rb = lo + seg * a.seg_rows;
        re = min(hi, rb + a.seg_rows);
        break;
      }
      seg -= ns;
    }
    s_rb = rb;
    s_re = re;
  }
  for (int i = tid; i < EX_BATCH + EX_KMAX; i += 256) m[i] = ~0ull;
  __syncthreads();
  const int rb = s_rb, re = s_re;
  uint64_t thr = ~0ull;
  unsigned long long* bound = a.bound + blockIdx.x;
  for (int base = rb; base < re; base += EX_BATCH) {
    // keys at or above the query's bound (another segment's kr-th key) cannot
    // be among its kr smallest
    if (tid == 0) s_any = 0;
    thr = min(thr, (uint64_t)__hip_atomic_load(bound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    __syncthreads();
    for (int i = g; i < EX_BATCH; i += 16) {
      const int r = base + i;
      uint64_t key = ~0ull;
      if (r < re) {
        const float dv = exact_metric(a.x32, a.q32, a.xn2, a.qn2, a.d16, a.metric, r, q, j);
        key = pack_key(dv, (uint32_t)r);
        if (key >= thr) key = ~0ull;
      }
      if (j == 0) {
        m[i] = key;
        if (key != ~0ull) s_any = 1;
      }
    }
    __syncthreads();
    if (s_any) {
      bitonic_sort(m, EX_BATCH + EX_KMAX);
      const uint64_t keep = tid < kr ? m[tid] : ~0ull;
      __syncthreads();
      for (int i = tid; i < EX_BATCH + EX_KMAX; i += 256) m[i] = ~0ull;
      __syncthreads();
      if (tid < kr) m[EX_BATCH + tid] = keep;
      __syncthreads();
      thr = m[EX_BATCH + kr - 1];
      if (tid == 0 && thr != ~0ull) atomicMin(bound, (unsigned long long)thr);
    }
  }
  if (tid < kr) out[tid] = m[EX_BATCH + tid];
}

// Exact fallback, stage 2: merge the per-segment top kr of one query and write
// its result (self removal by id for the kNN graph).
__global__ __launch_bounds__(256) void exact_merge_kernel(ExactArgs a) {
  const int slot = a.slot_base + (int)blockIdx.x;
  if (slot >= *a.nq_dev) return;
  const int q = a.qlist[slot];
  const int tid = threadIdx.x;
  const int kr = a.k_req < EX_KMAX ? a.k_req : EX_KMAX;
  __shared__ uint64_t m[EX_MCAP];
  __shared__ int s_n;
  __shared__ uint64_t s_cut;
  const uint64_t* part = a.partial + (int64_t)blockIdx.x * a.nseg * kr;
  const int64_t total = (int64_t)a.nseg * kr;
  // Every segment key at or below the cut is collected in parallel, EX_MPER
  // keys per thread per round; the kr smallest keys are among them.  The cut
  // starts at the bound (another segment's kr-th key) and, whenever the
  // buffer has no room for another round, the buffer is sorted, cut back to
  // its kr smallest and the cut lowered to the kr-th of them.
  if (tid == 0) {
    s_n = 0;
    s_cut = (uint64_t)a.bound[blockIdx.x];
  }
  for (int i = tid; i < EX_MCAP; i += 256) m[i] = ~0ull;
  __syncthreads();
  for (int64_t b0 = 0; b0 < total; b0 += 256 * EX_MPER) {
    const uint64_t cut = s_cut;
    uint64_t key[EX_MPER];
#pragma unroll
    for (int u = 0; u < EX_MPER; ++u) {
      const int64_t i = b0 + u * 256 + tid;
      key[u] = i < total ? part[i] : ~0ull;
    }
#pragma unroll
    for (int u = 0; u < EX_MPER; ++u)
      if (key[u] != ~0ull && key[u] <= cut) m[atomicAdd(&s_n, 1)] = key[u];
    __syncthreads();
    const bool full = s_n > EX_MCAP - 256 * EX_MPER;
    __syncthreads();  // every thread has read s_n before the next round adds to it
    if (full) {
      bitonic_sort(m, EX_MCAP);
      for (int i = kr + tid; i < EX_MCAP; i += 256) m[i] = ~0ull;
      const uint64_t kth = m[kr - 1];
      __syncthreads();
      if (tid == 0) {
        s_n = kr;
        if (kth < s_cut) s_cut = kth;
      }
      __syncthreads();
    }
  }
  {
    int P = 256;
    while (P < s_n) P <<= 1;
    bitonic_sort(m, P);
    // the kr smallest now lead m: move them where the write-out reads them
    const uint64_t keep = tid < kr ? m[tid] : ~0ull;
    __syncthreads();
    if (tid < kr) m[EX_BATCH + tid] = keep;
    __syncthreads();
  }
  if (tid == 0) {
    int o = 0;
    int64_t* oi = a.out_ids + (int64_t)q * a.k;
    float* od = a.out_dist + (int64_t)q * a.k;
    for (int i = 0; i < kr && o < a.k; ++i) {
      uint64_t key = m[EX_BATCH + i];
      if (key == ~0ull) break;
      uint32_t r = (uint32_t)(key & 0xffffffffu);
      int64_t id = a.perm ? (int64_t)a.perm[r] : (int64_t)r;
      if (a.exclude_self && id == q) continue;
      oi[o] = id + 1;
      od[o] = ord2f((uint32_t)(key >> 32));
      ++o;
    }
    if (a.out_count) a.out_count[q] = o;
    for (int i = o; i < a.k; ++i) {
      oi[i] = 0;
      od[i] = INFINITY;
    }
  }
}

}  // namespace

int select_cap_max() { return SEL_CAP; }
int select_exmax() { return EXMAX; }
int exact_kmax() { return EX_KMAX; }

void launch_tau_from_sample(const TauArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  if (a.m > 8192)  // a wavefront per row would stream too long a row alone
    hipLaunchKernelGGL(tau_from_sample_block_kernel, dim3((unsigned)a.nq), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(tau_from_sample_kernel, dim3((unsigned)ceil_div(a.nq, 4)), dim3(256), 0,
                       st, a);
}

void launch_select_rerank(const SelectArgs& a, int nq_max, hipStream_t st) {
  if (nq_max <= 0) return;
  hipLaunchKernelGGL(select_rerank_kernel, dim3((unsigned)ceil_div(nq_max, 4)), dim3(256), 0, st,
                     a);
}

int exact_seg_rows() { return EX_SEG; }
int exact_merge_cap() { return EX_MCAP; }

void launch_exact_topk(const ExactArgs& a, int nslots, hipStream_t st) {
  if (nslots <= 0 || a.nseg <= 0) return;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)ceil_div(nslots, 256)), dim3(256), 0, st,
                     a.bound, nslots, ~0ull);
  hipLaunchKernelGGL(exact_partial_kernel, dim3(nslots, a.nseg), dim3(256), 0, st, a);
  hipLaunchKernelGGL(exact_merge_kernel, dim3(nslots), dim3(256), 0, st, a);
}

}  // namespace mann
