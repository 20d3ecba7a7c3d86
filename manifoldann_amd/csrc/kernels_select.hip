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

namespace mann {

namespace {

constexpr int SEL_CAP = 1024;   // max candidates kept per query (LDS sort size)
constexpr int EX_BATCH = 256;   // rows per batch in the exact fallback
constexpr int EX_KMAX = 256;    // max k_req of the exact fallback

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

__global__ __launch_bounds__(256) void tau_from_sample_kernel(TauArgs a) {
  const int q = blockIdx.x;
  if (q >= a.nq) return;
  __shared__ unsigned hist[256];
  __shared__ unsigned s_sel, s_want;
  constexpr int PER = 16;  // m <= 4096
  uint32_t v[PER];
  const float* row = a.dense + (int64_t)q * a.ldd;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    int i = threadIdx.x + 256 * p;
    v[p] = i < a.m ? f2ord(row[i]) : 0xffffffffu;
  }
  uint32_t prefix = 0, mask = 0;
  if (threadIdx.x == 0) s_want = (unsigned)a.rank;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p)
      if ((v[p] & mask) == prefix) atomicAdd(&hist[(v[p] >> shift) & 255u], 1u);
    __syncthreads();
    if (threadIdx.x < 64) {
      const int l = threadIdx.x;
      unsigned h0 = hist[4 * l], h1 = hist[4 * l + 1], h2 = hist[4 * l + 2], h3 = hist[4 * l + 3];
      unsigned tot = h0 + h1 + h2 + h3, inc = tot;
      for (int o = 1; o < 64; o <<= 1) {
        unsigned t = __shfl_up(inc, o);
        if (l >= o) inc += t;
      }
      unsigned excl = inc - tot, want = s_want;
      bool here = excl < want && inc >= want;
      if (here) {
        unsigned c = excl, b = 4 * l;
        if (c + h0 >= want) b = 4 * l;
        else if (c + h0 + h1 >= want) { b = 4 * l + 1; c += h0; }
        else if (c + h0 + h1 + h2 >= want) { b = 4 * l + 2; c += h0 + h1; }
        else { b = 4 * l + 3; c += h0 + h1 + h2; }
        s_sel = b;
        s_want = want - c;
      }
    }
    __syncthreads();
    prefix |= s_sel << shift;
    mask |= 255u << shift;
    __syncthreads();
  }
  if (threadIdx.x == 0) a.tau[q] = ord2f(prefix);
}

__global__ __launch_bounds__(256) void select_rerank_kernel(SelectArgs a) {
  const int nq = a.nq_dev ? *a.nq_dev : a.nq;
  if ((int)blockIdx.x >= nq) return;
  const int q = a.qlist ? a.qlist[blockIdx.x] : (int)blockIdx.x;
  __shared__ uint64_t keys[SEL_CAP];
  __shared__ uint64_t ex[SEL_CAP];
  __shared__ int s_m1;
  __shared__ double s_T;
  const int tid = threadIdx.x, g = tid >> 4, j = tid & 15;

  int c = 0;
  bool overflow = false;
  for (int p = 0; p < a.parts; ++p) {
    const int cp = a.cnt[(int64_t)q * a.parts + p];
    overflow |= cp > a.cap;
    c += cp < a.cap ? cp : a.cap;
  }
  const int P = next_pow2(c > 0 ? c : 1);
  for (int i = tid; i < P; i += 256) {
    uint64_t key = ~0ull;
    if (i < c) {
      int p = 0, base = 0;
      for (;; ++p) {
        int cp = a.cnt[(int64_t)q * a.parts + p];
        cp = cp < a.cap ? cp : a.cap;
        if (i < base + cp) break;
        base += cp;
      }
      const unsigned* cq = a.cand + ((int64_t)q * a.parts + p) * (int64_t)a.cap * 2;
      uint2 e = *(const uint2*)(cq + 2 * (i - base));
      key = pack_key(__uint_as_float(e.x), e.y);
    }
    keys[i] = key;
  }
  __syncthreads();
  bitonic_sort(keys, P);

  const double qn = a.qn[q], qe = a.qe[q], qr = a.qr[q];
  const int k_eff = a.k_eff_q ? a.k_eff_q[q] : a.k_eff;
  int done = 0;
  int target = c < a.shortlist ? c : a.shortlist;
  double T = INFINITY;
  for (int round = 0; round < 4; ++round) {
    for (int i = done + g; i < target; i += 16) {
      uint32_t r = (uint32_t)(keys[i] & 0xffffffffu);
      float dv = exact_metric(a.x32, a.q32, a.xn2, a.qn2, a.d16, a.metric, r, q, j);
      if (j == 0) ex[i] = pack_key(dv, r);
    }
    done = target;
    int P2 = next_pow2(done > 0 ? done : 1);
    for (int i = done + tid; i < P2; i += 256) ex[i] = ~0ull;
    __syncthreads();
    bitonic_sort(ex, P2);
    if (tid == 0) {
      double Tn = INFINITY;
      if (done >= k_eff && k_eff > 0) {
        float dk = ord2f((uint32_t)(ex[k_eff - 1] >> 32));
        Tn = exclusion_threshold(dk, a.metric, qn, qe, qr, a.ex, a.rx, a.d);
      }
      // rows with a' <= Tn must be re-ranked: count them in the sorted keys
      int lo = 0, hi = c;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        float av = ord2f((uint32_t)(keys[mid] >> 32));
        if ((double)av <= Tn) lo = mid + 1;
        else hi = mid;
      }
      s_m1 = lo;
      s_T = Tn;
    }
    __syncthreads();
    T = s_T;
    const int m1 = s_m1;
    __syncthreads();
    if (m1 <= done) break;
    target = m1;
  }

  if (tid == 0) {
    const float tq = a.tau[q];
    bool certified;
    if (overflow) certified = false;
    else if (tq == INFINITY) certified = true;
    else certified = (done >= k_eff) && ((double)tq > T);
    if (!certified) {
      int f = atomicAdd(a.flag_count, 1);
      a.flag_list[f] = q;
      float tn = INFINITY;
      if (done >= k_eff && T < 3.0e38) tn = nextafterf((float)T, INFINITY);
      if (!(tn > tq) && !overflow) tn = INFINITY;  // no progress possible: emit everything
      a.tau_new[q] = tn;
    }
    if (a.stats) {
      atomicAdd(a.stats, done);
      if (overflow) atomicAdd(a.stats + 1, 1);
    }
    int o = 0;
    int64_t* oi = a.out_ids + (int64_t)q * a.k;
    float* od = a.out_dist + (int64_t)q * a.k;
    int lim = done < a.k_req ? done : a.k_req;
    for (int i = 0; i < lim && o < a.k; ++i) {
      uint32_t r = (uint32_t)(ex[i] & 0xffffffffu);
      int64_t id = a.perm ? (int64_t)a.perm[r] : (int64_t)r;
      if (a.exclude_self && id == q) continue;
      oi[o] = id + 1;
      od[o] = ord2f((uint32_t)(ex[i] >> 32));
      ++o;
    }
    if (a.out_count) a.out_count[q] = o;
    for (int i = o; i < a.k; ++i) {
      oi[i] = 0;
      od[i] = INFINITY;
    }
  }
}

__global__ __launch_bounds__(256) void exact_topk_kernel(ExactArgs a) {
  const int nq = *a.nq_dev;
  if ((int)blockIdx.x >= nq) return;
  const int q = a.qlist[blockIdx.x];
  __shared__ uint64_t m[EX_BATCH + EX_KMAX];
  const int tid = threadIdx.x, g = tid >> 4, j = tid & 15;
  const int kr = a.k_req < EX_KMAX ? a.k_req : EX_KMAX;
  for (int i = tid; i < EX_BATCH + EX_KMAX; i += 256) m[i] = ~0ull;
  __syncthreads();
  int nr = a.range_cnt ? a.range_cnt[q] : 1;
  for (int ri = 0; ri < nr; ++ri) {
    int rb = 0, re = a.n_rows;
    if (a.range_cnt) {
      int2 rg = a.ranges[(int64_t)q * a.range_stride + ri];
      rb = rg.x;
      re = rg.y;
    }
    for (int base = rb; base < re; base += EX_BATCH) {
      for (int i = g; i < EX_BATCH; i += 16) {
        int r = base + i;
        uint64_t key = ~0ull;
        if (r < re) {
          float dv = exact_metric(a.x32, a.q32, a.xn2, a.qn2, a.d16, a.metric, r, q, j);
          key = pack_key(dv, (uint32_t)r);
        }
        if (j == 0) m[i] = key;
      }
      __syncthreads();
      bitonic_sort(m, EX_BATCH + EX_KMAX);
      // keep the kr smallest at [EX_BATCH, EX_BATCH + kr)
      uint64_t keep = tid < kr ? m[tid] : ~0ull;
      __syncthreads();
      for (int i = tid; i < EX_BATCH + EX_KMAX; i += 256) m[i] = ~0ull;
      __syncthreads();
      if (tid < kr) m[EX_BATCH + tid] = keep;
      __syncthreads();
    }
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
int exact_kmax() { return EX_KMAX; }

void launch_tau_from_sample(const TauArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  hipLaunchKernelGGL(tau_from_sample_kernel, dim3(a.nq), dim3(256), 0, st, a);
}

void launch_select_rerank(const SelectArgs& a, int grid, hipStream_t st) {
  if (grid <= 0) return;
  hipLaunchKernelGGL(select_rerank_kernel, dim3(grid), dim3(256), 0, st, a);
}

void launch_exact_topk(const ExactArgs& a, int grid, hipStream_t st) {
  if (grid <= 0) return;
  hipLaunchKernelGGL(exact_topk_kernel, dim3(grid), dim3(256), 0, st, a);
}

}  // namespace mann
