// k-means kernels: top-1 certified selection (E-step), stable counting sort of
// assignments, FP64 segmented centroid sums (M-step), movement, farthest point,
// and the kmeans++ D^2-sampling rounds.  Every FP32/FP64 operation follows the
// order of the oracle's restatement of kmeans/{distance,lloyd,init}.jl.
#include "kmeans.h"

namespace mann {

namespace {

// A16 dot product of two rows held in global memory (16-lane group).
__device__ __forceinline__ float a16_dot(const float* __restrict__ x, const float* __restrict__ y,
                                         int d16, int j) {
  float acc = 0.0f;
  int t = 0;
  // eight steps per batch, their loads issued first (the chain keeps its order)
  for (; t + 128 <= d16; t += 128) {
    float xv[8], yv[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      xv[s] = x[t + 16 * s + j];
      yv[s] = y[t + 16 * s + j];
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) acc = __builtin_fmaf(xv[s], yv[s], acc);
  }
  for (; t < d16; t += 16) acc = __builtin_fmaf(x[t + j], y[t + j], acc);
  return group16_tree(acc);
}

// norms form, kmeans/distance.jl:99-101
__device__ __forceinline__ float norms_form(float cn, float xn, float dot) {
  float t = cn + xn;
  float sq = t - 2.0f * dot;
  return __builtin_sqrtf(sq > 0.0f ? sq : 0.0f);
}

// E-step selection, one 16-lane group per point: exact norms-form distance of
// every emitted centroid, argmin with ties to the lower centroid (strict '<'
// of assign_clusters!, distance.jl:299).  A point whose candidate buffer
// overflowed is resolved by the exact scan over all centroids.  The point's
// row stays in registers (lane j holds elements j, j + 16, ...) and centroids
// are taken four at a time with all their loads issued first; each dot
// product keeps the A16 order (one FMA chain per lane in ascending element,
// then the fixed tree).  NS bounds d16 / 16 (registers of the row).
template <int NS>
__global__ __launch_bounds__(256) void top1_select_kernel(Top1Args a) {
  const int j = threadIdx.x & 15;
  const int64_t p = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  if (p >= a.n) return;
  const int ns = a.d16 >> 4;
  const float* x = a.x32 + p * a.d16;
  float xv[NS];
#pragma unroll
  for (int s2 = 0; s2 < NS; ++s2) xv[s2] = s2 < ns ? x[16 * s2 + j] : 0.0f;
  const float xn = a.xn2[p];
  const int craw = a.cnt[p];
  const bool exact = craw > a.cap;
  const int total = exact ? a.k : craw;
  const unsigned* cq = a.cand + p * a.cap * 2;
  float best = INFINITY;
  int bi = 0x7fffffff;
  for (int i0 = 0; i0 < total; i0 += 4) {
    int cid[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = min(i0 + u, total - 1);
      cid[u] = exact ? i : (int)cq[2 * i + 1];
    }
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float cv[4][NS];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float* cr = a.c32 + (int64_t)cid[u] * a.d16 + j;
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) cv[u][s2] = s2 < ns ? cr[16 * s2] : 0.0f;
    }
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2)
      if (s2 < ns)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = __builtin_fmaf(cv[u][s2], xv[s2], acc[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (i0 + u >= total) break;  // (uniform in the group)
      const float v = norms_form(a.cn2[cid[u]], xn, group16_tree(acc[u]));
      if (v < best || (v == best && cid[u] < bi)) {
        best = v;
        bi = cid[u];
      }
    }
  }
  if (j == 0) {
    if (exact) atomicAdd(a.overflow, 1);
    a.assign[p] = bi;
    a.mindist[p] = best;
  }
}

// Rows longer than 64 x 16 elements: the same selection with the row read
// from memory for every centroid.
__global__ __launch_bounds__(256) void top1_select_long_kernel(Top1Args a) {
  const int j = threadIdx.x & 15;
  const int64_t p = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  if (p >= a.n) return;
  const float* x = a.x32 + p * a.d16;
  const float xn = a.xn2[p];
  const int craw = a.cnt[p];
  const bool exact = craw > a.cap;
  const int total = exact ? a.k : craw;
  const unsigned* cq = a.cand + p * a.cap * 2;
  float best = INFINITY;
  int bi = 0x7fffffff;
  for (int i = 0; i < total; ++i) {
    const int c = exact ? i : (int)cq[2 * i + 1];
    const float v = norms_form(a.cn2[c], xn, a16_dot(a.c32 + (int64_t)c * a.d16, x, a.d16, j));
    if (v < best || (v == best && c < bi)) {
      best = v;
      bi = c;
    }
  }
  if (j == 0) {
    if (exact) atomicAdd(a.overflow, 1);
    a.assign[p] = bi;
    a.mindist[p] = best;
  }
}

// Emission threshold of the next E-step: the exclusion threshold of the exact
// distance to the centroid the point held before (an upper bound of the new
// minimum), so the scan emits only centroids that can still win.
__global__ __launch_bounds__(256) void prime_threshold_kernel(Top1Args a, const float* qe,
                                                              const float* qr,
                                                              float ex, float rx, float ra, int d,
                                                              float* tau) {
  const int j = threadIdx.x & 15;
  const int64_t p = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  if (p >= a.n) return;
  const int c = a.assign[p];
  float v = norms_form(a.cn2[c], a.xn2[p],
                       a16_dot(a.c32 + (int64_t)c * a.d16, a.x32 + p * a.d16, a.d16, j));
  if (j == 0) {
    const double u = 1.0 / 16777216.0;
    double R = (double)qr[p] + rx;
    double G = (2.0 * d + 8.0) * u * R * R;
    // FP32 norms-form error: uncentred norms (the bf16 operands may be centred)
    double qa = sqrt((double)a.xn2[p] * (1.0 + (d + 8.0) * u)) * (1.0 + 4.0 * u);
    double R32 = qa + ra;
    double H = (d + 8.0) * u * R32 * R32;
    double f = (double)v / (1.0 - 8.0 * u);
    double L = sqrt(f * f + H) + qe[p] + ex;
    double t = (L * L + G) * (1.0 + 1e-9) + 1e-30;
    tau[p] = nextafterf((float)t * (1.0f + 1e-6f), INFINITY);
  }
}

// ---- stable counting sort of assignments (ascending point id per cluster) ----
constexpr int SORT_BLOCK = 1024;

__global__ __launch_bounds__(256) void block_hist_kernel(const int* __restrict__ assign, int64_t n,
                                                         int k, int* __restrict__ bcount) {
  // bcount[c * nb + b]: points of block b in cluster c
  extern __shared__ int h[];
  const int b = blockIdx.x, nb = gridDim.x;
  for (int c = threadIdx.x; c < k; c += 256) h[c] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < SORT_BLOCK; i += 256) {
    int64_t p = (int64_t)b * SORT_BLOCK + i;
    if (p < n) atomicAdd(&h[assign[p]], 1);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += 256) bcount[(int64_t)c * nb + b] = h[c];
}

// Exclusive scan of the cluster-major (c, b) table in three passes: tile sums
// (SCAN_TILE elements per block), an exclusive scan of the tile sums in one
// block, then each tile's own scan plus its offset.  Integer sums: exact in any
// order.  Also writes the cluster offsets offsets[c] = scanned t[c * nb]
// (offsets[k] = 0; callers set the end themselves).
constexpr int SCAN_TILE = 4096;  // 256 threads x 16 elements

__global__ __launch_bounds__(256) void scan_tile_sums_kernel(const int* __restrict__ t,
                                                             int64_t len,
                                                             long long* __restrict__ tsum) {
  __shared__ long long part[256];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * 16;
  long long s = 0;
#pragma unroll
  for (int u = 0; u < 16; ++u) s += base + u < len ? t[base + u] : 0;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tsum[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(1024) void scan_tile_prefix_kernel(long long* __restrict__ tsum,
                                                                int ntiles) {
  __shared__ long long part[1024];
  const int tid = threadIdx.x;
  const int per = (ntiles + 1023) / 1024;
  const int s0 = tid * per, s1 = min(ntiles, s0 + per);
  long long sum = 0;
  for (int i = s0; i < s1; ++i) sum += tsum[i];
  part[tid] = sum;
  __syncthreads();
  for (int w = 1; w < 1024; w <<= 1) {  // inclusive Hillis-Steele scan
    const long long v = tid >= w ? part[tid - w] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  long long run = part[tid] - sum;  // exclusive
  for (int i = s0; i < s1; ++i) {
    const long long v = tsum[i];
    tsum[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(256) void scan_tile_apply_kernel(int* __restrict__ t, int64_t len,
                                                              const long long* __restrict__ tsum,
                                                              int nb, int k,
                                                              int* __restrict__ offsets) {
  __shared__ long long part[256];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * 16;
  int v[16];
  long long s = 0;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    v[u] = base + u < len ? t[base + u] : 0;
    s += v[u];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 1; w < 256; w <<= 1) {
    const long long x = threadIdx.x >= (unsigned)w ? part[threadIdx.x - w] : 0;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  long long run = tsum[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int64_t i = base + u;
    if (i < len) {
      t[i] = (int)run;
      if (i % nb == 0) offsets[i / nb] = (int)run;
    }
    run += v[u];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) offsets[k] = 0;
}

__global__ __launch_bounds__(256) void scatter_kernel(const int* __restrict__ assign, int64_t n,
                                                      int k, const int* __restrict__ start,
                                                      int* __restrict__ members) {
  extern __shared__ int sh[];  // [k] running counters, then [256] chunk clusters
  int* run = sh;
  int* cl = sh + k;
  const int b = blockIdx.x, nb = gridDim.x;
  for (int c = threadIdx.x; c < k; c += 256) run[c] = start[(int64_t)c * nb + b];
  __syncthreads();
  for (int base = 0; base < SORT_BLOCK; base += 256) {
    int64_t p = (int64_t)b * SORT_BLOCK + base + threadIdx.x;
    int c = p < n ? assign[p] : -1;
    cl[threadIdx.x] = c;
    __syncthreads();
    if (c >= 0) {
      int rank = 0;
      for (int t = 0; t < (int)threadIdx.x; ++t) rank += cl[t] == c;
      members[run[c] + rank] = (int)p;
    }
    __syncthreads();
    if (c >= 0) {
      // the last thread of each cluster in this chunk advances its counter
      bool last = true;
      for (int t = threadIdx.x + 1; t < 256; ++t)
        if (cl[t] == c) {
          last = false;
          break;
        }
      if (last) {
        int cntc = 0;
        for (int t = 0; t <= (int)threadIdx.x; ++t) cntc += cl[t] == c;
        run[c] += cntc;
      }
    }
    __syncthreads();
  }
}

// ---- FP64 segmented sums: chunk partials, then per-cluster chunk order ----
__global__ __launch_bounds__(256) void chunk_sums_kernel(const float* __restrict__ x32, int d,
                                                         int d16, const int* __restrict__ members,
                                                         const int2* __restrict__ chunks,
                                                         int nchunks, double* __restrict__ part) {
  const int ci = blockIdx.x;
  if (ci >= nchunks) return;
  const int2 ch = chunks[ci];  // member range [x, y)
  for (int r = threadIdx.x; r < d; r += 256) {
    double s = 0.0;
    for (int m = ch.x; m < ch.y; ++m) s += (double)x32[(int64_t)members[m] * d16 + r];
    part[(int64_t)ci * d + r] = s;
  }
}

__global__ __launch_bounds__(256) void cluster_means_kernel(const double* __restrict__ part,
                                                            const int* __restrict__ chunk_off,
                                                            const int* __restrict__ counts, int k,
                                                            int d, int d16,
                                                            float* __restrict__ newc) {
  const int c = blockIdx.x;
  if (c >= k) return;
  const int cnt = counts[c];
  for (int r = threadIdx.x; r < d16; r += 256) {
    if (r >= d) {
      newc[(int64_t)c * d16 + r] = 0.0f;
      continue;
    }
    double s = 0.0;
    for (int t = chunk_off[c]; t < chunk_off[c + 1]; ++t) s += part[(int64_t)t * d + r];
    if (cnt > 0) newc[(int64_t)c * d16 + r] = (float)(s / (double)cnt);
  }
}

// movement[c] = Euclidean(old_c, new_c), difference form (lloyd.jl:113-118)
__global__ __launch_bounds__(256) void movement_kernel(const float* __restrict__ oldc,
                                                       const float* __restrict__ newc, int k,
                                                       int d16, float* __restrict__ mv) {
  const int j = threadIdx.x & 15;
  const int c = blockIdx.x * 16 + (threadIdx.x >> 4);
  if (c >= k) return;
  float acc = 0.0f;
  for (int t = 0; t < d16; t += 16) {
    float df = oldc[(int64_t)c * d16 + t + j] - newc[(int64_t)c * d16 + t + j];
    acc = __builtin_fmaf(df, df, acc);
  }
  acc = group16_tree(acc);
  if (j == 0) mv[c] = __builtin_sqrtf(acc);
}

// farthest point: max mindist, first index on ties (lloyd.jl:88-96)
__global__ __launch_bounds__(1024) void farthest_kernel(const float* __restrict__ md, int64_t n,
                                                        unsigned long long* __restrict__ out) {
  __shared__ unsigned long long best[1024];
  unsigned long long b = 0;  // key = (ord(dist) << 32) | ~index : max key = max dist, min index
  for (int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 1024) {
    unsigned long long key = ((unsigned long long)f2ord(md[i]) << 32) |
                             (unsigned long long)(0xffffffffu - (unsigned)i);
    b = key > b ? key : b;
  }
  best[threadIdx.x] = b;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s && best[threadIdx.x + s] > best[threadIdx.x])
      best[threadIdx.x] = best[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, best[0]);
}

// ---- kmeans++ (kmeans/init.jl:56-128) ----
constexpr int PP_BLOCK = 1024;

// One workgroup per PP_BLOCK points: the running minimum squared distance to
// the chosen centres (_update_distances_euclidean!, init.jl:116-127: FP32
// norms form of the newest centre, clamped at 0, widened to FP64), then the
// block's sequential FP64 sum in ascending point order (the blocked
// cumulative sum's block total).
__global__ __launch_bounds__(256) void pp_update_kernel(const float* __restrict__ x32,
                                                        const float* __restrict__ xn2, int64_t n,
                                                        int d16, const float* __restrict__ c,
                                                        double* __restrict__ mins,
                                                        double* __restrict__ bsum) {
  __shared__ double m[PP_BLOCK];
  const int j = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int64_t b0 = (int64_t)blockIdx.x * PP_BLOCK;
  // centre norm in A16 order (recomputed per group: cheaper than a launch)
  const float cn = a16_dot(c, c, d16, j);
  for (int i = g; i < PP_BLOCK; i += 16) {
    const int64_t p = b0 + i;
    double v = 0.0;
    if (p < n) {
      const float dot = a16_dot(c, x32 + p * d16, d16, j);
      const float sq = (cn + xn2[p]) - 2.0f * dot;
      const double cur = mins[p];
      v = sq > 0.0f ? (double)sq : 0.0;
      v = v < cur ? v : cur;
      if (j == 0) mins[p] = v;
    }
    if (j == 0) m[i] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int e = (int)(n - b0 < PP_BLOCK ? n - b0 : PP_BLOCK);
    double acc = 0.0;
    for (int i = 0; i < e; ++i) acc += m[i];
    bsum[blockIdx.x] = acc;
  }
}

// One workgroup: total = sequential sum of the block totals; the first block
// whose running end reaches u * total; inside it the first point whose
// cumulative sum reaches it (searchsortedfirst, init.jl:92-96, clamped to the
// last point); then the new centre is copied.  The inputs are staged in LDS
// first, so the sequential FP64 chains run on LDS reads.
constexpr int PP_PICK_LDS = 8192;  // block totals staged at a time
__global__ __launch_bounds__(1024) void pp_pick_kernel(const double* __restrict__ mins,
                                                       int64_t n, int nb,
                                                       const double* __restrict__ bsum,
                                                       const double* __restrict__ draws,
                                                       int round, const float* __restrict__ x32,
                                                       int d16, float* __restrict__ centres,
                                                       int* __restrict__ picked) {
  __shared__ double sb[PP_PICK_LDS];
  __shared__ double s_total, s_prefix;
  __shared__ int s_hit, s_idx;
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_total = 0.0;
    s_prefix = 0.0;
    s_hit = -1;
  }
  // total, in ascending block order, one staged slice at a time
  for (int c0 = 0; c0 < nb; c0 += PP_PICK_LDS) {
    const int cn = min(PP_PICK_LDS, nb - c0);
    __syncthreads();
    for (int i = tid; i < cn; i += 1024) sb[i] = bsum[c0 + i];
    __syncthreads();
    if (tid == 0) {
      double t = s_total;
      for (int i = 0; i < cn; ++i) t += sb[i];
      s_total = t;
    }
  }
  __syncthreads();
  const double thr = draws[round] * s_total;
  for (int c0 = 0; c0 < nb; c0 += PP_PICK_LDS) {
    const int cn = min(PP_PICK_LDS, nb - c0);
    __syncthreads();
    if (s_hit >= 0) break;
    for (int i = tid; i < cn; i += 1024) sb[i] = bsum[c0 + i];
    __syncthreads();
    if (tid == 0) {
      double prefix = s_prefix;
      for (int i = 0; i < cn; ++i) {
        const double end = prefix + sb[i];
        if (end >= thr) {
          s_hit = c0 + i;
          break;
        }
        prefix = end;
      }
      s_prefix = prefix;
    }
  }
  __syncthreads();
  const int hit = s_hit;
  if (hit >= 0) {
    const int64_t s0 = (int64_t)hit * PP_BLOCK;
    const int e = (int)(n - s0 < PP_BLOCK ? n - s0 : PP_BLOCK);
    for (int i = tid; i < e; i += 1024) sb[i] = mins[s0 + i];
    __syncthreads();
    if (tid == 0) {
      const double prefix = s_prefix;
      double loc = 0.0;
      int64_t idx = n - 1;
      for (int i = 0; i < e; ++i) {
        loc += sb[i];
        if (prefix + loc >= thr) {
          idx = s0 + i;
          break;
        }
      }
      s_idx = (int)idx;
    }
  } else if (tid == 0) {
    s_idx = (int)(n - 1);
  }
  __syncthreads();
  const int idx = s_idx;
  if (tid == 0) picked[round] = idx;
  for (int r = tid; r < d16; r += 1024)
    centres[(int64_t)round * d16 + r] = x32[(int64_t)idx * d16 + r];
}

__global__ void pp_first_kernel(const double* __restrict__ draws, int64_t n,
                                const float* __restrict__ x32, int d16, float* __restrict__ centres,
                                int* __restrict__ picked, double* __restrict__ mins) {
  __shared__ int s_idx;
  if (threadIdx.x == 0) {
    int64_t f = (int64_t)(draws[0] * (double)n);
    if (f >= n) f = n - 1;
    s_idx = (int)f;
    picked[0] = (int)f;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < d16; r += blockDim.x) centres[r] = x32[(int64_t)s_idx * d16 + r];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += blockDim.x)
    mins[i] = INFINITY;
}

}  // namespace

void launch_top1_select(const Top1Args& a, hipStream_t st) {
  if (a.n <= 0) return;
  const dim3 grid((unsigned)ceil_div(a.n, 16)), block(256);
  const int ns = a.d16 / 16;
  if (ns <= 16) hipLaunchKernelGGL(top1_select_kernel<16>, grid, block, 0, st, a);
  else if (ns <= 32) hipLaunchKernelGGL(top1_select_kernel<32>, grid, block, 0, st, a);
  else if (ns <= 64) hipLaunchKernelGGL(top1_select_kernel<64>, grid, block, 0, st, a);
  else hipLaunchKernelGGL(top1_select_long_kernel, grid, block, 0, st, a);
}

void launch_prime_threshold(const Top1Args& a, const float* qe, const float* qr, float ex, float rx,
                            float ra, int d, float* tau, hipStream_t st) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(prime_threshold_kernel, dim3((unsigned)ceil_div(a.n, 16)), dim3(256), 0, st,
                     a, qe, qr, ex, rx, ra, d, tau);
}

int sort_block() { return SORT_BLOCK; }

int64_t counting_sort_table_ints(int64_t n, int k) {
  const int64_t len = (int64_t)k * ceil_div(n, SORT_BLOCK);
  return round_up(len, 2) + 2 * ceil_div(len, SCAN_TILE);
}

void launch_counting_sort(const int* assign, int64_t n, int k, int* table, int* offsets,
                          int* members, hipStream_t st) {
  const int nb = (int)ceil_div(n, SORT_BLOCK);
  hipLaunchKernelGGL(block_hist_kernel, dim3(nb), dim3(256), sizeof(int) * k, st, assign, n, k,
                     table);
  const int64_t len = (int64_t)k * nb;
  const int ntiles = (int)ceil_div(len, SCAN_TILE);
  // the tile sums follow the table (see counting_sort_table_ints)
  long long* tsum = (long long*)(table + round_up(len, 2));
  hipLaunchKernelGGL(scan_tile_sums_kernel, dim3(ntiles), dim3(256), 0, st, table, len, tsum);
  hipLaunchKernelGGL(scan_tile_prefix_kernel, dim3(1), dim3(1024), 0, st, tsum, ntiles);
  hipLaunchKernelGGL(scan_tile_apply_kernel, dim3(ntiles), dim3(256), 0, st, table, len, tsum, nb,
                     k, offsets);
  hipLaunchKernelGGL(scatter_kernel, dim3(nb), dim3(256), sizeof(int) * (k + 256), st, assign, n,
                     k, table, members);
}

void launch_chunk_sums(const float* x32, int d, int d16, const int* members, const int2* chunks,
                       int nchunks, double* part, hipStream_t st) {
  if (nchunks <= 0) return;
  hipLaunchKernelGGL(chunk_sums_kernel, dim3(nchunks), dim3(256), 0, st, x32, d, d16, members,
                     chunks, nchunks, part);
}

void launch_cluster_means(const double* part, const int* chunk_off, const int* counts, int k, int d,
                          int d16, float* newc, hipStream_t st) {
  hipLaunchKernelGGL(cluster_means_kernel, dim3(k), dim3(256), 0, st, part, chunk_off, counts, k, d,
                     d16, newc);
}

void launch_movement(const float* oldc, const float* newc, int k, int d16, float* mv,
                     hipStream_t st) {
  hipLaunchKernelGGL(movement_kernel, dim3((unsigned)ceil_div(k, 16)), dim3(256), 0, st, oldc, newc,
                     k, d16, mv);
}

void launch_farthest(const float* md, int64_t n, unsigned long long* out, hipStream_t st) {
  int grid = (int)std::min<int64_t>(ceil_div(n, 1024), 1024);
  hipLaunchKernelGGL(farthest_kernel, dim3(grid), dim3(1024), 0, st, md, n, out);
}

void launch_pp_first(const double* draws, int64_t n, const float* x32, int d16, float* centres,
                     int* picked, double* mins, hipStream_t st) {
  hipLaunchKernelGGL(pp_first_kernel, dim3(1), dim3(1024), 0, st, draws, n, x32, d16, centres,
                     picked, mins);
}

void launch_pp_round(const float* x32, const float* xn2, int64_t n, int d16, float* centres,
                     int round, double* mins, double* bsum, const double* draws, int* picked,
                     hipStream_t st) {
  const float* c = centres + (int64_t)(round - 1) * d16;
  const int nb = (int)ceil_div(n, PP_BLOCK);
  hipLaunchKernelGGL(pp_update_kernel, dim3((unsigned)nb), dim3(256), 0, st, x32, xn2, n, d16, c,
                     mins, bsum);
  hipLaunchKernelGGL(pp_pick_kernel, dim3(1), dim3(1024), 0, st, mins, n, nb, bsum, draws, round,
                     x32, d16, centres, picked);
}

}  // namespace mann
