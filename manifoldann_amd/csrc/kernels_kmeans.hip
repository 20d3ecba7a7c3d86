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
// overflowed is resolved by the exact scan over all centroids.
__global__ __launch_bounds__(256) void top1_select_kernel(Top1Args a) {
  const int j = threadIdx.x & 15;
  const int64_t p = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  if (p >= a.n) return;
  const float* x = a.x32 + p * a.d16;
  const float xn = a.xn2[p];
  const int craw = a.cnt[p];
  float best = INFINITY;
  int bi = 0x7fffffff;
  if (craw <= a.cap) {
    const unsigned* cq = a.cand + p * a.cap * 2;
    int cj = 0;  // candidate i0 + j of this group, fetched 16 at a time
    for (int i = 0; i < craw; ++i) {
      if ((i & 15) == 0) cj = i + j < craw ? (int)cq[2 * (i + j) + 1] : 0;
      const int c = __shfl(cj, (threadIdx.x & 48) + (i & 15));
      float v = norms_form(a.cn2[c], xn, a16_dot(a.c32 + (int64_t)c * a.d16, x, a.d16, j));
      if (v < best || (v == best && c < bi)) {
        best = v;
        bi = c;
      }
    }
  } else {
    for (int c = 0; c < a.k; ++c) {
      float v = norms_form(a.cn2[c], xn, a16_dot(a.c32 + (int64_t)c * a.d16, x, a.d16, j));
      if (v < best) {
        best = v;
        bi = c;
      }
    }
    if (j == 0) atomicAdd(a.overflow, 1);
  }
  if (j == 0) {
    a.assign[p] = bi;
    a.mindist[p] = best;
  }
}

// Emission threshold of the next E-step: the exclusion threshold of the exact
// distance to the centroid the point held before (an upper bound of the new
// minimum), so the scan emits only centroids that can still win.
__global__ __launch_bounds__(256) void prime_threshold_kernel(Top1Args a, const float* qe,
                                                              const float* qr,
                                                              float ex, float rx, int d,
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
    double R32 = R + qe[p] + ex;
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

// exclusive scan over the cluster-major (c, b) table, one block (sequential
// per thread chunk + block scan); also writes cluster offsets
__global__ __launch_bounds__(1024) void scan_table_kernel(int* __restrict__ t, int64_t len, int nb,
                                                          int k, int* __restrict__ offsets) {
  __shared__ long long part[1024];
  const int tid = threadIdx.x;
  const int64_t per = (len + 1023) / 1024;
  const int64_t s = tid * per, e = s + per < len ? s + per : len;
  long long sum = 0;
  for (int64_t i = s; i < e; ++i) sum += t[i];
  part[tid] = sum;
  __syncthreads();
  if (tid == 0) {
    long long run = 0;
    for (int i = 0; i < 1024; ++i) {
      long long v = part[i];
      part[i] = run;
      run += v;
    }
  }
  __syncthreads();
  long long run = part[tid];
  for (int64_t i = s; i < e; ++i) {
    int v = t[i];
    t[i] = (int)run;
    run += v;
  }
  __syncthreads();
  for (int c = tid; c <= k; c += 1024) offsets[c] = c < k ? t[(int64_t)c * nb] : 0;
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
__global__ __launch_bounds__(256) void pp_update_kernel(const float* __restrict__ x32,
                                                        const float* __restrict__ xn2, int64_t n,
                                                        int d16, const float* __restrict__ c,
                                                        double* __restrict__ mins) {
  const int j = threadIdx.x & 15;
  const int64_t p = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  // centre norm in A16 order (recomputed per group: cheaper than a launch)
  float cn = a16_dot(c, c, d16, j);
  if (p >= n) return;
  float dot = a16_dot(c, x32 + p * d16, d16, j);
  if (j == 0) {
    float sq = (cn + xn2[p]) - 2.0f * dot;
    double v = sq > 0.0f ? (double)sq : 0.0;
    if (v < mins[p]) mins[p] = v;
  }
}

constexpr int PP_BLOCK = 1024;

__global__ __launch_bounds__(256) void pp_block_sums_kernel(const double* __restrict__ mins,
                                                            int64_t n, int nb,
                                                            double* __restrict__ bsum) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nb) return;
  int64_t s = (int64_t)b * PP_BLOCK, e = s + PP_BLOCK < n ? s + PP_BLOCK : n;
  double acc = 0.0;
  // ascending order, eight loads in flight (clamped)
  for (int64_t i0 = s; i0 < e; i0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = mins[i0 + u < e ? i0 + u : e - 1];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i0 + u < e) acc += v[u];
  }
  bsum[b] = acc;
}

// single workgroup: block prefix, threshold, position, copy the new centre
__global__ __launch_bounds__(256) void pp_pick_kernel(const double* __restrict__ mins, int64_t n,
                                                      int nb, const double* __restrict__ bsum,
                                                      const double* __restrict__ draws, int round,
                                                      const float* __restrict__ x32, int d16,
                                                      float* __restrict__ centres,
                                                      int* __restrict__ picked) {
  __shared__ int s_idx;
  if (threadIdx.x == 0) {
    // sequential FP64 sums in ascending order, loads eight at a time (clamped)
    double prefix = 0.0, total = 0.0;
    for (int b0 = 0; b0 < nb; b0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = bsum[b0 + u < nb ? b0 + u : nb - 1];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (b0 + u < nb) total += v[u];
    }
    const double thr = draws[round] * total;
    int64_t idx = n - 1;
    int hit = -1;  // first block whose end reaches the threshold
    for (int b0 = 0; b0 < nb && hit < 0; b0 += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = bsum[b0 + u < nb ? b0 + u : nb - 1];
      for (int u = 0; u < 8 && b0 + u < nb; ++u) {
        const double end = prefix + v[u];
        if (end >= thr) {
          hit = b0 + u;
          break;
        }
        prefix = end;
      }
    }
    if (hit >= 0) {
      const int64_t s = (int64_t)hit * PP_BLOCK, e = s + PP_BLOCK < n ? s + PP_BLOCK : n;
      double loc = 0.0;
      bool found = false;
      for (int64_t i0 = s; i0 < e && !found; i0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = mins[i0 + u < e ? i0 + u : e - 1];
        for (int u = 0; u < 8 && i0 + u < e; ++u) {
          loc += v[u];
          if (prefix + loc >= thr) {
            idx = i0 + u;
            found = true;
            break;
          }
        }
      }
    }
    s_idx = (int)idx;
    picked[round] = (int)idx;
  }
  __syncthreads();
  const int idx = s_idx;
  for (int r = threadIdx.x; r < d16; r += 256)
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
  hipLaunchKernelGGL(top1_select_kernel, dim3((unsigned)ceil_div(a.n, 16)), dim3(256), 0, st, a);
}

void launch_prime_threshold(const Top1Args& a, const float* qe, const float* qr, float ex, float rx,
                            int d, float* tau, hipStream_t st) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(prime_threshold_kernel, dim3((unsigned)ceil_div(a.n, 16)), dim3(256), 0, st,
                     a, qe, qr, ex, rx, d, tau);
}

int sort_block() { return SORT_BLOCK; }

void launch_counting_sort(const int* assign, int64_t n, int k, int* table, int* offsets,
                          int* members, hipStream_t st) {
  const int nb = (int)ceil_div(n, SORT_BLOCK);
  hipLaunchKernelGGL(block_hist_kernel, dim3(nb), dim3(256), sizeof(int) * k, st, assign, n, k,
                     table);
  hipLaunchKernelGGL(scan_table_kernel, dim3(1), dim3(1024), 0, st, table, (int64_t)k * nb, nb, k,
                     offsets);
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
  hipLaunchKernelGGL(pp_update_kernel, dim3((unsigned)ceil_div(n, 16)), dim3(256), 0, st, x32, xn2,
                     n, d16, c, mins);
  const int nb = (int)ceil_div(n, PP_BLOCK);
  hipLaunchKernelGGL(pp_block_sums_kernel, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, st,
                     mins, n, nb, bsum);
  hipLaunchKernelGGL(pp_pick_kernel, dim3(1), dim3(256), 0, st, mins, n, nb, bsum, draws, round,
                     x32, d16, centres, picked);
}

}  // namespace mann
