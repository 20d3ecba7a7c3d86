// IVF search kernels: routing distances (transform(::KMeansTransform)),
// TopKRouting probe selection with empty-bucket removal, inversion of the
// probe lists into per-list query lists, the persistent scan's work items, and
// the certified emission threshold from an exact row sample.
#include "bounds.h"
#include "ivf.h"

namespace mann {

namespace {

// R[q][c] = norms form to centroid c (kmeans/distance.jl:213-242), A16 order.
// Block = one query x 16 centroids (one 16-lane group each).
__global__ __launch_bounds__(256) void route_dist_kernel(const float* __restrict__ q32,
                                                         const float* __restrict__ qn2, int nq,
                                                         const float* __restrict__ c32,
                                                         const float* __restrict__ cn2, int nlist,
                                                         int d16, float* __restrict__ R) {
  const int cblocks = (nlist + 15) / 16;
  const int q = blockIdx.x / cblocks;
  const int c = (blockIdx.x - q * cblocks) * 16 + (threadIdx.x >> 4);
  const int j = threadIdx.x & 15;
  if (q >= nq || c >= nlist) return;
  const float* x = c32 + (int64_t)c * d16;
  const float* y = q32 + (int64_t)q * d16;
  float acc = 0.0f;
  for (int t = 0; t < d16; t += 16) acc = __builtin_fmaf(x[t + j], y[t + j], acc);
  acc = group16_tree(acc);
  if (j == 0) {
    float tt = cn2[c] + qn2[q];
    float sq = tt - 2.0f * acc;
    R[(int64_t)q * nlist + c] = __builtin_sqrtf(sq > 0.0f ? sq : 0.0f);
  }
}

__device__ __forceinline__ void bitonic_sort_u64(uint64_t* s, int P) {
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint64_t a = s[i], b = s[ixj];
          if ((a > b) == ((i & k) == 0)) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
}

constexpr int ROUTE_MAX = 4096;

// select_indices(::TopKRouting) (routing.jl:94-118): the min(nprobe, avail)
// smallest routing distances among the first avail = min(n_buckets, nlist),
// ties to the lower bucket; then _resolve_child_positions (query.jl:108-121)
// drops buckets without a child.  probes[q][r] in probe order, -1 padded.
__global__ __launch_bounds__(256) void probe_select_kernel(const float* __restrict__ R, int nq,
                                                           int nlist, int avail, int nprobe,
                                                           const int* __restrict__ list_off,
                                                           int* __restrict__ probes,
                                                           int* __restrict__ pcount,
                                                           int* __restrict__ prows) {
  __shared__ uint64_t keys[ROUTE_MAX];
  const int q = blockIdx.x;
  if (q >= nq) return;
  int P = 1;
  while (P < avail) P <<= 1;
  for (int i = threadIdx.x; i < P; i += 256)
    keys[i] = i < avail ? pack_key(R[(int64_t)q * nlist + i], (uint32_t)i) : ~0ull;
  __syncthreads();
  bitonic_sort_u64(keys, P);
  if (threadIdx.x == 0) {
    const int kk = nprobe < avail ? nprobe : avail;
    int m = 0, rows = 0;
    for (int r = 0; r < kk; ++r) {
      const int b = (int)(keys[r] & 0xffffffffu);
      const int sz = list_off[b + 1] - list_off[b];
      if (sz == 0) continue;
      probes[(int64_t)q * nprobe + m++] = b;
      rows += sz;
    }
    for (int r = m; r < nprobe; ++r) probes[(int64_t)q * nprobe + r] = -1;
    pcount[q] = m;
    prows[q] = rows;
  }
}

// per-list count of probing queries
__global__ void list_count_kernel(const int* __restrict__ probes, const int* __restrict__ pcount,
                                  int nq, int nprobe, int* __restrict__ qcount) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nq * nprobe) return;
  const int q = (int)(i / nprobe), r = (int)(i - (int64_t)q * nprobe);
  if (r < pcount[q]) atomicAdd(&qcount[probes[i]], 1);
}

// one block: exclusive scan of qcount (list order) and of the work-item counts
// (size-descending list order); writes n_items.
__global__ __launch_bounds__(1024) void plan_kernel(const int* __restrict__ qcount, int nb,
                                                    const int* __restrict__ list_off,
                                                    const int* __restrict__ lorder, int bn,
                                                    int rows_item, int* __restrict__ qoff,
                                                    int* __restrict__ ioff,
                                                    int* __restrict__ n_items) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  // queries offsets in bucket order
  int s0 = tid * per, e0 = min(nb, s0 + per), sum = 0;
  for (int i = s0; i < e0; ++i) sum += qcount[i];
  part[tid] = sum;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int i = 0; i < 1024; ++i) {
      int v = part[i];
      part[i] = run;
      run += v;
    }
  }
  __syncthreads();
  int run = part[tid];
  for (int i = s0; i < e0; ++i) {
    qoff[i] = run;
    run += qcount[i];
  }
  __syncthreads();
  // items in size-descending list order
  sum = 0;
  for (int i = s0; i < e0; ++i) {
    const int b = lorder[i];
    const int nqc = (qcount[b] + bn - 1) / bn;
    const int sz = list_off[b + 1] - list_off[b];
    const int ns = (sz + rows_item - 1) / rows_item;
    sum += nqc * ns;
  }
  part[tid] = sum;
  __syncthreads();
  if (tid == 0) {
    int r2 = 0;
    for (int i = 0; i < 1024; ++i) {
      int v = part[i];
      part[i] = r2;
      r2 += v;
    }
    *n_items = r2;
  }
  __syncthreads();
  run = part[tid];
  for (int i = s0; i < e0; ++i) {
    ioff[i] = run;
    const int b = lorder[i];
    const int nqc = (qcount[b] + bn - 1) / bn;
    const int sz = list_off[b + 1] - list_off[b];
    run += nqc * ((sz + rows_item - 1) / rows_item);
  }
}

__global__ void list_fill_kernel(const int* __restrict__ probes, const int* __restrict__ pcount,
                                 int nq, int nprobe, const int* __restrict__ qoff,
                                 int* __restrict__ qfill, int* __restrict__ qlist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nq * nprobe) return;
  const int q = (int)(i / nprobe), r = (int)(i - (int64_t)q * nprobe);
  if (r >= pcount[q]) return;
  const int b = probes[i];
  const int pos = atomicAdd(&qfill[b], 1);
  qlist[qoff[b] + pos] = q;
}

__global__ void items_fill_kernel(const int* __restrict__ qcount, int nb,
                                  const int* __restrict__ list_off, const int* __restrict__ lorder,
                                  const int* __restrict__ qoff, const int* __restrict__ ioff,
                                  int bn, int rows_item, ScanItem* __restrict__ items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int b = lorder[i];
  const int nqc = (qcount[b] + bn - 1) / bn;
  const int lo = list_off[b], hi = list_off[b + 1];
  const int ns = (hi - lo + rows_item - 1) / rows_item;
  int w = ioff[i];
  for (int s = 0; s < ns; ++s)
    for (int c = 0; c < nqc; ++c) {
      ScanItem it;
      it.qoff = qoff[b] + c * bn;
      it.nq = min(bn, qcount[b] - c * bn);
      it.row_begin = lo + s * rows_item;
      it.row_end = min(hi, it.row_begin + rows_item);
      it.region = 0;
      items[w++] = it;
    }
}

constexpr int TAU_M = 2048;

// Emission threshold of each query from the exact distances of the first
// TAU_M rows of its probed lists (probe order): the k-th smallest exact value
// is an upper bound of the final k-th, so T* of it certifies the emitted set.
__global__ __launch_bounds__(256) void ivf_tau_kernel(IvfTauArgs a) {
  __shared__ uint64_t keys[TAU_M];
  const int q = blockIdx.x;
  if (q >= a.nq) return;
  const int tid = threadIdx.x, g = tid >> 4, j = tid & 15;
  const int np = a.pcount[q];
  const int total = a.prows[q];
  const int m = total < TAU_M ? total : TAU_M;
  const int k_eff = a.k < total ? a.k : total;
  if (tid == 0) a.k_eff_q[q] = k_eff;
  if (m <= 0 || k_eff <= 0 || total <= a.emit_all_below) {
    if (tid == 0) a.tau[q] = INFINITY;
    return;
  }
  for (int i = g; i < TAU_M; i += 16) {
    uint64_t key = ~0ull;
    if (i < m) {
      // locate sample position i in the probe-order concatenation
      int r = 0, base = 0, row = -1;
      for (; r < np; ++r) {
        const int b = a.probes[(int64_t)q * a.nprobe + r];
        const int sz = a.list_off[b + 1] - a.list_off[b];
        if (i < base + sz) {
          row = a.list_off[b] + (i - base);
          break;
        }
        base += sz;
      }
      float dv = exact_metric(a.x32, a.q32, a.xn2, a.qn2, a.d16, a.metric, row, q, j);
      key = pack_key(dv, (uint32_t)row);
    }
    if (j == 0) keys[i] = key;
  }
  __syncthreads();
  int P = 1;
  while (P < m) P <<= 1;
  bitonic_sort_u64(keys, P);
  if (tid == 0) {
    float dk = ord2f((uint32_t)(keys[k_eff - 1] >> 32));
    double T = exclusion_threshold(dk, a.metric, a.qn[q], a.qe[q], a.qr[q], a.ex, a.rx, a.d);
    a.tau[q] = T < 3.0e38 ? nextafterf((float)T * (1.0f + 1e-6f), INFINITY) : INFINITY;
  }
}

// probed row ranges per query (for the exact fallback)
__global__ void probe_ranges_kernel(const int* __restrict__ probes, const int* __restrict__ pcount,
                                    int nq, int nprobe, const int* __restrict__ list_off,
                                    int2* __restrict__ ranges) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nq * nprobe) return;
  const int q = (int)(i / nprobe), r = (int)(i - (int64_t)q * nprobe);
  int2 v = make_int2(0, 0);
  if (r < pcount[q]) {
    const int b = probes[i];
    v = make_int2(list_off[b], list_off[b + 1]);
  }
  ranges[i] = v;
}

}  // namespace

int route_max() { return ROUTE_MAX; }

void launch_route_dist(const float* q32, const float* qn2, int nq, const float* c32,
                       const float* cn2, int nlist, int d16, float* R, hipStream_t st) {
  if (nq <= 0) return;
  const int cblocks = (nlist + 15) / 16;
  hipLaunchKernelGGL(route_dist_kernel, dim3((unsigned)((int64_t)nq * cblocks)), dim3(256), 0, st,
                     q32, qn2, nq, c32, cn2, nlist, d16, R);
}

void launch_probe_select(const float* R, int nq, int nlist, int avail, int nprobe,
                         const int* list_off, int* probes, int* pcount, int* prows,
                         hipStream_t st) {
  if (nq <= 0) return;
  hipLaunchKernelGGL(probe_select_kernel, dim3(nq), dim3(256), 0, st, R, nq, nlist, avail, nprobe,
                     list_off, probes, pcount, prows);
}

void launch_ivf_plan(const int* probes, const int* pcount, int nq, int nprobe, int nb,
                     const int* list_off, const int* lorder, int bn, int rows_item, int* qcount,
                     int* qfill, int* qoff, int* ioff, int* qlist, ScanItem* items, int* n_items,
                     hipStream_t st) {
  const int64_t tot = (int64_t)nq * nprobe;
  const unsigned g = (unsigned)ceil_div(tot, 256);
  hipLaunchKernelGGL(list_count_kernel, dim3(g), dim3(256), 0, st, probes, pcount, nq, nprobe,
                     qcount);
  hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(1024), 0, st, qcount, nb, list_off, lorder, bn,
                     rows_item, qoff, ioff, n_items);
  hipLaunchKernelGGL(list_fill_kernel, dim3(g), dim3(256), 0, st, probes, pcount, nq, nprobe, qoff,
                     qfill, qlist);
  hipLaunchKernelGGL(items_fill_kernel, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, st, qcount,
                     nb, list_off, lorder, qoff, ioff, bn, rows_item, items);
}

void launch_ivf_tau(const IvfTauArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  hipLaunchKernelGGL(ivf_tau_kernel, dim3(a.nq), dim3(256), 0, st, a);
}

void launch_probe_ranges(const int* probes, const int* pcount, int nq, int nprobe,
                         const int* list_off, int2* ranges, hipStream_t st) {
  const int64_t tot = (int64_t)nq * nprobe;
  if (tot <= 0) return;
  hipLaunchKernelGGL(probe_ranges_kernel, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st,
                     probes, pcount, nq, nprobe, list_off, ranges);
}

}  // namespace mann
