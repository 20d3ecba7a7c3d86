// Exact search for k beyond the certified selection's capacity (the reference
// accepts any k: bruteforce.jl:46-71, partialsortperm over all distances).
// Every exact FP32 distance of a chunk of queries (the A16 order), packed with
// its row as a (distance, row) key, a segmented radix sort per query, and the
// first k keys (the query itself dropped for a self-join, knn_graph.jl:105-130)
// written out: ties go to the lower row, as partialsortperm's.  Nothing is
// approximated, so no certification is needed; the cost is one full pass and
// one sort per query, the path of large k only.
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "bounds.h"
#include "core.h"

namespace mann {

namespace {

constexpr int ALL_ROWS = 64;  // rows per workgroup (16 groups of 16 lanes)

__global__ __launch_bounds__(256) void all_keys_kernel(const float* __restrict__ x32,
                                                       const float* __restrict__ q32,
                                                       const float* __restrict__ xn2,
                                                       const float* __restrict__ qn2, int d16,
                                                       int metric, int n, int q0,
                                                       uint64_t* __restrict__ keys) {
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  const int qi = blockIdx.y;
  const int r0 = blockIdx.x * ALL_ROWS;
  const int r1 = min(n, r0 + ALL_ROWS);
  for (int r = r0 + g; r < r1; r += 16) {  // (uniform in the 16-lane group)
    const float v = exact_metric(x32, q32, xn2, qn2, d16, metric, r, (int64_t)q0 + qi, j);
    if (j == 0) keys[(int64_t)qi * n + r] = pack_key(v, (uint32_t)r);
  }
}

__global__ void all_offsets_kernel(int* off, int m, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= m) off[i] = i * n;
}

__global__ void all_out_kernel(const uint64_t* __restrict__ sorted, int n, int q0, int qc, int k,
                               int exclude_self, int64_t* __restrict__ out_ids,
                               float* __restrict__ out_dist, int* __restrict__ out_count) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= qc) return;
  const int64_t q = (int64_t)q0 + qi;
  const uint64_t* s = sorted + (int64_t)qi * n;
  int64_t* oi = out_ids + q * k;
  float* od = out_dist + q * k;
  int o = 0;
  for (int i = 0; i < n && o < k; ++i) {
    const uint64_t key = s[i];
    const int64_t r = (int64_t)(key & 0xffffffffu);
    if (exclude_self && r == q) continue;
    oi[o] = r + 1;
    od[o] = ord2f((uint32_t)(key >> 32));
    ++o;
  }
  if (out_count) out_count[q] = o;
  for (; o < k; ++o) {
    oi[o] = 0;
    od[o] = INFINITY;
  }
}

}  // namespace

void bf_search_all(const mann_dataset& ds, const QueryOps& Q, int k, int metric, int exclude_self,
                   int64_t* out_ids, float* out_dist, int* out_count) {
  ThreadCtx& c = ctx();
  MANN_REQUIRE(!c.capture, ERR_ARGUMENT,
               "k = " + std::to_string(k) + " exceeds what a captured search holds (" +
                   std::to_string(select_exmax() / 2) + ")");
  hipStream_t st = c.stream;
  const int n = (int)ds.n;
  const int nq = (int)Q.nq;
  // query chunks of at most 2^26 keys (512 MiB, twice: unsorted and sorted)
  const int qc = (int)std::max<int64_t>(
      1, std::min<int64_t>({(int64_t)nq, 65535, ((int64_t)1 << 26) / std::max(n, 1)}));
  uint64_t* k0 = c.buf("all_k0").as<uint64_t>((size_t)qc * n);
  uint64_t* k1 = c.buf("all_k1").as<uint64_t>((size_t)qc * n);
  int* off = c.buf("all_off").as<int>((size_t)qc + 1);
  size_t need = 0;
  MANN_HIP(rocprim::segmented_radix_sort_keys(nullptr, need, k0, k1, (unsigned)((size_t)qc * n),
                                              (unsigned)qc, off, off + 1, 0u, 64u, st));
  void* tmp = c.buf("all_tmp").get(std::max<size_t>(need, 1));
  ScopedTime tm("exact_all", st);
  for (int q0 = 0; q0 < nq; q0 += qc) {
    const int m = std::min(qc, nq - q0);
    hipLaunchKernelGGL(all_offsets_kernel, dim3((unsigned)ceil_div(m + 1, 256)), dim3(256), 0, st,
                       off, m, n);
    hipLaunchKernelGGL(all_keys_kernel, dim3((unsigned)ceil_div(n, ALL_ROWS), (unsigned)m),
                       dim3(256), 0, st, ds.x32, Q.q32, ds.xn2, Q.qn2, (int)ds.d16, metric, n,
                       q0, k0);
    size_t nb = need;
    MANN_HIP(rocprim::segmented_radix_sort_keys(tmp, nb, k0, k1, (unsigned)((size_t)m * n),
                                                (unsigned)m, off, off + 1, 0u, 64u, st));
    hipLaunchKernelGGL(all_out_kernel, dim3((unsigned)ceil_div(m, 64)), dim3(64), 0, st, k1, n,
                       q0, m, k, exclude_self, out_ids, out_dist, out_count);
  }
  // callers read the outputs (and counts) with blocking copies, as after bf_search
  MANN_HIP(hipStreamSynchronize(st));
  MANN_HIP(hipGetLastError());
  SearchStats& S = last_stats();
  S = SearchStats();
  S.exact = nq;
}

}  // namespace mann
