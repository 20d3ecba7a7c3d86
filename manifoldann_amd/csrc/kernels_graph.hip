// Undirected kNN graph on the GPU: _symmetrize_adjacency
// (src/graphs/knn_graph.jl:196-210).  Every directed edge i -> j is emitted in
// both directions as one packed key (src << b | dst), the keys are radix
// sorted (rocPRIM), duplicates are dropped and the survivors written as CSR:
// each row is the union of out- and in-neighbours in ascending id, as the
// reference's sort!(collect(Set)) gives.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include "runtime.h"

namespace mann {

// keys[2t], keys[2t+1] = both directions of edge t = (i, adj[t]); a neighbour id
// outside [1, n] sets *bad and is replaced by a self edge so that the sort
// stays in range (the host raises before any output is used)
__global__ void sym_emit_kernel(const int64_t* adj, int64_t n, int64_t k, int b, uint64_t* keys,
                                int* bad) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  int64_t i = t / k;
  int64_t j = adj[t] - 1;
  if (j < 0 || j >= n) {
    atomicOr(bad, 1);
    j = i;
  }
  keys[2 * t] = ((uint64_t)i << b) | (uint64_t)j;
  keys[2 * t + 1] = ((uint64_t)j << b) | (uint64_t)i;
}

// flag[t] = key t differs from key t-1 (first of its run); deg[src] counts them
__global__ void sym_mark_kernel(const uint64_t* keys, int64_t m, int b, int* flag,
                                unsigned long long* deg) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  uint64_t key = keys[t];
  uint64_t prev = keys[t > 0 ? t - 1 : 0];
  int f = (t == 0) || (prev != key);
  flag[t] = f;
  if (f) atomicAdd(&deg[key >> b], 1ull);
}

// ids[pos[t]] = dst + 1 for every first-of-run key
__global__ void sym_write_kernel(const uint64_t* keys, const int* flag, const int64_t* pos,
                                 int64_t m, int b, int64_t* ids) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  uint64_t key = keys[t];
  int64_t p = pos[t];
  if (flag[t]) ids[p] = (int64_t)(key & ((1ull << b) - 1)) + 1;
}

static inline unsigned grid_of(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void knn_graph_symmetrize_gpu(const int64_t* adj, int64_t n, int64_t k, int64_t* offsets,
                              int64_t* ids, int64_t* total, int64_t* max_degree) {
  ThreadCtx& c = ctx();
  hipStream_t st = c.stream;
  MANN_REQUIRE(n < (int64_t(1) << 31), ERR_ARGUMENT, "graph too large");
  int b = 1;
  while ((int64_t(1) << b) < n) ++b;
  const int64_t m = 2 * n * k;
  if (m == 0) {
    MANN_HIP(hipMemsetAsync(offsets, 0, sizeof(int64_t) * (n + 1), st));
    MANN_HIP(hipStreamSynchronize(st));
    if (total) *total = 0;
    if (max_degree) *max_degree = 0;
    return;
  }
  uint64_t* k0 = c.buf("sym_k0").as<uint64_t>(m);
  uint64_t* k1 = c.buf("sym_k1").as<uint64_t>(m);
  int* flag = c.buf("sym_flag").as<int>(m);
  int64_t* pos = c.buf("sym_pos").as<int64_t>(m);
  unsigned long long* deg = c.buf("sym_deg").as<unsigned long long>(n + 1);
  int* bad = c.buf("sym_bad").as<int>(2);
  unsigned long long* dmax = c.buf("sym_max").as<unsigned long long>(1);
  MANN_HIP(hipMemsetAsync(bad, 0, sizeof(int), st));
  MANN_HIP(hipMemsetAsync(deg, 0, sizeof(unsigned long long) * (n + 1), st));
  {
    ScopedTime tm("sym_emit", st);
    sym_emit_kernel<<<grid_of(n * k, 256), 256, 0, st>>>(adj, n, k, b, k0, bad);
    MANN_HIP(hipGetLastError());
  }
  int hbad = 0;
  MANN_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  MANN_REQUIRE(hbad == 0, ERR_ARGUMENT, "neighbour id out of range");

  size_t need = 0, tmp_bytes = 0;
  MANN_HIP(rocprim::radix_sort_keys(nullptr, need, k0, k1, (size_t)m, 0u, (unsigned)(2 * b), st));
  tmp_bytes = std::max(tmp_bytes, need);
  MANN_HIP(rocprim::exclusive_scan(nullptr, need, flag, pos, (int64_t)0, (size_t)m,
                                   rocprim::plus<int64_t>(), st));
  tmp_bytes = std::max(tmp_bytes, need);
  MANN_HIP(rocprim::exclusive_scan(nullptr, need, deg, (unsigned long long*)offsets, 0ull,
                                   (size_t)(n + 1), rocprim::plus<unsigned long long>(), st));
  tmp_bytes = std::max(tmp_bytes, need);
  MANN_HIP(rocprim::reduce(nullptr, need, deg, dmax, 0ull, (size_t)n,
                           rocprim::maximum<unsigned long long>(), st));
  tmp_bytes = std::max(tmp_bytes, need);
  void* tmp = c.buf("sym_tmp").get(tmp_bytes);

  {
    ScopedTime tm("sym_sort", st);
    need = tmp_bytes;
    MANN_HIP(rocprim::radix_sort_keys(tmp, need, k0, k1, (size_t)m, 0u, (unsigned)(2 * b), st));
  }
  {
    ScopedTime tm("sym_mark", st);
    sym_mark_kernel<<<grid_of(m, 256), 256, 0, st>>>(k1, m, b, flag, deg);
    MANN_HIP(hipGetLastError());
  }
  need = tmp_bytes;
  MANN_HIP(rocprim::exclusive_scan(tmp, need, flag, pos, (int64_t)0, (size_t)m,
                                   rocprim::plus<int64_t>(), st));
  need = tmp_bytes;
  MANN_HIP(rocprim::exclusive_scan(tmp, need, deg, (unsigned long long*)offsets, 0ull,
                                   (size_t)(n + 1), rocprim::plus<unsigned long long>(), st));
  need = tmp_bytes;
  MANN_HIP(rocprim::reduce(tmp, need, deg, dmax, 0ull, (size_t)n,
                           rocprim::maximum<unsigned long long>(), st));
  {
    ScopedTime tm("sym_write", st);
    sym_write_kernel<<<grid_of(m, 256), 256, 0, st>>>(k1, flag, pos, m, b, ids);
    MANN_HIP(hipGetLastError());
  }
  int64_t h[2] = {0, 0};
  MANN_HIP(hipMemcpyAsync(&h[0], offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipMemcpyAsync(&h[1], dmax, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MANN_HIP(hipStreamSynchronize(st));
  if (total) *total = h[0];
  if (max_degree) *max_degree = h[1];
}

}  // namespace mann
