// IVF search kernels: routing distances (transform(::KMeansTransform)),
// TopKRouting probe selection with empty-bucket removal, inversion of the
// probe lists into per-list query lists, the persistent scan's work items, and
// the certified emission threshold from an exact row sample.
#include "bounds.h"
#include "radix.h"
#include "ivf.h"
#include "util.h"

namespace mann {

namespace {

// R[q][c] = norms form to centroid c (kmeans/distance.jl:213-242), A16 order:
// lane j of a 16-lane group runs the FMA chain of the dims t = j (mod 16) in
// ascending t, then the fixed tree.  A block computes a 32 x 32 (query,
// centroid) tile from LDS; group g owns 8 queries x 8 centroids (centroids
// g%4 + 4b, so the four groups of a wavefront read four adjacent LDS rows), so
// each lane holds 64 chains and loads 16 operands per 64 FMAs.  The d axis is
// staged in chunks of RT_DC floats (the chains continue across chunks, order
// unchanged).
constexpr int RT_T = 32;
constexpr int RT_DC = 64;
constexpr int RT_LD = RT_DC + 16;  // row stride: adjacent rows 16 banks apart
// Two LDS buffers: chunk c+1 is loaded into registers while chunk c is
// multiplied out from LDS, then written to the other buffer (one barrier per
// chunk).
__global__ __launch_bounds__(256) void route_dist_kernel(const float* __restrict__ q32,
                                                         const float* __restrict__ qn2, int nq,
                                                         const float* __restrict__ c32,
                                                         const float* __restrict__ cn2, int nlist,
                                                         int d16, int metric,
                                                         float* __restrict__ R) {
  __shared__ float ys[2][RT_T][RT_LD];
  __shared__ float xs[2][RT_T][RT_LD];
  // consecutive workgroups share a query tile (the centroids stay L2-resident)
  const int c0 = blockIdx.x * RT_T, q0 = blockIdx.y * RT_T;
  const int tid = threadIdx.x, g = tid >> 4, j = tid & 15;
  const int qa = (g >> 2) * 8, cb = g & 3;
  float acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = 0.0f;
  // staging: float4 i = tid + 256 p (p < 2) covers row i / 16, floats 4 (i % 16)
  float4 ry[2], rx[2];
  auto load = [&](int t0) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int i = tid + 256 * p, r = i >> 4, c4 = (i & 15) * 4;
      const bool in = t0 + c4 < d16;  // d16 is a multiple of 16
      const int q = min(q0 + r, nq - 1), c = min(c0 + r, nlist - 1);
      ry[p] = in ? *(const float4*)(q32 + (int64_t)q * d16 + t0 + c4) : make_float4(0, 0, 0, 0);
      rx[p] = in ? *(const float4*)(c32 + (int64_t)c * d16 + t0 + c4) : make_float4(0, 0, 0, 0);
    }
  };
  auto put = [&](int b) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int i = tid + 256 * p, r = i >> 4, c4 = (i & 15) * 4;
      *(float4*)&ys[b][r][c4] = ry[p];
      *(float4*)&xs[b][r][c4] = rx[p];
    }
  };
  // the norms of the lane's queries and centroids, loaded up front (a load
  // inside each writer's branch of the epilogue would wait out its latency
  // 64 times in a row)
  float qnl[8], cnl[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) qnl[a] = qn2[min(q0 + qa + a, nq - 1)];
#pragma unroll
  for (int b = 0; b < 8; ++b) cnl[b] = cn2[min(c0 + cb + 4 * b, nlist - 1)];
  const int nch = (d16 + RT_DC - 1) / RT_DC;
  load(0);
  put(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ++ch) {
    const int t0 = ch * RT_DC, b = ch & 1;
    if (ch + 1 < nch) load(t0 + RT_DC);
    const int w = min(RT_DC, d16 - t0);  // multiple of 16
    for (int t = j; t < w; t += 16) {
      float yv[8], xv[8];
#pragma unroll
      for (int a = 0; a < 8; ++a) yv[a] = ys[b][qa + a][t];
#pragma unroll
      for (int bb = 0; bb < 8; ++bb) xv[bb] = xs[b][cb + 4 * bb][t];
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int bb = 0; bb < 8; ++bb) acc[a][bb] = __builtin_fmaf(xv[bb], yv[a], acc[a][bb]);
    }
    if (ch + 1 < nch) put(b ^ 1);  // the other buffer: last read before the previous barrier
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int q = q0 + qa + a;
    const float qn = qnl[a];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const float dot = group16_tree(acc[a][b]);
      const int c = c0 + cb + 4 * b;
      if (j == ((a * 8 + b) & 15) && q < nq && c < nlist) {
        float v;
        if (metric == COSINE) {
          // CosineDist of centroid and query: max(1 - c.q / (|c| |q|), 0)
          v = 1.0f - dot / (__builtin_sqrtf(cnl[b]) * __builtin_sqrtf(qn));
          v = v > 0.0f ? v : 0.0f;
        } else {
          float tt = cnl[b] + qn;
          float sq = tt - 2.0f * dot;
          sq = sq > 0.0f ? sq : 0.0f;
          v = metric == SQEUCLIDEAN ? sq : __builtin_sqrtf(sq);
        }
        R[(int64_t)q * nlist + c] = v;
      }
    }
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

// Same selection for nprobe <= 64 and at most 1024 routing distances: one
// wavefront per query holds 16 keys per lane and extracts the kk smallest
// (dist, bucket) keys by repeated wavefront minima, i.e. in sorted order.
// mass > 0 (learned routing): the distances are negated logits; lists are
// taken in the same order until their softmax probabilities sum to `mass`
// (at most nprobe non-empty lists).
constexpr int PS_KPL = 16;
__global__ __launch_bounds__(256) void probe_select_wave_kernel(
    const float* __restrict__ R, int nq, int nlist, int avail, int nprobe, float mass,
    const int* __restrict__ list_off, int* __restrict__ probes, int* __restrict__ pcount,
    int* __restrict__ prows) {
  const int lane = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= nq) return;
  // the lane's distances and list sizes, every load issued before the first
  // use (clamped indices)
  uint64_t key[PS_KPL];
  int lsz[PS_KPL];
  float rv[PS_KPL];
#pragma unroll
  for (int u = 0; u < PS_KPL; ++u) {
    const int ic = min(lane + 64 * u, avail - 1);
    rv[u] = R[(int64_t)q * nlist + ic];
    lsz[u] = list_off[ic + 1] - list_off[ic];
  }
  uint64_t lmin = ~0ull;
#pragma unroll
  for (int u = 0; u < PS_KPL; ++u) {
    const int i = lane + 64 * u;
    key[u] = i < avail ? pack_key(rv[u], (uint32_t)i) : ~0ull;
    lmin = key[u] < lmin ? key[u] : lmin;
  }
  // softmax normaliser over the negated distances (mass mode)
  float rmin = 0.0f, zsum = 1.0f;
  if (mass > 0.0f) {
    uint64_t g = lmin;
    for (int o = 32; o >= 1; o >>= 1) {
      const uint64_t t = __shfl_xor(g, o);
      g = t < g ? t : g;
    }
    rmin = ord2f((uint32_t)(g >> 32));
    float zs = 0.0f;
#pragma unroll
    for (int u = 0; u < PS_KPL; ++u)
      if (key[u] != ~0ull) zs += __expf(rmin - ord2f((uint32_t)(key[u] >> 32)));
    for (int o = 32; o >= 1; o >>= 1) zs += __shfl_xor(zs, o);
    zsum = zs;
  }
  const int kk = mass > 0.0f ? avail : (nprobe < avail ? nprobe : avail);
  int m = 0, rows = 0;
  float cum = 0.0f;
  for (int r = 0; r < kk && m < nprobe && !(mass > 0.0f && cum >= mass); ++r) {
    uint64_t g = lmin;
    for (int o = 32; o >= 1; o >>= 1) {
      const uint64_t t = __shfl_xor(g, o);
      g = t < g ? t : g;
    }
    if (mass > 0.0f) cum += __expf(rmin - ord2f((uint32_t)(g >> 32))) / zsum;
    const bool own = lmin == g;
    int osz = 0;
    if (own) {  // the owner removes it (keys are unique: bucket in the low bits)
      lmin = ~0ull;
#pragma unroll
      for (int u = 0; u < PS_KPL; ++u) {
        if (key[u] == g) {
          key[u] = ~0ull;
          osz = lsz[u];
        }
        lmin = key[u] < lmin ? key[u] : lmin;
      }
    }
    const int b = (int)(g & 0xffffffffu);
    const int sz = __shfl(osz, __ffsll((long long)__ballot(own)) - 1);
    if (sz > 0) {
      if (lane == 0) probes[(int64_t)q * nprobe + m] = b;
      ++m;
      rows += sz;
    }
  }
  if (lane == 0) {
    for (int r = m; r < nprobe; ++r) probes[(int64_t)q * nprobe + r] = -1;
    pcount[q] = m;
    prows[q] = rows;
  }
}

// TopKRouting from a bf16 first pass (certified; the result equals the exact
// route followed by probe_select_wave_kernel).  One wavefront per query over
// the dense first-pass values S^ of the first `avail` centroids:
// 1. the kk-th smallest S^ (kk = min(nprobe, avail)) and its norms-form
//    exclusion threshold T: every centroid with S^ > T has an FP32 norms-form
//    distance above that of each of the kk centroids at or below the kk-th
//    S^, so it cannot be among the kk nearest;
// 2. the exact FP32 norms form (route_dist_kernel's arithmetic: A16 dot, then
//    (|c|^2 + |q|^2) - 2 c.q, clamped, square root unless SQEUCLIDEAN) of every
//    centroid with S^ <= T;
// 3. the kk smallest (distance, centroid) keys in order, empty buckets
//    skipped, as probe_select_wave_kernel writes them.
constexpr int RS_CAP = 64 * PS_KPL;  // centroids per query (avail <= RS_CAP)
__global__ __launch_bounds__(256) void route_select_kernel(RouteSelArgs a) {
  __shared__ uint32_t buf_all[4][256];
  __shared__ unsigned hist_all[4][256];
  __shared__ int cidx_all[4][RS_CAP];
  __shared__ float cval_all[4][RS_CAP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int qi = blockIdx.x * 4 + w;
  if (qi >= a.nq) return;
  const int q = a.q0 + qi;
  int* cidx = cidx_all[w];
  float* cval = cval_all[w];
  const float* row = a.dense + (int64_t)qi * a.ldd;
  const int avail = a.avail;
  const int kk = a.nprobe < avail ? a.nprobe : avail;
  float sv[PS_KPL];
  int lsz[PS_KPL];
#pragma unroll
  for (int u = 0; u < PS_KPL; ++u) {
    const int i = lane + 64 * u;
    const int ic = min(i, avail - 1);
    sv[u] = i < avail ? row[i] : INFINITY;
    lsz[u] = a.list_off[ic + 1] - a.list_off[ic];
  }
  const uint32_t kkey = wave_rank_key(row, avail, kk, buf_all[w], hist_all[w], lane);
  const double T = norms_form_threshold((double)ord2f(kkey), a.qe[q], a.qr[q], a.qn2[q], a.ex,
                                        a.rx, a.ra, a.d);
  const float Tf = T < 3.0e38 ? nextafterf((float)T * (1.0f + 1e-6f), INFINITY) : INFINITY;
  // candidates, compacted in lane order
  int pos[PS_KPL];
  int cnt = 0;
#pragma unroll
  for (int u = 0; u < PS_KPL; ++u) {
    const int i = lane + 64 * u;
    const bool c = i < avail && sv[u] <= Tf;
    const uint64_t m = __ballot(c);
    const int p = cnt + __popcll(m & ((1ull << lane) - 1ull));
    pos[u] = c ? p : -1;
    if (c) cidx[p] = i;
    cnt += __popcll(m);
  }
  wave_fence_sync();
  // exact norms form of the candidates, four per 16-lane group at a time, the
  // loads of 8 chunks issued before their multiply-adds (each candidate's A16
  // sum keeps its chunk order)
  const int g = lane >> 4, j = lane & 15;
  const float* y = a.q32 + (int64_t)q * a.d16;
  const int nch = a.d16 >> 4;
  for (int i0 = 4 * g; i0 < cnt; i0 += 16) {
    int cc[4];
    const float* x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cc[u] = cidx[i0 + u < cnt ? i0 + u : i0];
      x[u] = a.c32 + (int64_t)cc[u] * a.d16;
    }
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int c0 = 0;
    for (; c0 + 8 <= nch; c0 += 8) {
      float yv[8], xv[8][4];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int e = ((c0 + t) << 4) + j;
        yv[t] = y[e];
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[t][u] = x[u][e];
      }
#pragma unroll
      for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = __builtin_fmaf(xv[t][u], yv[t], acc[u]);
    }
    for (; c0 < nch; ++c0) {
      const int e = (c0 << 4) + j;
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = __builtin_fmaf(x[u][e], y[e], acc[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float dot = group16_tree(acc[u]);
      const float tt = a.cn2[cc[u]] + a.qn2[q];
      float sq = tt - 2.0f * dot;
      sq = sq > 0.0f ? sq : 0.0f;
      if (j == 0 && i0 + u < cnt) cval[i0 + u] = a.sq ? sq : __builtin_sqrtf(sq);
    }
  }
  wave_fence_sync();
  uint64_t key[PS_KPL];
  uint64_t lmin = ~0ull;
#pragma unroll
  for (int u = 0; u < PS_KPL; ++u) {
    key[u] = pos[u] >= 0 ? pack_key(cval[pos[u]], (uint32_t)(lane + 64 * u)) : ~0ull;
    lmin = key[u] < lmin ? key[u] : lmin;
  }
  int m = 0, rows = 0;
  for (int r = 0; r < kk; ++r) {
    uint64_t gk = lmin;
    for (int o = 32; o >= 1; o >>= 1) {
      const uint64_t t = __shfl_xor(gk, o);
      gk = t < gk ? t : gk;
    }
    const bool own = lmin == gk;
    int osz = 0;
    if (own) {  // the owner removes it (keys are unique: centroid in the low bits)
      lmin = ~0ull;
#pragma unroll
      for (int u = 0; u < PS_KPL; ++u) {
        if (key[u] == gk) {
          key[u] = ~0ull;
          osz = lsz[u];
        }
        lmin = key[u] < lmin ? key[u] : lmin;
      }
    }
    const int b = (int)(gk & 0xffffffffu);
    const int sz = __shfl(osz, __ffsll((long long)__ballot(own)) - 1);
    if (sz > 0) {
      if (lane == 0) a.probes[(int64_t)q * a.nprobe + m] = b;
      ++m;
      rows += sz;
    }
  }
  if (lane == 0) {
    for (int r = m; r < a.nprobe; ++r) a.probes[(int64_t)q * a.nprobe + r] = -1;
    a.pcount[q] = m;
    a.prows[q] = rows;
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

// Block-aggregated forms of the two kernels above, for nb lists that fit in
// LDS: counts (and slots) are taken in LDS first, then one global atomic per
// (block, list) reserves the block's range, so a list probed by thousands of
// queries does not serialize thousands of atomics on one address.
constexpr int LF_EPB = 1024;  // entries per block
__global__ __launch_bounds__(256) void list_count_agg_kernel(const int* __restrict__ probes,
                                                             const int* __restrict__ pcount,
                                                             int nq, int nprobe, int nb,
                                                             int* __restrict__ qcount) {
  extern __shared__ int lh[];
  for (int b = threadIdx.x; b < nb; b += 256) lh[b] = 0;
  __syncthreads();
  const int64_t tot = (int64_t)nq * nprobe, e0 = (int64_t)blockIdx.x * LF_EPB;
  for (int u = threadIdx.x; u < LF_EPB; u += 256) {
    const int64_t i = e0 + u;
    if (i >= tot) break;
    const int q = (int)(i / nprobe), r = (int)(i - (int64_t)q * nprobe);
    if (r < pcount[q]) atomicAdd(&lh[probes[i]], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += 256)
    if (lh[b]) atomicAdd(&qcount[b], lh[b]);
}

__global__ __launch_bounds__(256) void list_fill_agg_kernel(
    const int* __restrict__ probes, const int* __restrict__ pcount, int nq, int nprobe, int nb,
    const int* __restrict__ qoff, int* __restrict__ qfill, int* __restrict__ qlist,
    int* __restrict__ qlist_list, int* __restrict__ slot_of, int phase) {
  extern __shared__ int lh[];  // [nb] local counts, then the block's base slot
  for (int b = threadIdx.x; b < nb; b += 256) lh[b] = 0;
  __syncthreads();
  const int64_t tot = (int64_t)nq * nprobe, e0 = (int64_t)blockIdx.x * LF_EPB;
  constexpr int PER = LF_EPB / 256;
  int rank[PER], lst[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int64_t i = e0 + threadIdx.x + 256 * j;
    lst[j] = -1;
    if (i >= tot) continue;
    const int q = (int)(i / nprobe), r = (int)(i - (int64_t)q * nprobe);
    if (r >= pcount[q] || (r == 0) != (phase == 0)) continue;
    lst[j] = probes[i];
    rank[j] = atomicAdd(&lh[lst[j]], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += 256)
    if (lh[b]) lh[b] = atomicAdd(&qfill[b], lh[b]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    if (lst[j] < 0) continue;
    const int64_t i = e0 + threadIdx.x + 256 * j;
    const int q = (int)(i / nprobe), b = lst[j];
    const int pos = qoff[b] + lh[b] + rank[j];
    qlist[pos] = q;
    qlist_list[pos] = b;
    slot_of[i] = pos;
  }
}

// exclusive scan of one value per thread over a 1024-thread block
__device__ __forceinline__ int block_exclusive_scan(int v, int* wsum, int* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int i = 0; i < 16; ++i) {
      const int t = wsum[i];
      wsum[i] = run;
      run += t;
    }
    *total = run;
  }
  __syncthreads();
  const int r = wsum[w] + inc - v;
  __syncthreads();
  return r;
}

// one block: exclusive scan of qcount (list order) and of the work-item counts
// (size-descending list order); writes n_items and the entry count.
__global__ __launch_bounds__(1024) void plan_kernel(const int* __restrict__ qcount, int nb,
                                                    const int* __restrict__ list_off,
                                                    const int* __restrict__ lorder, int bn,
                                                    int rows_item, int* __restrict__ qoff,
                                                    int* __restrict__ ioff,
                                                    int* __restrict__ n_items,
                                                    int* __restrict__ n_entries) {
  __shared__ int wsum[16];
  __shared__ int total;
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int s0 = tid * per, e0 = min(nb, s0 + per);
  // query offsets in bucket order
  int sum = 0;
  for (int i = s0; i < e0; ++i) sum += qcount[i];
  int run = block_exclusive_scan(sum, wsum, &total);
  if (tid == 0) *n_entries = total;
  for (int i = s0; i < e0; ++i) {
    qoff[i] = run;
    run += qcount[i];
  }
  // items in size-descending list order
  sum = 0;
  for (int i = s0; i < e0; ++i) {
    const int b = lorder[i];
    const int nqc = (qcount[b] + bn - 1) / bn;
    const int sz = list_off[b + 1] - list_off[b];
    sum += nqc * ((sz + rows_item - 1) / rows_item);
  }
  run = block_exclusive_scan(sum, wsum, &total);
  if (tid == 0) *n_items = total;
  for (int i = s0; i < e0; ++i) {
    ioff[i] = run;
    const int b = lorder[i];
    const int nqc = (qcount[b] + bn - 1) / bn;
    const int sz = list_off[b + 1] - list_off[b];
    run += nqc * ((sz + rows_item - 1) / rows_item);
  }
}

// phase 0 places the entries of first-probed lists, phase 1 the others, so a
// list's entries for which it is the query's first probe come first
__global__ void list_fill_kernel(const int* __restrict__ probes, const int* __restrict__ pcount,
                                 int nq, int nprobe, const int* __restrict__ qoff,
                                 int* __restrict__ qfill, int* __restrict__ qlist,
                                 int* __restrict__ qlist_list, int* __restrict__ slot_of,
                                 int phase) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nq * nprobe) return;
  const int q = (int)(i / nprobe), r = (int)(i - (int64_t)q * nprobe);
  if (r >= pcount[q] || (r == 0) != (phase == 0)) return;
  const int b = probes[i];
  const int pos = atomicAdd(&qfill[b], 1);
  qlist[qoff[b] + pos] = q;
  qlist_list[qoff[b] + pos] = b;
  slot_of[i] = qoff[b] + pos;
}

// Refinement items: for each flagged query, its probed lists in ROWS_ITEM
// chunks as single-query items (slot = the query's entry in the list).
__global__ void refine_items_kernel(const int* __restrict__ flist, const int* __restrict__ nflag,
                                    int nprobe, const int* __restrict__ probes,
                                    const int* __restrict__ pcount,
                                    const int* __restrict__ slot_of,
                                    const int* __restrict__ list_off, int rows_item,
                                    ScanItem* __restrict__ items, int* __restrict__ n_items) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = (int)(i / nprobe), r = (int)(i - (int64_t)f * nprobe);
  if (f >= *nflag) return;
  const int q = flist[f];
  if (r >= pcount[q]) return;
  const int b = probes[(int64_t)q * nprobe + r];
  const int slot = slot_of[(int64_t)q * nprobe + r];
  const int lo = list_off[b], hi = list_off[b + 1];
  const int ns = (hi - lo + rows_item - 1) / rows_item;
  const int base = atomicAdd(n_items, ns);
  for (int s = 0; s < ns; ++s) {
    ScanItem it;
    it.qoff = slot;
    it.nq = 1;
    it.row_begin = lo + s * rows_item;
    it.row_end = min(hi, it.row_begin + rows_item);
    it.region = 0;
    items[base + s] = it;
  }
}

// Residual query rows for every (list, query) entry: bf16(q - c_list), its
// squared norm (FP64 sum), and per query the maxima of |r - bf16(r)| and
// |bf16(r)| over its entries (the bounds of the certification).
__global__ __launch_bounds__(256) void residual_queries_kernel(
    const float* __restrict__ q32, const float* __restrict__ c32, int d, int d16, int kp,
    int normalize,
    const int* __restrict__ qlist, const int* __restrict__ qlist_list, const int* __restrict__ n_dev,
    bf16_t* __restrict__ qres, float* __restrict__ rnorm, unsigned* __restrict__ qe_max,
    unsigned* __restrict__ qr_max) {
  const int lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= *n_dev) return;
  const int q = qlist[e], b = qlist_list[e];
  const float* y = q32 + (int64_t)q * d16;
  const float* cen = c32 + (int64_t)b * d16;
  double scale = 1.0;
  if (normalize) {  // cosine: the residual of the unit vector (as the base rows)
    double ss = 0.0;
    for (int c = lane; c < d; c += 64) ss += (double)y[c] * (double)y[c];
    for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o);
    scale = ss > 0.0 ? 1.0 / sqrt(ss) : 0.0;
  }
  double nb = 0.0, ee = 0.0;
  // four columns per lane per pass, their loads issued together (clamped)
  for (int c0 = 0; c0 < kp; c0 += 256) {
    float yv[4], cv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int cc = min(c0 + 64 * u + lane, d - 1);
      yv[u] = y[cc];
      cv[u] = cen[cc];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = c0 + 64 * u + lane;
      if (c >= kp) break;
      const double t = c < d ? (double)yv[u] * scale - (double)cv[u] : 0.0;
      const bf16_t v = (bf16_t)(float)t;
      qres[e * kp + c] = v;
      const double bd = (double)(float)v;
      nb += bd * bd;
      ee += (t - bd) * (t - bd);
    }
  }
  for (int o = 32; o >= 1; o >>= 1) {
    nb += __shfl_xor(nb, o);
    ee += __shfl_xor(ee, o);
  }
  if (lane == 0) {
    rnorm[e] = (float)nb;
    const float er = (float)(sqrt(ee) * (1.0 + 1e-6)) + 1e-30f;
    const float nr = (float)(sqrt(nb) * (1.0 + 1e-6)) + 1e-30f;
    atomicMax(qe_max + q, __float_as_uint(er));
    atomicMax(qr_max + q, __float_as_uint(nr));
  }
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

// one block: per list (size-descending order) the item count of its first-probe
// queries, exclusive scan into ioff, total into *n_items; then per list the items
__global__ __launch_bounds__(1024) void first_plan_kernel(const int* __restrict__ first_cnt,
                                                          int nb, const int* __restrict__ list_off,
                                                          const int* __restrict__ lorder, int bn,
                                                          int rows_item, int max_rows,
                                                          int* __restrict__ ioff,
                                                          int* __restrict__ n_items) {
  __shared__ int wsum[16];
  __shared__ int total;
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int s0 = tid * per, e0 = min(nb, s0 + per);
  // rows of a list taken for the threshold: all, or its first max_rows
  auto rows_of = [&](int b) {
    const int sz = list_off[b + 1] - list_off[b];
    return max_rows > 0 && sz > max_rows ? max_rows : sz;
  };
  int sum = 0;
  for (int i = s0; i < e0; ++i) {
    const int b = lorder[i];
    sum += ((first_cnt[b] + bn - 1) / bn) * ((rows_of(b) + rows_item - 1) / rows_item);
  }
  int run = block_exclusive_scan(sum, wsum, &total);
  if (tid == 0) *n_items = total;
  for (int i = s0; i < e0; ++i) {
    ioff[i] = run;
    const int b = lorder[i];
    run += ((first_cnt[b] + bn - 1) / bn) * ((rows_of(b) + rows_item - 1) / rows_item);
  }
}

__global__ void first_items_kernel(const int* __restrict__ first_cnt, int nb,
                                   const int* __restrict__ list_off,
                                   const int* __restrict__ lorder, const int* __restrict__ qoff,
                                   const int* __restrict__ ioff, int bn, int rows_item,
                                   int max_rows, ScanItem* __restrict__ items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int b = lorder[i];
  const int c0 = first_cnt[b];
  const int nqc = (c0 + bn - 1) / bn;
  const int lo = list_off[b];
  const int hi = max_rows > 0 ? min(list_off[b + 1], lo + max_rows) : list_off[b + 1];
  const int ns = (hi - lo + rows_item - 1) / rows_item;
  int w = ioff[i];
  for (int s = 0; s < ns; ++s)
    for (int c = 0; c < nqc; ++c) {
      ScanItem it;
      it.qoff = qoff[b] + c * bn;
      it.nq = min(bn, c0 - c * bn);
      it.row_begin = lo + s * rows_item;
      it.row_end = min(hi, it.row_begin + rows_item);
      it.region = lo;
      items[w++] = it;
    }
}

// Upper bound, in the metric's own domain, of the FP32 exact value of a row
// whose residual first-pass value is S: |q - x| <= sqrt(S + G) + qe + ex, then
// the FP32 rounding of the exact forms on top.
__device__ inline double metric_upper_bound(double S, int metric, double qe, double qr,
                                            double ex, double rx, int d) {
  const double u = 1.0 / 16777216.0;
  const double R = qr + rx;
  const double G = (2.0 * d + 8.0) * u * R * R;
  const double t = sqrt(S + G > 0.0 ? S + G : 0.0) + qe + ex;
  const double rho = (d + 8.0) * u;
  if (metric == EUCLIDEAN) return t * (1.0 + rho) * (1.0 + 4.0 * u);
  if (metric == SQEUCLIDEAN) return t * t * (1.0 + rho) * (1.0 + 4.0 * u);
  // cosine: the residual rows are unit vectors minus a common centre, so
  // |q^ - x^|^2 = 2 (1 - cos); FP32 cosine form error below (2d + 16) u
  return 0.5 * t * t + (2.0 * d + 16.0) * u;
}

// one wavefront per query, four per workgroup
__global__ __launch_bounds__(256) void ivf_tau_dense_kernel(IvfDenseTauArgs a) {
  __shared__ uint32_t buf_all[4][256];
  __shared__ unsigned hist_all[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = blockIdx.x * 4 + w;
  if (q >= a.nq) return;
  const int np = a.pcount[q];
  const int total = a.prows[q];
  const int k_eff = a.k < total ? a.k : total;
  int m = 0;
  if (np > 0) {
    const int b = a.probes[(int64_t)q * a.nprobe];
    m = a.list_off[b + 1] - a.list_off[b];
    if (a.max_rows > 0 && m > a.max_rows) m = a.max_rows;  // the rows the scan covered
  }
  if (lane == 0) a.k_eff_q[q] = k_eff;
  if (k_eff <= 0 || total <= a.emit_all_below || m < k_eff) {
    // nothing to bound with (or everything fits): emit every probed row
    if (lane == 0) a.tau[q] = INFINITY;
    return;
  }
  // 4-row minima (8 per 32-row block, +inf past the list): each is the value
  // of a distinct row, so their k-th smallest bounds the k-th smallest value
  // from above, as the k-th value itself does (grp_min >= 4 k leaves at least
  // k finite minima)
  const int mv = (a.grp_min > 0 && m >= a.grp_min) ? 8 * ((m + 31) / 32) : m;
  const uint32_t key =
      wave_rank_key(a.dense + (int64_t)q * a.ldd, mv, k_eff, buf_all[w], hist_all[w], lane);
  if (lane == 0) {
    const double dub = metric_upper_bound((double)ord2f(key), a.metric, a.qe[q], a.qr[q], a.ex,
                                          a.rx, a.d);
    const double T = exclusion_threshold(dub, a.metric, a.qe[q], a.qr[q], a.ex, a.rx, a.d);
    a.tau[q] = T < 3.0e38 ? nextafterf((float)T * (1.0f + 1e-6f), INFINITY) : INFINITY;
  }
}

constexpr int TAU_M = 1024;

// Emission threshold of each query from the exact distances of the first
// TAU_M rows of its probed lists (probe order): the k-th smallest exact value
// is an upper bound of the final k-th, so T* of it certifies the emitted set.
__global__ __launch_bounds__(256) void ivf_tau_kernel(IvfTauArgs a) {
  constexpr int MAXL = 64;  // probed lists the sample may span
  __shared__ float vals[TAU_M];
  __shared__ int cum[MAXL + 1];
  __shared__ int s_nl, s_m;
  __shared__ unsigned hist[256];
  __shared__ unsigned s_sel, s_want;
  const int64_t ent = blockIdx.x;
  if (ent >= *a.n_entries) return;
  const int q = a.qlist[ent];
  if (a.probes[(int64_t)q * a.nprobe] != a.qlist_list[ent]) return;
  const int tid = threadIdx.x, g = tid >> 4, j = tid & 15;
  const int np = a.pcount[q];
  const int total = a.prows[q];
  const int mmax = a.sample_rows < TAU_M ? a.sample_rows : TAU_M;
  const int m0 = total < mmax ? total : mmax;
  const int k_eff = a.k < total ? a.k : total;
  if (tid == 0) a.k_eff_q[q] = k_eff;
  if (m0 <= 0 || k_eff <= 0 || total <= a.emit_all_below) {
    if (tid == 0) a.tau[q] = INFINITY;
    return;
  }
  if (tid == 0) {  // prefix sums of the list sizes until the sample is covered
    int run = 0, r = 0;
    for (; r < np && r < MAXL && run < m0; ++r) {
      cum[r] = run;
      const int b = a.probes[(int64_t)q * a.nprobe + r];
      run += a.list_off[b + 1] - a.list_off[b];
    }
    cum[r] = run;
    s_nl = r;
    s_m = run < m0 ? run : m0;
  }
  __syncthreads();
  const int nl = s_nl, m = s_m;
  if (m < k_eff) {  // sample too small to bound the k-th distance: emit everything
    if (tid == 0) a.tau[q] = INFINITY;
    return;
  }
  auto row_of = [&](int i) {
    int r = 0;
    while (r + 1 < nl && cum[r + 1] <= i) ++r;
    const int b = a.probes[(int64_t)q * a.nprobe + r];
    return a.list_off[b] + (i - cum[r]);
  };
  const float* y = a.q32 + (int64_t)q * a.d16;
  const bool diff = a.metric == EUCLIDEAN || a.metric == SQEUCLIDEAN;
  const float qn = diff ? 0.0f : a.qn2[q];
  for (int i0 = 4 * g; i0 < m; i0 += 64) {
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int row[4];
    const float* x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      row[u] = row_of(min(i0 + u, m - 1));
      x[u] = a.x32 + (int64_t)row[u] * a.d16;
    }
    if (diff) {
      for (int t = 0; t < a.d16; t += 16) {
        const float yv = y[t + j];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float df = x[u][t + j] - yv;
          acc[u] = __builtin_fmaf(df, df, acc[u]);
        }
      }
    } else {
      for (int t = 0; t < a.d16; t += 16) {
        const float yv = y[t + j];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = __builtin_fmaf(x[u][t + j], yv, acc[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float s = group16_tree(acc[u]);
      const float v = finish_metric(s, a.metric, diff ? 0.0f : a.xn2[row[u]], qn);
      if (j == 0 && i0 + u < m) vals[i0 + u] = v;
    }
  }
  __syncthreads();
  // k_eff-th smallest value: radix select on the float order keys
  uint32_t prefix = 0, mask = 0;
  if (tid == 0) s_want = (unsigned)k_eff;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    for (int i = tid; i < m; i += 256) {
      const uint32_t o = f2ord(vals[i]);
      if ((o & mask) == prefix) atomicAdd(&hist[(o >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid < 64) {
      const unsigned h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2],
                     h3 = hist[4 * tid + 3];
      const unsigned tot = h0 + h1 + h2 + h3;
      unsigned inc = tot;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(inc, o);
        if (tid >= o) inc += t;
      }
      const unsigned excl = inc - tot, want = s_want;
      if (excl < want && inc >= want) {
        unsigned cc = excl, b;
        if (cc + h0 >= want) b = 4 * tid;
        else if (cc + h0 + h1 >= want) { b = 4 * tid + 1; cc += h0; }
        else if (cc + h0 + h1 + h2 >= want) { b = 4 * tid + 2; cc += h0 + h1; }
        else { b = 4 * tid + 3; cc += h0 + h1 + h2; }
        s_sel = b;
        s_want = want - cc;
      }
    }
    __syncthreads();
    prefix |= s_sel << shift;
    mask |= 255u << shift;
    __syncthreads();
  }
  if (tid == 0) {
    const float dk = ord2f(prefix);
    const double T = exclusion_threshold(dk, a.metric, a.qe[q], a.qr[q], a.ex, a.rx, a.d);
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
                       const float* cn2, int nlist, int d16, float* R, hipStream_t st,
                       int metric) {
  if (nq <= 0 || nlist <= 0) return;
  dim3 grid((unsigned)ceil_div(nlist, RT_T), (unsigned)ceil_div(nq, RT_T));
  hipLaunchKernelGGL(route_dist_kernel, grid, dim3(256), 0, st, q32, qn2, nq, c32, cn2, nlist, d16,
                     metric, R);
}

void launch_probe_select(const float* R, int nq, int nlist, int avail, int nprobe, float mass,
                         const int* list_off, int* probes, int* pcount, int* prows,
                         hipStream_t st) {
  if (nq <= 0) return;
  if (mass > 0.0f || (nprobe <= 64 && avail <= 64 * PS_KPL))
    hipLaunchKernelGGL(probe_select_wave_kernel, dim3((unsigned)ceil_div(nq, 4)), dim3(256), 0, st,
                       R, nq, nlist, avail, nprobe, mass, list_off, probes, pcount, prows);
  else
    hipLaunchKernelGGL(probe_select_kernel, dim3(nq), dim3(256), 0, st, R, nq, nlist, avail,
                       nprobe, list_off, probes, pcount, prows);
}

int route_select_cap() { return RS_CAP; }

void launch_route_select(const RouteSelArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  hipLaunchKernelGGL(route_select_kernel, dim3((unsigned)ceil_div(a.nq, 4)), dim3(256), 0, st, a);
}

void launch_ivf_plan(const int* probes, const int* pcount, int nq, int nprobe, int nb,
                     const int* list_off, const int* lorder, int bn, int rows_item, int* qcount,
                     int* qfill, int* qoff, int* ioff, int* qlist, int* qlist_list,
                     int* slot_of, ScanItem* items, int* n_items, int* n_entries,
                     int* first_cnt, hipStream_t st) {
  const int64_t tot = (int64_t)nq * nprobe;
  const unsigned g = (unsigned)ceil_div(tot, 256);
  // LDS-aggregated counting when the lists fit in 32 KiB of LDS
  const bool agg = nb <= 8192;
  const unsigned ga = (unsigned)ceil_div(tot, LF_EPB);
  const size_t lds = sizeof(int) * (size_t)nb;
  if (agg)
    hipLaunchKernelGGL(list_count_agg_kernel, dim3(ga), dim3(256), lds, st, probes, pcount, nq,
                       nprobe, nb, qcount);
  else
    hipLaunchKernelGGL(list_count_kernel, dim3(g), dim3(256), 0, st, probes, pcount, nq, nprobe,
                       qcount);
  hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(1024), 0, st, qcount, nb, list_off, lorder, bn,
                     rows_item, qoff, ioff, n_items, n_entries);
  for (int phase = 0; phase < 2; ++phase) {
    if (agg)
      hipLaunchKernelGGL(list_fill_agg_kernel, dim3(ga), dim3(256), lds, st, probes, pcount, nq,
                         nprobe, nb, qoff, qfill, qlist, qlist_list, slot_of, phase);
    else
      hipLaunchKernelGGL(list_fill_kernel, dim3(g), dim3(256), 0, st, probes, pcount, nq, nprobe,
                         qoff, qfill, qlist, qlist_list, slot_of, phase);
    if (phase == 0 && first_cnt)
      MANN_HIP(hipMemcpyAsync(first_cnt, qfill, sizeof(int) * nb, hipMemcpyDeviceToDevice, st));
  }
  hipLaunchKernelGGL(items_fill_kernel, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, st, qcount,
                     nb, list_off, lorder, qoff, ioff, bn, rows_item, items);
}

void launch_residual_queries(const float* q32, const float* c32, int d, int d16, int kp,
                             int normalize, const int* qlist, const int* qlist_list, const int* n_dev,
                             int64_t max_entries, bf16_t* qres, float* rnorm, unsigned* qe_max,
                             unsigned* qr_max, hipStream_t st) {
  if (max_entries <= 0) return;
  hipLaunchKernelGGL(residual_queries_kernel, dim3((unsigned)ceil_div(max_entries, 4)), dim3(256),
                     0, st, q32, c32, d, d16, kp, normalize, qlist, qlist_list, n_dev, qres, rnorm,
                     qe_max,
                     qr_max);
}

void launch_refine_items(const int* flist, const int* nflag, int nflag_host, int nprobe,
                         const int* probes, const int* pcount, const int* slot_of,
                         const int* list_off, int rows_item, ScanItem* items, int* n_items,
                         hipStream_t st) {
  const int64_t tot = (int64_t)nflag_host * nprobe;
  if (tot <= 0) return;
  hipLaunchKernelGGL(refine_items_kernel, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st,
                     flist, nflag, nprobe, probes, pcount, slot_of, list_off, rows_item, items,
                     n_items);
}

int ivf_tau_rows_max() { return TAU_M; }

void launch_first_items(const int* first_cnt, const int* qoff, int nb, const int* list_off,
                        const int* lorder, int bn, int rows_item, int max_rows, int* ioff,
                        ScanItem* items, int* n_items, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(first_plan_kernel, dim3(1), dim3(1024), 0, st, first_cnt, nb, list_off, lorder,
                     bn, rows_item, max_rows, ioff, n_items);
  hipLaunchKernelGGL(first_items_kernel, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, st,
                     first_cnt, nb, list_off, lorder, qoff, ioff, bn, rows_item, max_rows, items);
}

void launch_ivf_tau_dense(const IvfDenseTauArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  hipLaunchKernelGGL(ivf_tau_dense_kernel, dim3((unsigned)ceil_div(a.nq, 4)), dim3(256), 0, st,
                     a);
}

void launch_ivf_tau(const IvfTauArgs& a, hipStream_t st) {
  if (a.nq <= 0 || a.max_entries <= 0) return;
  // queries without an entry (nothing probed) keep these defaults
  fill_f32(a.tau, a.nq, INFINITY, st);
  fill_i32(a.k_eff_q, a.nq, 0, st);
  hipLaunchKernelGGL(ivf_tau_kernel, dim3((unsigned)a.max_entries), dim3(256), 0, st, a);
}

void launch_probe_ranges(const int* probes, const int* pcount, int nq, int nprobe,
                         const int* list_off, int2* ranges, hipStream_t st) {
  const int64_t tot = (int64_t)nq * nprobe;
  if (tot <= 0) return;
  hipLaunchKernelGGL(probe_ranges_kernel, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, st,
                     probes, pcount, nq, nprobe, list_off, ranges);
}

}  // namespace mann
