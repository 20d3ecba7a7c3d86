// Rank selection on float order keys by one wavefront, shared by the
// threshold kernels.
#pragma once

#include "launch.h"

namespace mann {

// ---- one wavefront per row ---------------------------------------------------

// Bitonic sort (ascending) of 64 R keys across the wavefront: element e lives
// in lane e % 64, register e / 64.
template <int R>
__device__ __forceinline__ void wave_sort_u32(uint32_t (&v)[R], int lane) {
  constexpr int N = 64 * R;
#pragma unroll
  for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= 64) {
        const int sr = stride >> 6;
#pragma unroll
        for (int jj = 0; jj < R; ++jj) {
          if (jj & sr) continue;
          const bool asc = ((jj * 64 + lane) & size) == 0;
          const uint32_t lo = v[jj], hi = v[jj | sr];
          if ((lo > hi) == asc) {
            v[jj] = hi;
            v[jj | sr] = lo;
          }
        }
      } else {
#pragma unroll
        for (int jj = 0; jj < R; ++jj) {
          const uint32_t o = (uint32_t)__shfl_xor((int)v[jj], stride);
          const bool asc = ((jj * 64 + lane) & size) == 0;
          const bool lower = (lane & stride) == 0;
          v[jj] = (lower == asc) ? min(v[jj], o) : max(v[jj], o);
        }
      }
    }
  }
}

// The rank-th smallest of the cnt (rank <= cnt <= 256) keys collected in buf.
__device__ __forceinline__ uint32_t wave_rank_of_collected(const uint32_t* buf, int cnt, int rank,
                                                           int lane) {
  wave_fence_sync();
  if (cnt <= 64) {  // the usual case: a one-register sort
    uint32_t v1[1] = {lane < cnt ? buf[lane] : 0xffffffffu};
    wave_fence_sync();
    wave_sort_u32<1>(v1, lane);
    return (uint32_t)__shfl((int)v1[0], rank - 1);
  }
  uint32_t v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = j * 64 + lane < cnt ? buf[j * 64 + lane] : 0xffffffffu;
  wave_fence_sync();
  wave_sort_u32<4>(v, lane);
  const int e = rank - 1;
  uint32_t r = v[0];
#pragma unroll
  for (int j = 1; j < 4; ++j)
    if (e / 64 == j) r = v[j];
  return (uint32_t)__shfl((int)r, e % 64);
}

// Rows of at most 64 P values (rank <= 64): the row stays in registers
// between the two passes.  The loads are unconditional (indices clamped into
// the row), so straight-line code keeps all of them in flight at once.
// Returns false when more than 256 values qualify (the caller falls back).
template <int P>
__device__ __forceinline__ bool wave_rank_cached(const float* row, int m, int rank, uint32_t* buf,
                                                 int lane, uint32_t& key) {
  float x[P];
#pragma unroll
  for (int t = 0; t < P; ++t) {
    const int i = t * 64 + lane;
    x[t] = row[i < m ? i : m - 1];
  }
  uint32_t o[P];
  uint32_t mn[1] = {0xffffffffu};
#pragma unroll
  for (int t = 0; t < P; ++t) {
    o[t] = t * 64 + lane < m ? f2ord(x[t]) : 0xffffffffu;
    mn[0] = min(mn[0], o[t]);
  }
  wave_sort_u32<1>(mn, lane);
  const uint32_t U = (uint32_t)__shfl((int)mn[0], rank - 1);
  const unsigned long long lt = (1ull << lane) - 1ull;
  int cnt = 0;
  // sentinels (0xffffffff) lie above every finite or infinite value's key
#pragma unroll
  for (int t = 0; t < P; ++t) {
    if (t * 64 >= m) break;
    const bool p = o[t] <= U;
    const unsigned long long b = __ballot(p);
    const int pos = cnt + __popcll(b & lt);
    if (p && pos < 256) buf[pos] = o[t];
    cnt += __popcll(b);
  }
  if (cnt > 256) return false;
  key = wave_rank_of_collected(buf, cnt, rank, lane);
  return true;
}

// rank-th smallest (1-based, clamped to [1, m]) of the m floats row[0..m) as
// an order key, by one wavefront; buf (256 keys) and hist (256 counters) are
// the wavefront's LDS.  For rank <= 64 the rank-th smallest of the 64 lane
// minima bounds the answer from above (they are rank distinct elements); the
// values at or below that bound (usually a few more than rank) are collected
// and sorted.  Otherwise, or when more than 256 qualify: a 4-pass radix select
// over all m values.  (Inlined into its kernels, so the row is known to be
// global memory and the scratch LDS.)
__device__ __forceinline__ uint32_t wave_rank_key(const float* row, int m, int rank, uint32_t* buf,
                                                  unsigned* hist, int lane) {
  if (m <= 0) return 0xffffffffu;  // no values: the largest key
  rank = rank < 1 ? 1 : (rank > m ? m : rank);
  if (rank <= 64) {
    uint32_t key;
    if (m <= 64 * 32) {
      if (wave_rank_cached<32>(row, m, rank, buf, lane, key)) return key;
    } else {
      // two streaming passes, eight loads in flight per lane (clamped indices)
      constexpr int U8 = 8;
      uint32_t mn[1] = {0xffffffffu};
      for (int i0 = 0; i0 < m; i0 += 64 * U8) {
        float x[U8];
#pragma unroll
        for (int u = 0; u < U8; ++u) {
          const int i = i0 + 64 * u + lane;
          x[u] = row[i < m ? i : m - 1];
        }
#pragma unroll
        for (int u = 0; u < U8; ++u)
          if (i0 + 64 * u + lane < m) mn[0] = min(mn[0], f2ord(x[u]));
      }
      wave_sort_u32<1>(mn, lane);
      const uint32_t U = (uint32_t)__shfl((int)mn[0], rank - 1);
      const unsigned long long lt = (1ull << lane) - 1ull;
      int cnt = 0;
      for (int i0 = 0; i0 < m; i0 += 64 * U8) {
        float x[U8];
#pragma unroll
        for (int u = 0; u < U8; ++u) {
          const int i = i0 + 64 * u + lane;
          x[u] = row[i < m ? i : m - 1];
        }
#pragma unroll
        for (int u = 0; u < U8; ++u) {
          const int i = i0 + 64 * u + lane;
          const uint32_t v = i < m ? f2ord(x[u]) : 0xffffffffu;
          const bool p = i < m && v <= U;
          const unsigned long long b = __ballot(p);
          const int pos = cnt + __popcll(b & lt);
          if (p && pos < 256) buf[pos] = v;
          cnt += __popcll(b);
        }
      }
      if (cnt <= 256) return wave_rank_of_collected(buf, cnt, rank, lane);
    }
  }
  uint32_t prefix = 0, mask = 0;
  unsigned want = (unsigned)rank;
  for (int shift = 24; shift >= 0; shift -= 8) {
#pragma unroll
    for (int b = 0; b < 4; ++b) hist[4 * lane + b] = 0;
    wave_fence_sync();
    for (int i = lane; i < m; i += 64) {
      const uint32_t o = f2ord(row[i]);
      if ((o & mask) == prefix) atomicAdd(&hist[(o >> shift) & 255u], 1u);
    }
    wave_fence_sync();
    const unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                   h3 = hist[4 * lane + 3];
    const unsigned tot = h0 + h1 + h2 + h3;
    unsigned inc = tot;
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned t = __shfl_up(inc, o);
      if (lane >= o) inc += t;
    }
    const unsigned excl = inc - tot;
    const bool here = excl < want && inc >= want;
    const unsigned long long hm = __ballot(here);
    const int src = __ffsll((long long)hm) - 1;
    unsigned bin = 0, nw = 0;
    if (here) {
      const unsigned c = excl;
      if (c + h0 >= want) { bin = 4 * lane; nw = want - c; }
      else if (c + h0 + h1 >= want) { bin = 4 * lane + 1; nw = want - c - h0; }
      else if (c + h0 + h1 + h2 >= want) { bin = 4 * lane + 2; nw = want - c - h0 - h1; }
      else { bin = 4 * lane + 3; nw = want - c - h0 - h1 - h2; }
    }
    bin = (unsigned)__shfl((int)bin, src);
    want = (unsigned)__shfl((int)nw, src);
    prefix |= bin << shift;
    mask |= 255u << shift;
    wave_fence_sync();
  }
  return prefix;
}

// rank-th smallest (1-based, clamped to [1, m]) of m floats by a whole
// 256-thread workgroup, for long rows.  An upper bound U of the answer: for
// rank <= 64 the smallest of the four wavefronts' rank-th smallest values of
// their 4,096-value slices (each is the rank-th smallest of a subset), else the
// rank-th smallest of the 256 thread minima (rank <= 256).  The values at or
// below U are collected in LDS (up to 1024) and sorted by wavefront 0;
// otherwise one wavefront runs the radix select of wave_rank_key over the row.
// All threads return the key.
struct BlockRankScratch {
  uint32_t mins[256];
  uint32_t buf[1024];
  unsigned hist[4][256];
  unsigned cnt;
  uint32_t key;
};

__device__ __forceinline__ uint32_t block_rank_key(const float* row, int m, int rank,
                                                   BlockRankScratch& sm) {
  rank = rank < 1 ? 1 : (rank > m ? m : rank);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (rank <= 64 && m >= 4 * 4096) {
    const uint32_t uw = wave_rank_key(row + w * 4096, 4096, rank, sm.buf + w * 256, sm.hist[w],
                                      lane);
    if (lane == 0) sm.mins[w] = uw;
    if (tid == 0) sm.cnt = 0;
    __syncthreads();
    if (tid == 0) sm.key = min(min(sm.mins[0], sm.mins[1]), min(sm.mins[2], sm.mins[3]));
  } else if (rank <= 256) {
    uint32_t mn = 0xffffffffu;
    for (int i0 = 0; i0 < m; i0 += 256 * 8) {
      float x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + 256 * u + tid;
        x[u] = row[i < m ? i : m - 1];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (i0 + 256 * u + tid < m) mn = min(mn, f2ord(x[u]));
    }
    sm.mins[tid] = mn;
    if (tid == 0) sm.cnt = 0;
    __syncthreads();
    if (w == 0) {
      uint32_t v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = sm.mins[j * 64 + lane];
      wave_sort_u32<4>(v, lane);
      const int e = rank - 1;
      uint32_t r = v[0];
#pragma unroll
      for (int j = 1; j < 4; ++j)
        if (e / 64 == j) r = v[j];
      r = (uint32_t)__shfl((int)r, e % 64);
      if (lane == 0) sm.key = r;
    }
  }
  if (rank <= 256) {
    __syncthreads();
    const uint32_t U = sm.key;
    // eight loads in flight per thread: the pass is bound by memory latency
    int i = tid;
    for (; i + 7 * 256 < m; i += 8 * 256) {
      uint32_t o[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) o[u] = f2ord(row[i + u * 256]);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (o[u] <= U) {
          const unsigned pos = atomicAdd(&sm.cnt, 1u);
          if (pos < 1024u) sm.buf[pos] = o[u];
        }
    }
    for (; i < m; i += 256) {
      const uint32_t o = f2ord(row[i]);
      if (o <= U) {
        const unsigned pos = atomicAdd(&sm.cnt, 1u);
        if (pos < 1024u) sm.buf[pos] = o;
      }
    }
    __syncthreads();
    const unsigned c = sm.cnt;
    if (c <= 256u) {
      if (w == 0) {
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          v[j] = (unsigned)(j * 64 + lane) < c ? sm.buf[j * 64 + lane] : 0xffffffffu;
        wave_sort_u32<4>(v, lane);
        const int e = rank - 1;
        uint32_t r = v[0];
#pragma unroll
        for (int j = 1; j < 4; ++j)
          if (e / 64 == j) r = v[j];
        r = (uint32_t)__shfl((int)r, e % 64);
        if (lane == 0) sm.key = r;
      }
      __syncthreads();
      return sm.key;
    }
    if (c <= 1024u) {
      if (w == 0) {
        uint32_t v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j)
          v[j] = (unsigned)(j * 64 + lane) < c ? sm.buf[j * 64 + lane] : 0xffffffffu;
        wave_sort_u32<16>(v, lane);
        const int e = rank - 1;
        uint32_t r = v[0];
#pragma unroll
        for (int j = 1; j < 16; ++j)
          if (e / 64 == j) r = v[j];
        r = (uint32_t)__shfl((int)r, e % 64);
        if (lane == 0) sm.key = r;
      }
      __syncthreads();
      return sm.key;
    }
  }
  if (w == 0) {
    const uint32_t r = wave_rank_key(row, m, rank, sm.buf, sm.hist[0], lane);
    if (lane == 0) sm.key = r;
  }
  __syncthreads();
  return sm.key;
}

}  // namespace mann
