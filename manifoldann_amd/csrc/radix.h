// Block-wide (256 threads) rank selection on float order keys, shared by the
// threshold kernels.
#pragma once

#include "launch.h"

namespace mann {

struct RadixScratch {
  unsigned hist[256];
  unsigned sel, want, cnt;
  uint32_t buf[4096];
};

// rank-th smallest of 256 x 16 order keys held in registers (4-pass radix
// select over the whole block; all threads return the same key)
__device__ inline uint32_t block_radix_select(const uint32_t (&v)[16], unsigned rank, unsigned* hist,
                                       unsigned* s_sel, unsigned* s_want) {
  uint32_t prefix = 0, mask = 0;
  if (threadIdx.x == 0) *s_want = rank;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int p = 0; p < 16; ++p)
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
      unsigned excl = inc - tot, want = *s_want;
      if (excl < want && inc >= want) {
        unsigned c = excl, b = 4 * l;
        if (c + h0 >= want) b = 4 * l;
        else if (c + h0 + h1 >= want) { b = 4 * l + 1; c += h0; }
        else if (c + h0 + h1 + h2 >= want) { b = 4 * l + 2; c += h0 + h1; }
        else { b = 4 * l + 3; c += h0 + h1 + h2; }
        *s_sel = b;
        *s_want = want - c;
      }
    }
    __syncthreads();
    prefix |= *s_sel << shift;
    mask |= 255u << shift;
    __syncthreads();
  }
  return prefix;
}

// rank-th smallest (1-based) of m floats at row[0..m) as an order key.
// m <= 4096: one register radix select.  Larger m: U = rank-th smallest of the
// first 4096 values bounds the answer from above; one pass collects the values
// <= U into LDS (about rank * m / 4096 of them) and a second register select
// finishes; if more than 4096 qualify, a 4-pass select over all m runs instead.
__device__ inline uint32_t select_rank_key(const float* row, int m, int rank, RadixScratch& sm) {
  constexpr int PER = 16, CAPL = 256 * PER;
  uint32_t v[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    int i = threadIdx.x + 256 * p;
    v[p] = i < m ? f2ord(row[i]) : 0xffffffffu;
  }
  const unsigned r0 = (unsigned)min(rank, min(m, CAPL));
  uint32_t key = block_radix_select(v, r0, sm.hist, &sm.sel, &sm.want);
  if (m <= CAPL) return key;
  const uint32_t U = key;
  if (threadIdx.x == 0) sm.cnt = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < m; i += 256) {
    const uint32_t o = f2ord(row[i]);
    if (o <= U) {
      const unsigned pos = atomicAdd(&sm.cnt, 1u);
      if (pos < (unsigned)CAPL) sm.buf[pos] = o;
    }
  }
  __syncthreads();
  const unsigned cnt = sm.cnt;
  if (cnt <= (unsigned)CAPL && cnt >= (unsigned)rank) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const unsigned i = threadIdx.x + 256 * p;
      v[p] = i < cnt ? sm.buf[i] : 0xffffffffu;
    }
    __syncthreads();
    return block_radix_select(v, (unsigned)rank, sm.hist, &sm.sel, &sm.want);
  }
  // rare: too many values at or below U; 4-pass select over all of them
  uint32_t prefix = 0, mask = 0;
  if (threadIdx.x == 0) sm.want = (unsigned)rank;
  for (int shift = 24; shift >= 0; shift -= 8) {
    sm.hist[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += 256) {
      const uint32_t o = f2ord(row[i]);
      if ((o & mask) == prefix) atomicAdd(&sm.hist[(o >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int l = threadIdx.x;
      unsigned h0 = sm.hist[4 * l], h1 = sm.hist[4 * l + 1], h2 = sm.hist[4 * l + 2],
               h3 = sm.hist[4 * l + 3];
      unsigned tot = h0 + h1 + h2 + h3, inc = tot;
      for (int o = 1; o < 64; o <<= 1) {
        unsigned t = __shfl_up(inc, o);
        if (l >= o) inc += t;
      }
      unsigned excl = inc - tot, want = sm.want;
      if (excl < want && inc >= want) {
        unsigned c = excl, b = 4 * l;
        if (c + h0 >= want) b = 4 * l;
        else if (c + h0 + h1 >= want) { b = 4 * l + 1; c += h0; }
        else if (c + h0 + h1 + h2 >= want) { b = 4 * l + 2; c += h0 + h1; }
        else { b = 4 * l + 3; c += h0 + h1 + h2; }
        sm.sel = b;
        sm.want = want - c;
      }
    }
    __syncthreads();
    prefix |= sm.sel << shift;
    mask |= 255u << shift;
    __syncthreads();
  }
  return prefix;
}

}  // namespace mann
