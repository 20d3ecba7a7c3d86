// Merge orders of the IVF search other than the batched SimpleMerge order
// (dist, bucket, id) that the certified selection produces directly:
//   MERGE_HEAP_ROWS   query(::IVFFlatIndex) (ivf_flat.jl:138-167): every row of
//                     the probed lists, lists in probe order and ascending id
//                     inside a list, pushed through BoundedMaxHeap
//                     (neighbor_heaps.jl:84-100), then sorted by (dist, id)
//                     (to_sorted_vector, :138-141)
//   MERGE_HEAP_LISTS  merge_results(::DisjointMerge) (merge.jl:150-169): each
//                     probed list's top-min(k, |list|) by (dist, id) (the
//                     BruteForce child, bruteforce.jl:65) pushed through the
//                     same heap; lists in bucket order for a batch
//                     (query.jl:382-396) or probe order for one query
//                     (query.jl:84-103)
//   MERGE_PROBE       merge_results(::SimpleMerge) (merge.jl:79-124) of one
//                     query: a stable sort by dist of the per-list results
//                     concatenated in probe order, i.e. (dist, probe rank, id)
//
// Input is the search's top-(k+1) in (dist, bucket, id) order.  Where the k-th
// and (k+1)-th distances differ, every order keeps the same k rows and only the
// order among equal distances changes: merge_fix_kernel re-sorts them.  A query
// whose k-th place is tied is replayed: replay_dist_kernel computes the exact
// distance of every probed row (a grid over the tied queries' rows, in replay
// order), then one wavefront per query runs the heap (or, for MERGE_PROBE,
// takes the first tied rows in probe order).
#include "bounds.h"
#include "ivf.h"
#include "radix.h"

namespace mann {

namespace {

// Bitonic sort (ascending) of P keys (a power of two, at most 512) in LDS by
// one wavefront.
__device__ void wave_lds_sort(uint64_t* s, int P, int lane) {
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t x = s[i], y = s[ixj];
          if ((x > y) == ((i & k) == 0)) {
            s[i] = y;
            s[ixj] = x;
          }
        }
      }
      wave_fence_sync();
    }
}

__device__ __forceinline__ int pow2_at_least(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Position of bucket b in the query's probe order (pc if absent).
__device__ __forceinline__ int probe_rank(const int* pr, int pc, int b) {
  for (int i = 0; i < pc; ++i)
    if (pr[i] == b) return i;
  return pc;
}

// Sort key of entry r of a query's (k+1) list.  HEAP kinds: (dist, id);
// PROBE: (dist, probe rank, r) -- r orders entries of one list by id, since the
// input lists each list's rows in ascending id.
__device__ __forceinline__ uint64_t entry_key(const MergeArgs& a, int q, int64_t id, float dist,
                                              int r) {
  if (a.kind == MERGE_PROBE) {
    const int b = a.bucket_of[id - 1];
    const int rk = probe_rank(a.probes + (int64_t)q * a.nprobe, a.pcount[q], b);
    return ((uint64_t)f2ord(dist) << 32) | ((uint64_t)rk << 20) | (uint64_t)r;
  }
  return ((uint64_t)f2ord(dist) << 32) | (uint64_t)(id - 1);
}

// Sort the first m entries of query q's list by its kind's key and write them
// to out[0..m); keys is the wavefront's LDS (256 keys).
__device__ void write_sorted(const MergeArgs& a, int q, int m, uint64_t* keys, int lane) {
  const int64_t* ri = a.ti + (int64_t)q * (a.k + 1);
  const float* rd = a.td + (int64_t)q * (a.k + 1);
  const int P = pow2_at_least(m > 0 ? m : 1);
  for (int r = lane; r < P; r += 64) keys[r] = r < m ? entry_key(a, q, ri[r], rd[r], r) : ~0ull;
  wave_fence_sync();
  wave_lds_sort(keys, P, lane);
  int64_t* oi = a.out_ids + (int64_t)q * a.k;
  float* od = a.out_dist + (int64_t)q * a.k;
  for (int r = lane; r < m; r += 64) {
    const uint64_t kv = keys[r];
    const float dist = ord2f((uint32_t)(kv >> 32));
    int64_t id;
    if (a.kind == MERGE_PROBE) id = ri[(int)(kv & 0xfffffu)];
    else id = (int64_t)(kv & 0xffffffffu) + 1;
    oi[r] = id;
    od[r] = dist;
  }
  wave_fence_sync();
}

// One wavefront per query: untied queries are written here, tied ones listed.
__global__ __launch_bounds__(256) void merge_fix_kernel(MergeArgs a) {
  __shared__ uint64_t keys_all[4][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = blockIdx.x * 4 + w;
  if (q >= a.nq) return;
  const int K1 = a.k + 1;
  const int64_t* ri = a.ti + (int64_t)q * K1;
  const float* rd = a.td + (int64_t)q * K1;
  int have = 0;  // valid entries form a prefix (the selection pads with id 0)
  for (int r = lane; r < K1; r += 64) have += ri[r] != 0 ? 1 : 0;
  for (int o = 32; o >= 1; o >>= 1) have += __shfl_xor(have, o);
  if (have == K1 && rd[a.k] == rd[a.k - 1]) {
    if (lane == 0) a.tied_list[atomicAdd(a.tied_count, 1)] = q;
    return;
  }
  const int m = have < a.k ? have : a.k;
  write_sorted(a, q, m, keys_all[w], lane);
  int64_t* oi = a.out_ids + (int64_t)q * a.k;
  float* od = a.out_dist + (int64_t)q * a.k;
  for (int r = m + lane; r < a.k; r += 64) {
    oi[r] = 0;
    od[r] = INFINITY;
  }
}

// Per tied query (slot s of the batch): its probed lists in replay order,
// seq[s][0..pc), and the exclusive prefix of their sizes, off[s][0..pc].
__global__ __launch_bounds__(256) void replay_plan_kernel(ReplayArgs a) {
  __shared__ int sb[ROUTE_MAX_MERGE];
  const int s = blockIdx.x;
  const int q = a.tied[a.slot0 + s];
  const int pc = a.pcount[q];
  const int* pr = a.probes + (int64_t)q * a.nprobe;
  int* seq = a.seq + (int64_t)s * a.nprobe;
  int* off = a.seqoff + (int64_t)s * (a.nprobe + 1);
  if (a.bucket_order) {
    const int P = pow2_at_least(pc > 0 ? pc : 1);
    for (int i = threadIdx.x; i < P; i += blockDim.x) sb[i] = i < pc ? pr[i] : 0x7fffffff;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
          const int ixj = i ^ j;
          if (ixj > i) {
            const int x = sb[i], y = sb[ixj];
            if ((x > y) == ((i & k) == 0)) {
              sb[i] = y;
              sb[ixj] = x;
            }
          }
        }
        __syncthreads();
      }
    for (int i = threadIdx.x; i < pc; i += blockDim.x) seq[i] = sb[i];
  } else {
    for (int i = threadIdx.x; i < pc; i += blockDim.x) seq[i] = pr[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < pc; ++i) {
      off[i] = acc;
      const int b = seq[i];
      acc += a.list_off[b + 1] - a.list_off[b];
    }
    off[pc] = acc;
  }
}

// Exact distance of every probed row of the tied queries, in replay order:
// block (s, y) covers positions [y RP_CH, (y + 1) RP_CH) of slot s; a 16-lane
// group computes one row (A16 order, the selection's exact arithmetic).
constexpr int RP_CH = 1024;
__global__ __launch_bounds__(256) void replay_dist_kernel(ReplayArgs a) {
  const int s = blockIdx.x;
  const int q = a.tied[a.slot0 + s];
  const int pc = a.pcount[q];
  const int* seq = a.seq + (int64_t)s * a.nprobe;
  const int* off = a.seqoff + (int64_t)s * (a.nprobe + 1);
  const int total = off[pc];
  const int p0 = blockIdx.y * RP_CH;
  if (p0 >= total || pc == 0) return;
  const int p1 = min(total, p0 + RP_CH);
  float* dst = a.dist + a.dbase[s];
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15;
  for (int p = p0 + g; p < p1; p += 16) {
    int lo = 0, hi = pc - 1;  // the last list with off[i] <= p
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= p) lo = mid;
      else hi = mid - 1;
    }
    const int r = a.list_off[seq[lo]] + (p - off[lo]);
    const float dv = exact_metric(a.x32, a.q32, a.xn2, a.qn2, a.d16, a.metric, r, q, j);
    if (j == 0) dst[p] = dv;
  }
}

struct Heap {
  float* hd;   // [cap] distances, 0-based array of the 1-based heap
  int* hid;    // [cap] 1-based ids
  int len;
  int cap;
};

// BoundedMaxHeap push! (neighbor_heaps.jl:84-100, sift-up :156-168,
// sift-down :170-190); run by one lane.  Returns false when rejected.
__device__ bool heap_push(Heap& h, float dist, int id) {
  if (h.len < h.cap) {
    int i = ++h.len;
    h.hd[i - 1] = dist;
    h.hid[i - 1] = id;
    while (i > 1) {
      const int p = i >> 1;
      if (h.hd[i - 1] > h.hd[p - 1]) {
        const float td = h.hd[i - 1];
        h.hd[i - 1] = h.hd[p - 1];
        h.hd[p - 1] = td;
        const int ti = h.hid[i - 1];
        h.hid[i - 1] = h.hid[p - 1];
        h.hid[p - 1] = ti;
        i = p;
      } else {
        break;
      }
    }
    return true;
  }
  if (!(dist < h.hd[0])) return false;
  h.hd[0] = dist;
  h.hid[0] = id;
  int i = 1;
  for (;;) {
    const int l = i << 1, rg = l + 1;
    int big = i;
    if (l <= h.len && h.hd[l - 1] > h.hd[big - 1]) big = l;
    if (rg <= h.len && h.hd[rg - 1] > h.hd[big - 1]) big = rg;
    if (big == i) break;
    const float td = h.hd[i - 1];
    h.hd[i - 1] = h.hd[big - 1];
    h.hd[big - 1] = td;
    const int ti = h.hid[i - 1];
    h.hid[i - 1] = h.hid[big - 1];
    h.hid[big - 1] = ti;
    i = big;
  }
  return true;
}

// One wavefront per tied query over the distances of replay_dist_kernel.
__global__ __launch_bounds__(64) void replay_heap_kernel(ReplayArgs a, MergeArgs m) {
  __shared__ float hd[256];
  __shared__ int hid[256];
  __shared__ float cd[64];
  __shared__ int cid[64];
  __shared__ uint64_t ck[256];
  __shared__ uint32_t rbuf[256];
  __shared__ unsigned rhist[256];
  __shared__ int s_full, s_len;
  __shared__ float s_root;
  const int s = blockIdx.x, lane = threadIdx.x;
  const int q = a.tied[a.slot0 + s];
  const int pc = a.pcount[q];
  const int* seq = a.seq + (int64_t)s * a.nprobe;
  const int* off = a.seqoff + (int64_t)s * (a.nprobe + 1);
  const float* dist = a.dist + a.dbase[s];
  const int k = m.k;
  int64_t* oi = m.out_ids + (int64_t)q * k;
  float* od = m.out_dist + (int64_t)q * k;

  if (m.kind == MERGE_PROBE) {
    // the entries below the tied value D_k are the k-list's own; the rest of
    // the k places go to the first rows at D_k in probe order
    const int64_t* ri = m.ti + (int64_t)q * (k + 1);
    const float* rd = m.td + (int64_t)q * (k + 1);
    const float dk = rd[k - 1];
    int clt = 0;
    for (int r = lane; r < k; r += 64) clt += rd[r] < dk ? 1 : 0;
    for (int o = 32; o >= 1; o >>= 1) clt += __shfl_xor(clt, o);
    write_sorted(m, q, clt, ck, lane);
    int filled = clt;
    for (int i = 0; i < pc && filled < k; ++i) {
      const int b = seq[i];
      const int r0 = m.list_off[b];
      const int nl = off[i + 1] - off[i];
      for (int t = 0; t < nl && filled < k; t += 64) {
        const bool in = t + lane < nl;
        const bool eq = in && dist[off[i] + t + lane] == dk;
        const unsigned long long bal = __ballot(eq);
        const int pos = filled + __popcll(bal & ((1ull << lane) - 1ull));
        if (eq && pos < k) {
          oi[pos] = (int64_t)m.perm[r0 + t + lane] + 1;
          od[pos] = dk;
        }
        filled += __popcll(bal);
      }
    }
    // (at least k - clt rows sit at D_k: the (k+1)-list held k + 1 - clt)
    return;
  }

  Heap h{hd, hid, 0, k};  // a (k+1)-th row exists, so the capacity is k
  if (lane == 0) {
    s_full = 0;
    s_root = INFINITY;
  }
  wave_fence_sync();
  for (int i = 0; i < pc; ++i) {
    const int b = seq[i];
    const int r0 = m.list_off[b];
    const int nl = off[i + 1] - off[i];
    const float* ld = dist + off[i];
    if (m.kind == MERGE_HEAP_ROWS) {
      for (int t = 0; t < nl; t += 64) {
        const bool in = t + lane < nl;
        const float dv = in ? ld[t + lane] : INFINITY;
        cd[lane] = dv;
        cid[lane] = in ? m.perm[r0 + t + lane] + 1 : 0;
        // the root only decreases: a row at or above it now never enters
        const bool maybe = in && !(s_full && dv >= s_root);
        const unsigned long long bal = __ballot(maybe);
        wave_fence_sync();
        if (lane == 0) {
          for (unsigned long long mk = bal; mk != 0; mk &= mk - 1) {
            const int e = __builtin_ctzll(mk);
            heap_push(h, cd[e], cid[e]);
          }
          s_full = h.len == h.cap;
          s_root = hd[0];
        }
        wave_fence_sync();
      }
      continue;
    }
    // MERGE_HEAP_LISTS: the list's kk = min(k, nl) smallest (dist, id), pushed
    // in that order; once one is rejected the rest are too (ascending order,
    // non-increasing root).  With a full heap only rows below the root count.
    if (nl == 0) continue;
    const int kk = nl < k ? nl : k;
    const bool full = s_full != 0;
    const float root = s_root;
    if (full) {
      int below = 0;
      for (int t = lane; t < nl; t += 64) below += ld[t] < root ? 1 : 0;
      for (int o = 32; o >= 1; o >>= 1) below += __shfl_xor(below, o);
      if (below == 0) continue;
    }
    // D* = the kk-th smallest distance of the list; take every row below it
    // and the first rows (ascending id) at it until kk are taken
    const float dstar = ord2f(wave_rank_key(ld, nl, kk, rbuf, rhist, lane));
    int clt = 0;
    for (int t = lane; t < nl; t += 64) clt += ld[t] < dstar ? 1 : 0;
    for (int o = 32; o >= 1; o >>= 1) clt += __shfl_xor(clt, o);
    int taken = 0, eqs = 0;
    for (int t = 0; t < nl && taken < kk; t += 64) {
      const bool in = t + lane < nl;
      const float dv = in ? ld[t + lane] : INFINITY;
      const bool lt = in && dv < dstar;
      const bool eq = in && dv == dstar;
      const unsigned long long beq = __ballot(eq);
      const int eqpos = eqs + __popcll(beq & ((1ull << lane) - 1ull));
      const bool take = lt || (eq && eqpos < kk - clt);
      const unsigned long long bt = __ballot(take);
      const int pos = taken + __popcll(bt & ((1ull << lane) - 1ull));
      if (take) ck[pos] = ((uint64_t)f2ord(dv) << 32) | (uint64_t)(t + lane);
      taken += __popcll(bt);
      eqs += __popcll(beq);
    }
    const int P = pow2_at_least(kk);
    for (int r = kk + lane; r < P; r += 64) ck[r] = ~0ull;
    wave_fence_sync();
    wave_lds_sort(ck, P, lane);
    if (lane == 0) {
      for (int r = 0; r < kk; ++r) {
        const uint64_t kv = ck[r];
        const int t = (int)(kv & 0xffffffffu);
        if (!heap_push(h, ord2f((uint32_t)(kv >> 32)), m.perm[r0 + t] + 1)) break;
      }
      s_full = h.len == h.cap;
      s_root = hd[0];
    }
    wave_fence_sync();
  }
  if (lane == 0) s_len = h.len;
  wave_fence_sync();
  // to_sorted_vector: (dist, id)
  const int len = s_len;
  const int P = pow2_at_least(len);
  for (int r = lane; r < P; r += 64)
    ck[r] = r < len ? (((uint64_t)f2ord(hd[r]) << 32) | (uint64_t)(uint32_t)(hid[r] - 1)) : ~0ull;
  wave_fence_sync();
  wave_lds_sort(ck, P, lane);
  for (int r = lane; r < k; r += 64) {
    const uint64_t kv = ck[r < P ? r : 0];
    oi[r] = r < len ? (int64_t)(kv & 0xffffffffu) + 1 : 0;
    od[r] = r < len ? ord2f((uint32_t)(kv >> 32)) : INFINITY;
  }
}

}  // namespace

void launch_merge_fix(const MergeArgs& a, hipStream_t st) {
  if (a.nq <= 0 || a.k <= 0) return;
  hipLaunchKernelGGL(merge_fix_kernel, dim3((unsigned)ceil_div(a.nq, 4)), dim3(256), 0, st, a);
  MANN_HIP(hipGetLastError());
}

void launch_replay(const ReplayArgs& a, const MergeArgs& m, int nslots, int max_rows,
                   hipStream_t st) {
  if (nslots <= 0) return;
  hipLaunchKernelGGL(replay_plan_kernel, dim3(nslots), dim3(256), 0, st, a);
  if (max_rows > 0)
    hipLaunchKernelGGL(replay_dist_kernel, dim3(nslots, (unsigned)ceil_div(max_rows, RP_CH)),
                       dim3(256), 0, st, a);
  hipLaunchKernelGGL(replay_heap_kernel, dim3(nslots), dim3(64), 0, st, a, m);
  MANN_HIP(hipGetLastError());
}

}  // namespace mann
