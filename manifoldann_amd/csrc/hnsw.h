// IVF-HNSW terminals (kernels_hnsw.hip): per-list HNSW graphs built and
// searched on the GPU, one wavefront per list (build) or per (query, list)
// pair (search).
#pragma once

#include "common.h"

namespace mann {

// (distance, local id) of a graph search
struct HnCand {
  float d;
  int id;
};

constexpr int HN_EF_MAX = 512;   // ef_construction and max(ef_search, k)
constexpr int HN_M_MAX = 63;     // M (lists hold M + 1 slots)

// Graph storage over the list-contiguous rows: base-layer lists nbr0[row][cap]
// and deg0[row]; upper layer l >= 1 holds the rows of level >= 1 in compact
// slots: nbru[(l - 1) n1 + upslot[row]][cap], degu likewise.  Ids in the
// lists are local (row - first row of the list).
struct HnGraph {
  int* nbr0;
  int* deg0;
  const int* upslot;
  int* nbru;
  int* degu;
  int64_t n1;
  int cap;
};

struct HnswBuildArgs {
  const float* x32;     // reordered base rows, d16 floats each
  int d16;
  const int* list_off;  // nb + 1
  const int* lorder;    // lists, largest first
  int nb;
  const int* level;     // per row
  HnGraph g;
  int M, ef, policy;    // policy: 0 heuristic, 1 diversified
  HnCand* pend_g;       // list b: pend_g + list_off[b] + b (|list| + 1 entries)
  unsigned* vis_g;      // list b: vis_g + (list_off[b] >> 5) + b words
  int* entry;           // per bucket: entry point (local id), -1 when empty
  int* maxl;            // per bucket: top layer, -1 when empty
};

struct HnswSearchArgs {
  const float* x32;
  const float* q32;     // queries, d16 floats each
  int d16;
  const int* list_off;
  const int* probes;    // [nq][nprobe] buckets (0-based) in probe order
  const int* pcount;
  int nq, nprobe;
  const int* entry;
  const int* maxl;
  HnGraph g;
  int k, ef;
  int pend_lds;         // pending-heap entries held in LDS before the spill
  int vis_words;        // visited words in LDS (0: the global bitmap)
  HnCand* pend_g;       // per workgroup: pend_g + blockIdx.x * pend_stride
  int64_t pend_stride;  // >= largest list + 1
  unsigned* vis_g;      // per workgroup: vis_g + blockIdx.x * vis_stride
  int64_t vis_stride;
  int* out_row;         // [nq][nprobe][k] reordered rows
  float* out_d;
  int* out_cnt;         // [nq][nprobe]
  unsigned long long* ndist;  // distances computed, added per pair (cleared by the caller)
};

struct HnswMergeArgs {
  const int* out_row;
  const float* out_d;
  const int* out_cnt;
  const int* probes;
  const int* pcount;
  int nq, nprobe, k;
  int merge;            // 0 SimpleMerge, 1 DisjointMerge
  int probe_order;      // 1: lists in probe order (single query), 0: bucket order
  const int* perm;      // reordered row -> point id (0-based)
  HnCand* heap;         // [nq][k] scratch
  int* scratch;         // [nq][2 nprobe] list order and cursors
  int64_t* ids;         // [nq][k] 1-based, 0 padded
  float* dist;
};

void launch_hnsw_build(const HnswBuildArgs& a, int grid, hipStream_t st);
// grid workgroups of one wavefront loop over the nq * nprobe pairs
void launch_hnsw_search(const HnswSearchArgs& a, int grid, hipStream_t st);
void launch_hnsw_merge(const HnswMergeArgs& a, hipStream_t st);
size_t hnsw_search_lds(int ef, int pend_lds, int vis_words);

}  // namespace mann
