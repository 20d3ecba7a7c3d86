// IVF kernel interface (routing, probe selection, work planning, thresholds).
#pragma once

#include "launch.h"

namespace mann {

struct IvfTauArgs {
  int nq, nprobe, k, d, d16, metric;
  int emit_all_below;       // queries probing at most this many rows emit everything
  int sample_rows;          // exact sample size (at most ivf_tau_rows_max())
  const int* probes;        // [nq][nprobe] buckets in probe order
  const int* pcount;        // probed (non-empty) lists per query
  const int* prows;         // rows in the probed lists per query
  const int* list_off;      // CSR offsets of the lists in the reordered base
  const float* x32;         // reordered base rows
  const float* xn2;
  const float* q32;
  const float* qn2;
  const float* qe;          // certification inputs of the queries
  const float* qr;
  float ex, rx;             // dataset bounds
  float* tau;               // out: emission threshold
  int* k_eff_q;             // out: min(k, rows probed)
  // block order: one block per (list, query) entry of the inverted probe
  // lists; the block of a query's first probed list does the work, so
  // consecutive blocks share their sample rows in L2
  const int* qlist;
  const int* qlist_list;
  const int* n_entries;
  int64_t max_entries;
};

int route_max();
int ivf_tau_rows_max();
// metric: EUCLIDEAN (norms form, the default), SQEUCLIDEAN or COSINE
void launch_route_dist(const float* q32, const float* qn2, int nq, const float* c32,
                       const float* cn2, int nlist, int d16, float* R, hipStream_t st,
                       int metric = EUCLIDEAN);
// mass > 0: probability-mass probing over negated logits (nprobe <= 64,
// avail <= 1024), see kernels_ivf.hip
void launch_probe_select(const float* R, int nq, int nlist, int avail, int nprobe, float mass,
                         const int* list_off, int* probes, int* pcount, int* prows,
                         hipStream_t st);
// TopKRouting over the dense bf16 first pass of the first `avail` centroids
// (certified: the probes equal those of the exact routing distances), queries
// [q0, q0 + nq) of the batch; dense row qi holds query q0 + qi.
struct RouteSelArgs {
  const float* dense;
  int64_t ldd;
  int avail, nprobe, q0, nq;
  const float* c32;  // centroids, d16 floats each, and their A16 |c|^2
  const float* cn2;
  const float* q32;  // queries, d16 floats each, and their A16 |q|^2
  const float* qn2;
  int d16, d;
  int sq;            // 1: SQEUCLIDEAN (no square root)
  const float* qe;   // query bf16 rounding error and rounded norm
  const float* qr;
  float ex, rx;      // centroid maxima of the same
  float ra;          // upper bound of the centroids' norms
  const int* list_off;
  int* probes;
  int* pcount;
  int* prows;
};
int route_select_cap();  // largest avail
void launch_route_select(const RouteSelArgs& a, hipStream_t st);
void launch_ivf_plan(const int* probes, const int* pcount, int nq, int nprobe, int nb,
                     const int* list_off, const int* lorder, int bn, int rows_item, int* qcount,
                     int* qfill, int* qoff, int* ioff, int* qlist, int* qlist_list,
                     int* slot_of, ScanItem* items, int* n_items, int* n_entries,
                     int* first_cnt, hipStream_t st);
// Work items of the threshold scan: for every list, the queries that probe it
// first (the leading first_cnt entries of its block) against all its rows;
// region = the list's first row.  n_items written on the device.
void launch_first_items(const int* first_cnt, const int* qoff, int nb, const int* list_off,
                        const int* lorder, int bn, int rows_item, int max_rows, int* ioff,
                        ScanItem* items, int* n_items, hipStream_t st);
// Emission threshold from the first-list scan: S^ values dense[q][0..|l1|)
struct IvfDenseTauArgs {
  int nq, nprobe, k, d, metric;
  int emit_all_below;
  const int* probes;
  const int* pcount;
  const int* prows;
  const int* list_off;
  const float* dense;
  int64_t ldd;
  const float* qe;
  const float* qr;
  float ex, rx;
  float* tau;
  int* k_eff_q;
  int max_rows;  // rows of the first list the DENSE scan covered (0: all)
  int grp_min;   // lists of at least this many covered rows hold 4-row minima (0: never)
};
void launch_ivf_tau_dense(const IvfDenseTauArgs& a, hipStream_t st);
// single-query items over the probed lists of the flagged queries
void launch_refine_items(const int* flist, const int* nflag, int nflag_host, int nprobe,
                         const int* probes, const int* pcount, const int* slot_of,
                         const int* list_off, int rows_item, ScanItem* items, int* n_items,
                         hipStream_t st);
// bf16 residual rows q - c_list for the entries of the inverted probe lists
void launch_residual_queries(const float* q32, const float* c32, int d, int d16, int kp,
                             int normalize, const int* qlist, const int* qlist_list, const int* n_dev,
                             int64_t max_entries, bf16_t* qres, float* rnorm, unsigned* qe_max,
                             unsigned* qr_max, hipStream_t st);
void launch_ivf_tau(const IvfTauArgs& a, hipStream_t st);
void launch_probe_ranges(const int* probes, const int* pcount, int nq, int nprobe,
                         const int* list_off, int2* ranges, hipStream_t st);

// Merge orders other than the batched SimpleMerge order (kernels_merge.hip).
enum MergeKind : int {
  MERGE_BATCHED = 0,     // (dist, bucket, id): the selection's own order
  MERGE_HEAP_ROWS = 1,   // IVF-Flat bounded heap over every probed row, probe order
  MERGE_HEAP_LISTS = 2,  // DisjointMerge: heap over each list's top-k
  MERGE_PROBE = 3        // single-query SimpleMerge: (dist, probe rank, id)
};
constexpr int ROUTE_MAX_MERGE = 4096;  // probe lists a replay plan can order
struct MergeArgs {
  const int64_t* ti;      // [nq][k + 1] the search's top-(k + 1), 1-based ids
  const float* td;
  int nq, k, kind;
  const int* probes;      // [nq][nprobe] the search's probed lists (probe order)
  const int* pcount;
  int nprobe;
  const int* bucket_of;   // point id (0-based) -> its bucket
  const int* perm;        // reordered row -> point id
  const int* list_off;
  int64_t* out_ids;       // [nq][k]
  float* out_dist;
  int* tied_list;         // out: queries whose k-th place is tied
  int* tied_count;
};
struct ReplayArgs {
  const int* tied;        // tied queries; this batch is tied[slot0 ..]
  int slot0;
  int nprobe;
  int bucket_order;       // replay lists in bucket order (else probe order)
  const int* probes;
  const int* pcount;
  const int* list_off;
  int* seq;               // [slot][nprobe] replay order of the lists
  int* seqoff;            // [slot][nprobe + 1] row offsets in that order
  float* dist;            // exact distances of the probed rows in replay order
  const int64_t* dbase;   // [slot] start of its rows in dist
  const float* x32;
  const float* q32;
  const float* xn2;
  const float* qn2;
  int d16, metric;
};
// Untied queries are written (their k rows re-sorted by the kind's order);
// tied ones are appended to tied_list.
void launch_merge_fix(const MergeArgs& a, hipStream_t st);
// nslots tied queries from a.slot0; max_rows = most probed rows of any of them
void launch_replay(const ReplayArgs& a, const MergeArgs& m, int nslots, int max_rows,
                   hipStream_t st);

}  // namespace mann
