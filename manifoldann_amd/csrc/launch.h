// Kernel launch interface shared by the kernel files and the C-ABI layer.
#pragma once

#include "common.h"

namespace mann {

enum ScanMode : int {
  SCAN_FILTER = 0,
  SCAN_DENSE = 1,
  SCAN_TOP1 = 2,
  SCAN_GATHER = 3,
  SCAN_SYM = 4  // FILTER over one triangle of a self-join, emitting both directions
};
enum ScanKind : int { SCAN_KIND_GRID = 0, SCAN_KIND_REFINE = 1, SCAN_KIND_ITEMS = 2 };

struct ScanItem {
  int qoff;       // offset into the query index list (or first query id)
  int nq;         // queries in this item (<= tile width)
  int row_begin;  // base rows [row_begin, row_end) of the (reordered) base
  int row_end;
  int region;     // candidate region of each query written by this item
};

struct ScanArgs {
  int cfg;            // tile configuration (index into the table of kernels_scan.hip)
  const bf16_t* xb;   // base rows, kp bf16 each (padded rows: zeros)
  const float* qnorm; // |q~|^2 per query (added to every value of its column)
  const float* col_norm;  // or per item column: [qoff + c] (IVF residual queries)
  int xcd_map;        // grid mode: 0 plain, 1 query-major per XCD, 2 part-major per XCD
  int xcd_band;       // grid mode 3: row partitions per band
  int cpol;           // cache policy of the operand loads: bit 0 rows, bit 1 queries
                      // non-temporal (streamed: inserted for early eviction in L2)
  int qtiles;         // grid mode: number of query tiles
  const float* xn;    // |x~|^2 per base row, +inf on padding
  const bf16_t* qb;   // query rows, kp bf16 each
  int kp;             // padded K (multiple of 32)
  int d;              // true dimension
  const int* qidx;    // query id list (nullptr: identity)
  int b_by_slot;      // qb row of column c is its slot qoff + c (IVF residual queries), not
                      // its query id
  // grid mode
  int nq;             // number of queries (or ...)
  const int* nq_dev;  // ... read from device memory when non-null
  int parts;          // row partitions per query tile
  int rows_per_part;  // multiple of the tile height
  int n_rows;         // rows scanned in grid mode: [0, n_rows)
  // persistent mode
  const ScanItem* items;
  const int* n_items_dev;
  int* item_counter;
  // FILTER / TOP1 outputs: query q, region r owns cand[(q * parts + r) * cap ..]
  const float* tau;   // per query threshold on a'
  unsigned* cand;     // pairs (float bits of a', row)
  int* cnt;           // [q * parts + r] emitted count (may exceed cap: overflow)
  int cap;            // capacity of one region
  // DENSE output: dense[q * ldd + row] (row: absolute index in the scanned rows),
  // or dense[q * ldd + row - region] when dense_rel (work items carrying their
  // list's first row as region); dense_t: dense[row * ldd + q] instead (the
  // scanned rows are the queries, the query columns the sample: coalesced)
  float* dense;
  int64_t ldd;
  int dense_rel;
  int dense_t;
  // dvec: ldd and the base are multiples of 4 floats and every row tile starts
  // 4-aligned relative to the stored origin (16-byte stores allowed)
  int dvec;
  // dense_bmin: dense[q * ldd + row / 32] = the minimum over each 32-row block
  // (row partitions start at multiples of 32)
  int dense_bmin;
  // dense_grp > 0: the minimum of every run of 4 consecutive rows is stored
  // instead of each value, 8 per 32-row block, dense[q * ldd + 8 b + slot]
  // (the slots of a block in lane order).  With dense_rel, blocks count from
  // the list's first row and only work items reaching at least dense_grp rows
  // past it do so (shorter lists store every value); otherwise blocks count
  // from row 0 and every item does (dense_grp = 1)
  int dense_grp;
  // TOP1 certification inputs: the queries' bf16 rounding errors and rounded
  // norms, their A16 |x|^2 (the FP32 norms form's error scales with the
  // uncentred norms), the rows' maxima of the first two and an upper bound of
  // the rows' uncentred norms
  const float* t1_qe;
  const float* t1_qr;
  const float* t1_qn2;
  float t1_ex, t1_rx, t1_ra;
  // Folded norms (FILTER / DENSE): the operands carry |x~|^2 and |q~|^2 in
  // spare K columns and the query side is -2 q~, so the MFMA accumulates S^
  // itself; xn and qnorm are not read
  int fold;
  // FP8 first pass (launch_scan8): E4M3 rows of kp bytes, and per query the
  // multiplier -2 * 2^-(s_x + s_q) of its accumulators
  const unsigned char* x8;
  const unsigned char* q8;
  const float* qmul;
  // candidate regions per query when not `parts` (SCAN_SYM: parts regions
  // for its direct emissions, then parts for the transposed ones)
  int rstride;
};

// Run-time tuning knobs (set through mann_tune_set; chosen by the autotuner).
struct Tuning {
  int scan_cfg = 5;          // exact-kNN scan tile config (table of kernels_scan.hip)
  int ivf_cfg = 5;           // IVF list-scan tile config
  int ivf_tau_cfg = 9;
  int ivf_rows_item = 2048;  // IVF work item: rows of one list chunk
  int ivf_first_rows = 0;    // IVF threshold: rows of the first list scanned (0: all)       // IVF first-list threshold scan tile config (few queries per list)
  int kmeans_cfg = 0;        // k-means assignment tile config
  int xcd_map = 3;           // grid order: 0 plain, 1 query-major per XCD, 2 part-major per
                             // XCD, 3 bands of xcd_band partitions x query tiles per XCD
  int xcd_band = 8;          // row partitions per band of grid order 3
  int scan_cpol = 0;         // operand loads non-temporal: bit 0 rows, bit 1 queries
  int target_groups = 2048;  // workgroups the row partitioning aims for
  int sample_rank = 1;       // multiplier of the emission target below
  int sample_grp = 1;        // exact kNN threshold sample: store 4-row minima (0: every value)
  int emit_per_k = 15;       // exact kNN: aim at emit_per_k * k_req first-pass emissions per query
  int fold_norms = 1;        // exact kNN: fold the norms into spare K columns when there are 6
  int shortlist_mult = 1;    // first re-rank batch = shortlist_mult * k_req (>= 32)
  int tau_rows = 1024;       // IVF: exact rows sampled per query for the emission threshold
  int ivf_tau_mode = 1;      // IVF threshold: 0 exact FP32 row sample, 1 bf16 scan of the first list
  int ivf_route_cert = 1;    // IVF: TopKRouting from a bf16 first pass over the centroids
                             // with FP32 re-rank of the uncertain ones (0: every routing
                             // distance in FP32)
  int ivf_route_cfg = 0;     // tile config of that first pass
  int ivf_tau_grp = 1;       // IVF first-list scan: store 4-row minima of lists of at least
                             // max(64, 4 k) rows (0: every value)
  int sample_rows = 2048;    // exact kNN: stride-sample rows (fixed when a dataset is created)
  int scan8_cfg = 1;         // FP8 first-pass tile config (table of kernels_fp8.hip)
  int fp8_emit_per_k = 2;    // FP8 first pass: bf16 sample rank aimed at fp8_emit_per_k * k_req
                             // rows, then widened by the FP8 band
  int sym_join = 0;          // kNN graph with folded norms: scan one triangle of the self-join
                             // (switched off: see mann_tune_set)
};
// the calling thread's snapshot of the knobs (refreshed by tuning_refresh at
// every API entry)
Tuning& tuning();
void tuning_refresh();
Tuning tuning_shared();             // a copy of the process-wide table
void tuning_publish(const Tuning& t);  // replace it (and refresh this thread)

int scan_num_cfg();
int scan_tile_rows(int cfg);
int scan_tile_queries(int cfg);
void launch_scan(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st);

void launch_prep_rows(const float* src, int64_t ld, int64_t n, int64_t n_pad, int d, int d16,
                      int kp, int normalize, float* x32, bf16_t* xb, float* xn, float* err,
                      float* nrm, unsigned* max_err, unsigned* max_nrm, hipStream_t st,
                      const float* center = nullptr, const int* center_row = nullptr,
                      int64_t center_ld = 0);
// atomicMax into *out (float bits, start at 0) of an upper bound of the largest
// norm sqrt(n2[i]), n2 holding A16 FP32 squared norms of d-vectors
void launch_norm_bound_max(const float* n2, int64_t n, int d, unsigned* out, hipStream_t st);
// FP32 mean of n rows (FP64 sums); tmp holds d doubles
void launch_mean_rows(const float* x, int64_t ld, int64_t n, int d, double* tmp, float* mean,
                      hipStream_t st);
void launch_a16_norms(const float* x32, int64_t n, int d16, float* out, hipStream_t st);
// folded-norm operands (see mann_dataset::fold): base rows in place, queries
// into out (nq_pad rows)
void launch_fold_base(bf16_t* xb, const float* xn, int64_t n, int kp, int d, hipStream_t st);
void launch_fold_queries(const bf16_t* qb, const float* qn, int64_t nq, int64_t nq_pad, int kp,
                         int d, bf16_t* out, hipStream_t st);

// ---- selection / exact re-rank -------------------------------------------------

struct TauArgs {
  const float* dense;  // [nq][m]
  int64_t ldd;
  int m;               // sample size
  int rank;            // 1-based rank of the threshold value
  float* tau;          // out
  int nq;
};
void launch_tau_from_sample(const TauArgs& a, hipStream_t st);

// Search statistics are accumulated by atomics spread over kStatSlots slots of
// 4 counters each (query slot % kStatSlots), so that same-address atomics do
// not serialize the selection; readers sum the slots.
constexpr int kStatSlots = 64;

struct SelectArgs {
  // queries processed: blockIdx -> qlist[blockIdx] (or identity), count on device
  const int* qlist;
  const int* nq_dev;
  int nq;
  // candidates
  const unsigned* cand;
  const int* cnt;
  int cap;             // capacity of one region
  int parts;           // regions per query
  const float* tau;
  // exact operands
  const float* x32;    // base rows (reordered space for IVF), d16 floats each
  const float* q32;    // query rows, d16 floats
  const float* xn2;    // A16 |x|^2 per base row (cosine / norms form)
  const float* qn2;    // A16 |q|^2 per query
  int d, d16;
  int metric;          // EUCLIDEAN / SQEUCLIDEAN / COSINE / NORMS_FORM
  const int* perm;     // base row -> original point id (nullptr: identity)
  // certification inputs
  const float* qe;     // |q - q~|
  const float* qr;     // |q~|
  float ex, rx;        // dataset bounds
  // request
  int k;               // neighbours returned per query
  int k_req;           // neighbours selected before self removal (k or k+1)
  int exclude_self;    // drop id == query id (kNN graph)
  int k_eff;           // min(k_req, rows available)
  const int* k_eff_q;  // per-query k_eff (IVF: rows in the probed lists), or null
  int shortlist;       // first re-rank batch (oversampling, >= k_req)
  // outputs
  int64_t* out_ids;    // [nq][k] 1-based, 0 = empty
  float* out_dist;     // [nq][k]
  int* out_count;      // optional: results per query (after self removal)
  // refinement
  int* flag_count;     // queries that failed certification
  int* flag_list;
  float* tau_new;      // refined threshold for a second pass
  int* stats;          // [kStatSlots][4]: reranked, overflowed queries, emitted (see below)
  // First pass in another domain (FP8 candidates re-scored on bf16): tau is
  // then that pass's threshold, certified against the exclusion threshold of
  // its own bounds (cert_*: |q - q~|, |q~| per query, dataset maxima, error
  // coefficient G / (|q~| + |x~|)^2); tau_new is in that domain too unless
  // tau_new_bf16 (a second pass on bf16 follows).
  // dropmin: per query, a lower bound of the re-scored values left out of the
  // candidates (they must lie above the bf16 exclusion threshold).
  const float* cert_qe;
  const float* cert_qr;
  float cert_ex, cert_rx;
  double cert_g;
  const float* dropmin;
  int no_emit_stat;    // emitted counts are recorded by the re-score instead
  int tau_new_bf16;
  // a second-pass threshold from the sample (null: none): used when it is
  // below the one derived from the k-th distance found, or when the first
  // pass left fewer than k candidates
  const float* tau_alt;
};
enum ExtraMetric : int { NORMS_FORM = 3 };
// one wavefront per query; nq_max bounds the device-side query count
void launch_select_rerank(const SelectArgs& a, int nq_max, hipStream_t st);

struct ExactArgs {
  const int* qlist;
  const int* nq_dev;
  const float* x32;
  const float* q32;
  const float* xn2;
  const float* qn2;
  int d, d16, metric;
  const int* perm;
  // row ranges of query q: ranges[q * range_stride + i], i < range_cnt[q];
  // range_cnt == nullptr -> the single range [0, n_rows)
  const int* range_cnt;
  const int2* ranges;
  int range_stride;
  int n_rows;
  int k, k_req, exclude_self;
  int64_t* out_ids;
  float* out_dist;
  int* out_count;
  // segment partials: slots [slot_base, slot_base + nslots) of qlist, nseg
  // segments each (seg_rows rows inside one range), kr keys per segment
  int slot_base;
  int nseg;
  int seg_rows;       // a multiple of 256, at most exact_seg_rows()
  uint64_t* partial;  // [nslots][nseg][min(k_req, exact_kmax())]
  // per slot of the launch: the smallest kr-th key any segment has found (an
  // upper bound of the query's kr-th key; reset by launch_exact_topk)
  unsigned long long* bound;
};
// Two launches: per-segment exact top-k, then the per-query merge.
void launch_exact_topk(const ExactArgs& a, int nslots, hipStream_t st);
int exact_seg_rows();
int exact_merge_cap();  // keys the merge sorts at once per query

}  // namespace mann
