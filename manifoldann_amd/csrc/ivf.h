// IVF kernel interface (routing, probe selection, work planning, thresholds).
#pragma once

#include "launch.h"

namespace mann {

struct IvfTauArgs {
  int nq, nprobe, k, d, d16, metric;
  int emit_all_below;       // queries probing at most this many rows emit everything
  const int* probes;        // [nq][nprobe] buckets in probe order
  const int* pcount;        // probed (non-empty) lists per query
  const int* prows;         // rows in the probed lists per query
  const int* list_off;      // CSR offsets of the lists in the reordered base
  const float* x32;         // reordered base rows
  const float* xn2;
  const float* q32;
  const float* qn2;
  const float* qn;          // certification inputs of the queries
  const float* qe;
  const float* qr;
  float ex, rx;             // dataset bounds
  float* tau;               // out: emission threshold
  int* k_eff_q;             // out: min(k, rows probed)
};

int route_max();
void launch_route_dist(const float* q32, const float* qn2, int nq, const float* c32,
                       const float* cn2, int nlist, int d16, float* R, hipStream_t st);
void launch_probe_select(const float* R, int nq, int nlist, int avail, int nprobe,
                         const int* list_off, int* probes, int* pcount, int* prows,
                         hipStream_t st);
void launch_ivf_plan(const int* probes, const int* pcount, int nq, int nprobe, int nb,
                     const int* list_off, const int* lorder, int bn, int rows_item, int* qcount,
                     int* qfill, int* qoff, int* ioff, int* qlist, ScanItem* items, int* n_items,
                     hipStream_t st);
void launch_ivf_tau(const IvfTauArgs& a, hipStream_t st);
void launch_probe_ranges(const int* probes, const int* pcount, int nq, int nprobe,
                         const int* list_off, int2* ranges, hipStream_t st);

}  // namespace mann
