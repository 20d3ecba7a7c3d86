// k-means kernel interface (E-step selection, counting sort, FP64 M-step,
// kmeans++ rounds).
#pragma once

#include "launch.h"

namespace mann {

struct Top1Args {
  int64_t n;              // points
  int k;                  // centroids
  int d16;
  const float* x32;       // points, d16 floats each
  const float* xn2;       // A16 |x|^2 of the points
  const float* c32;       // centroids, d16 floats each
  const float* cn2;       // A16 |c|^2 of the centroids
  const unsigned* cand;   // emitted (a', centroid) pairs, cap per point
  const int* cnt;
  int cap;
  int* assign;            // out: 0-based centroid
  float* mindist;         // out: norms-form distance to it
  int* overflow;          // points resolved by the exact scan
};

void launch_top1_select(const Top1Args& a, hipStream_t st);
// ra: an upper bound of the centroids' uncentred norms
void launch_prime_threshold(const Top1Args& a, const float* qe, const float* qr, float ex, float rx,
                            float ra, int d, float* tau, hipStream_t st);
int sort_block();
// ints of the `table` scratch of launch_counting_sort (block histograms and
// the scan's tile sums)
int64_t counting_sort_table_ints(int64_t n, int k);
void launch_counting_sort(const int* assign, int64_t n, int k, int* table, int* offsets,
                          int* members, hipStream_t st);
void launch_chunk_sums(const float* x32, int d, int d16, const int* members, const int2* chunks,
                       int nchunks, double* part, hipStream_t st);
void launch_cluster_means(const double* part, const int* chunk_off, const int* counts, int k, int d,
                          int d16, float* newc, hipStream_t st);
void launch_movement(const float* oldc, const float* newc, int k, int d16, float* mv,
                     hipStream_t st);
void launch_farthest(const float* md, int64_t n, unsigned long long* out, hipStream_t st);
void launch_pp_first(const double* draws, int64_t n, const float* x32, int d16, float* centres,
                     int* picked, double* mins, hipStream_t st);
// bsum: ceil(n / 1024) doubles
void launch_pp_round(const float* x32, const float* xn2, int64_t n, int d16, float* centres,
                     int round, double* mins, double* bsum, const double* draws, int* picked,
                     hipStream_t st);

}  // namespace mann
