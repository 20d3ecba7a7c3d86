// Exact FP32 metrics (A16 order) and certification bounds shared by the
// selection kernels.
#pragma once

#include "launch.h"

namespace mann {

// Smallest first-pass threshold T (squared-distance domain) such that every row
// whose first-pass value S^ exceeds T has an FP32 exact metric strictly above dk.
// Ingredients: |S^ - |q~ - x~|^2| <= G
// (FP32 accumulation of bf16 products), |x - x~| <= ex and |q - q~| <= qe
// (bf16 rounding, triangle inequality), relative error rho of the FP32
// difference form, absolute error of the cosine form, error H of the norms
// form.  metric: EUCLIDEAN / SQEUCLIDEAN / COSINE / NORMS_FORM.
// gcoef > 0 replaces the first-pass error (2d + 8) u by gcoef (another
// first-pass arithmetic, e.g. FP8).
__device__ inline double exclusion_threshold(double dk, int metric, double qe, double qr,
                                             double ex, double rx, int d, double gcoef = 0.0) {
  const double u = 1.0 / 16777216.0;
  double R = qr + rx;
  double G = (gcoef > 0.0 ? gcoef : (2.0 * d + 8.0) * u) * R * R;
  double rho = (d + 8.0) * u;
  double L;
  if (metric == EUCLIDEAN) {
    L = dk / (1.0 - rho) + qe + ex;
  } else if (metric == SQEUCLIDEAN) {
    L = sqrt(dk / (1.0 - rho)) + qe + ex;
  } else if (metric == COSINE) {
    double delta = (2.0 * d + 16.0) * u;
    L = sqrt(2.0 * (dk + delta)) + qe + ex;
  } else {
    // (R32 bounds the uncentred |q| + |x| only when the bf16 operands are
    // not centred; norms_form_threshold below takes the uncentred norms)
    double R32 = R + qe + ex;
    double H = (d + 8.0) * u * R32 * R32;
    double f = dk / (1.0 - 8.0 * u);
    L = sqrt(f * f + H) + qe + ex;
  }
  return (L * L + G) * (1.0 + 1e-9) + 1e-30;
}

// Norms form (kmeans/distance.jl:99-101): given a first-pass value v, every row
// whose first-pass value is at least the result has an FP32 norms-form
// distance strictly above that of any row whose first-pass value is at most v.
// Ingredients: |S^ - |q~ - x~|^2| <= G = (2d + 8) u (|q~| + max|x~|)^2, the
// bf16 rounding |q - q~| <= qe and |x - x~| <= ex, and the FP32 norms-form
// error |F^2 - |q - x|^2| <= H = (d + 8) u (|q| + |x|)^2.  H takes the
// uncentred norms (the bf16 operands may be centred, the FP32 form is not):
// |q| from its A16 square qn2, and ra >= max |x|.
__device__ inline double norms_form_threshold(double v, double qe, double qr, double qn2,
                                              double ex, double rx, double ra, int d) {
  const double u = 1.0 / 16777216.0;
  const double R = qr + rx;
  const double G = (2.0 * d + 8.0) * u * R * R;
  const double qa = sqrt((qn2 > 0.0 ? qn2 : 0.0) * (1.0 + (d + 8.0) * u)) * (1.0 + 4.0 * u);
  const double R32 = qa + ra;
  const double H = (d + 8.0) * u * R32 * R32;
  const double s = v + G;
  const double dub = sqrt(s > 0.0 ? s : 0.0) + qe + ex;      // >= |q - x| of such a row
  const double fub = sqrt(dub * dub + H) * (1.0 + 8.0 * u);  // >= its FP32 norms form
  const double f2 = fub / (1.0 - 8.0 * u);
  const double L = sqrt(f2 * f2 + H) + qe + ex;
  return (L * L + G) * (1.0 + 1e-9) + 1e-30;
}

// Final step of the exact FP32 metric from the A16 sum acc (the sum of squared
// differences for the Euclidean metrics, the dot product otherwise).
__device__ __forceinline__ float finish_metric(float acc, int metric, float xn2r, float qn2q) {
  if (metric == EUCLIDEAN) return __builtin_sqrtf(acc);
  if (metric == SQEUCLIDEAN) return acc;
  if (metric == COSINE) {
    float v = 1.0f - acc / (__builtin_sqrtf(xn2r) * __builtin_sqrtf(qn2q));
    return v > 0.0f ? v : 0.0f;
  }
  // norms form (kmeans/distance.jl:99-101): sqrt(max((|c|^2 + |x|^2) - 2 c.x, 0))
  float tt = xn2r + qn2q;
  float sq = tt - 2.0f * acc;
  return __builtin_sqrtf(sq > 0.0f ? sq : 0.0f);
}

// Exact FP32 metric of base row r against query q (16-lane group, A16 order).
__device__ __forceinline__ float exact_metric(const float* __restrict__ x32,
                                              const float* __restrict__ q32,
                                              const float* __restrict__ xn2,
                                              const float* __restrict__ qn2, int d16, int metric,
                                              int64_t r, int64_t q, int j) {
  const float* x = x32 + r * d16;
  const float* y = q32 + q * d16;
  float acc = 0.0f;
  if (metric == EUCLIDEAN || metric == SQEUCLIDEAN) {
    for (int t = 0; t < d16; t += 16) {
      float df = x[t + j] - y[t + j];
      acc = __builtin_fmaf(df, df, acc);
    }
  } else {
    for (int t = 0; t < d16; t += 16) acc = __builtin_fmaf(x[t + j], y[t + j], acc);
  }
  acc = group16_tree(acc);
  const bool diff = metric == EUCLIDEAN || metric == SQEUCLIDEAN;
  return finish_metric(acc, metric, diff ? 0.0f : xn2[r], diff ? 0.0f : qn2[q]);
}

}  // namespace mann
