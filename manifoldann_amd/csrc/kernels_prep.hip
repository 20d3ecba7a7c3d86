// Row preparation: FP32 padded copy, BF16 copy for the MFMA first pass, FP64
// norms / quantisation-error norms (the bounds that certify the re-rank), and
// A16-order FP32 squared norms (the norms-form arithmetic of the oracle).
#include "launch.h"

namespace mann {

// One wavefront per row.  src rows have stride ld (Julia column stride).
//  x32 : rows of d16 floats (zero padded)        -- exact re-rank operand
//  xb  : rows of kp bf16 (zero padded)           -- MFMA operand
//  xn  : |bf16(x)|^2 (FP64 sum rounded to FP32); +inf on padding rows so a
//        padding row can never pass a threshold
//  err : |x - bf16(x)| rounded up; nrm : |bf16(x)| rounded up
// normalize != 0 (cosine): the bf16 copy holds x/|x| (|x| in FP64) and the
// bounds refer to the normalised vector.
// center (optional): the bf16 copy holds x - center (center_row[row] selects a
// row of `center` per point, e.g. its IVF list centroid); distances are
// translation invariant, and the smaller norms tighten every bound.
__global__ __launch_bounds__(256) void prep_rows_kernel(const float* __restrict__ src, int64_t ld,
                                                        int64_t n, int64_t n_pad, int d, int d16,
                                                        int kp, int normalize,
                                                        const float* __restrict__ center,
                                                        const int* __restrict__ center_row,
                                                        int64_t center_ld,
                                                        float* __restrict__ x32,
                                                        bf16_t* __restrict__ xb,
                                                        float* __restrict__ xn,
                                                        float* __restrict__ err,
                                                        float* __restrict__ nrm,
                                                        unsigned* __restrict__ max_err,
                                                        unsigned* __restrict__ max_nrm) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n_pad) return;
  if (row >= n) {
    for (int c = lane; c < d16; c += 64) x32[row * d16 + c] = 0.0f;
    for (int c = lane; c < kp; c += 64) xb[row * kp + c] = (bf16_t)0.0f;
    if (lane == 0) xn[row] = INFINITY;
    return;
  }
  const float* s = src + row * ld;
  const float* cen = center ? (center_row ? center + (int64_t)center_row[row] * center_ld : center)
                            : nullptr;
  double scale = 1.0;
  if (normalize) {
    double ss = 0.0;
    for (int c = lane; c < d; c += 64) ss += (double)s[c] * (double)s[c];
    for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o);
    scale = ss > 0.0 ? 1.0 / sqrt(ss) : 0.0;
  }
  double nb = 0.0, ee = 0.0;
  for (int c = lane; c < kp; c += 64) {
    float v = c < d ? s[c] : 0.0f;
    if (c < d16) x32[row * d16 + c] = v;
    double t = (double)v * scale;
    if (cen && c < d) t -= (double)cen[c];
    bf16_t b = (bf16_t)(float)t;
    xb[row * kp + c] = b;
    double bd = (double)(float)b;
    nb += bd * bd;
    ee += (t - bd) * (t - bd);
  }
  for (int o = 32; o >= 1; o >>= 1) {
    nb += __shfl_xor(nb, o);
    ee += __shfl_xor(ee, o);
  }
  if (lane == 0) {
    xn[row] = (float)nb;
    float e = (float)(sqrt(ee) * (1.0 + 1e-6)) + 1e-30f;
    float r = (float)(sqrt(nb) * (1.0 + 1e-6)) + 1e-30f;
    if (err) err[row] = e;
    if (nrm) nrm[row] = r;
    if (max_err) atomicMax(max_err, __float_as_uint(e));
    if (max_nrm) atomicMax(max_nrm, __float_as_uint(r));
  }
}

// Upper bound of the largest norm sqrt(n2[i]) of n vectors given their A16
// FP32 squared norms (relative error below (d + 8) u), as float bits for
// atomicMax (non-negative floats order like their bits).
__global__ void norm_bound_max_kernel(const float* __restrict__ n2, int64_t n, int d,
                                      unsigned* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u = 1.0 / 16777216.0;
  const double v = n2[i] > 0.0f ? (double)n2[i] : 0.0;
  const float b = (float)(sqrt(v * (1.0 + (d + 8.0) * u)) * (1.0 + 4.0 * u)) + 1e-30f;
  atomicMax(out, __float_as_uint(b));
}

// A16-order FP32 squared norm of each row of x32 (16 lanes per row).
__global__ __launch_bounds__(256) void a16_norms_kernel(const float* __restrict__ x32, int64_t n,
                                                        int d16, float* __restrict__ out) {
  const int j = threadIdx.x & 15;
  const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  float acc = 0.0f;
  if (row < n) {
    const float* x = x32 + row * d16;
    for (int t = 0; t < d16; t += 16) acc = __builtin_fmaf(x[t + j], x[t + j], acc);
  }
  acc = group16_tree(acc);
  if (row < n && j == 0) out[row] = acc;
}

// v = h1 + h2 + h3 exactly, each a bf16 (round to nearest): h1 keeps the top
// 8 significant bits, the FP32 remainder v - h1 is exact and has at most 16,
// the next at most 8, so h3 represents the last remainder exactly.
__device__ __forceinline__ void split3_bf16(float v, bf16_t* o) {
  const bf16_t h1 = (bf16_t)v;
  const float r1 = v - (float)h1;
  const bf16_t h2 = (bf16_t)r1;
  const float r2 = r1 - (float)h2;
  o[0] = h1;
  o[1] = h2;
  o[2] = (bf16_t)r2;
}

__global__ __launch_bounds__(256) void fold_base_kernel(bf16_t* __restrict__ xb,
                                                        const float* __restrict__ xn, int64_t n,
                                                        int kp, int d) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  bf16_t* row = xb + r * kp + d;
  split3_bf16(xn[r], row);
  row[3] = row[4] = row[5] = (bf16_t)1.0f;
}

// out row q: -2 q~ (exact), then 1, 1, 1, then the parts of |q~|^2; padding
// rows (q >= nq, up to nq_pad) zero
__global__ __launch_bounds__(256) void fold_queries_kernel(const bf16_t* __restrict__ qb,
                                                           const float* __restrict__ qn,
                                                           int64_t nq, int64_t nq_pad, int kp,
                                                           int d, bf16_t* __restrict__ out) {
  const int64_t q = blockIdx.x;
  if (q >= nq_pad) return;
  bf16_t* o = out + q * kp;
  for (int c = threadIdx.x; c < kp; c += 256) {
    float v = 0.0f;
    if (q < nq) {
      if (c < d) v = -2.0f * (float)qb[q * kp + c];
      else if (c < d + 3) v = 1.0f;
      else if (c < d + 6) continue;  // the norm's parts, written below
    }
    o[c] = (bf16_t)v;
  }
  if (q < nq && threadIdx.x == 0) split3_bf16(qn[q], o + d + 3);
}

void launch_fold_base(bf16_t* xb, const float* xn, int64_t n, int kp, int d, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fold_base_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, xb, xn,
                     n, kp, d);
}

void launch_fold_queries(const bf16_t* qb, const float* qn, int64_t nq, int64_t nq_pad, int kp,
                         int d, bf16_t* out, hipStream_t st) {
  if (nq_pad <= 0) return;
  hipLaunchKernelGGL(fold_queries_kernel, dim3((unsigned)nq_pad), dim3(256), 0, st, qb, qn, nq,
                     nq_pad, kp, d, out);
}

void launch_prep_rows(const float* src, int64_t ld, int64_t n, int64_t n_pad, int d, int d16,
                      int kp, int normalize, float* x32, bf16_t* xb, float* xn, float* err,
                      float* nrm, unsigned* max_err, unsigned* max_nrm, hipStream_t st,
                      const float* center, const int* center_row, int64_t center_ld) {
  if (n_pad <= 0) return;
  hipLaunchKernelGGL(prep_rows_kernel, dim3((unsigned)ceil_div(n_pad, 4)), dim3(256), 0, st, src,
                     ld, n, n_pad, d, d16, kp, normalize, center, center_row, center_ld, x32, xb,
                     xn, err, nrm, max_err, max_nrm);
}

// column sums in FP64 (one block per 1024 rows, atomics across blocks)
__global__ __launch_bounds__(256) void colsum64_kernel(const float* __restrict__ x, int64_t ld,
                                                       int64_t n, int d, double* __restrict__ out) {
  const int64_t r0 = (int64_t)blockIdx.x * 1024;
  const int64_t r1 = r0 + 1024 < n ? r0 + 1024 : n;
  for (int c = threadIdx.x; c < d; c += 256) {
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) s += (double)x[r * ld + c];
    atomicAdd(out + c, s);
  }
}
__global__ void mean_from_sum_kernel(const double* __restrict__ sum, int64_t n, int d,
                                     float* __restrict__ mean) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < d) mean[c] = (float)(sum[c] / (double)n);
}

void launch_norm_bound_max(const float* n2, int64_t n, int d, unsigned* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(norm_bound_max_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, n2,
                     n, d, out);
}

void launch_mean_rows(const float* x, int64_t ld, int64_t n, int d, double* tmp, float* mean,
                      hipStream_t st) {
  MANN_HIP(hipMemsetAsync(tmp, 0, sizeof(double) * d, st));
  hipLaunchKernelGGL(colsum64_kernel, dim3((unsigned)ceil_div(n, 1024)), dim3(256), 0, st, x, ld, n,
                     d, tmp);
  hipLaunchKernelGGL(mean_from_sum_kernel, dim3((unsigned)ceil_div(d, 256)), dim3(256), 0, st, tmp,
                     n, d, mean);
}

void launch_a16_norms(const float* x32, int64_t n, int d16, float* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(a16_norms_kernel, dim3((unsigned)ceil_div(n, 16)), dim3(256), 0, st, x32, n,
                     d16, out);
}

}  // namespace mann
