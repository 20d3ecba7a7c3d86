// Small device utilities (fills, gathers, id conversion) used by the host layer.
#include "launch.h"
#include "util.h"

namespace mann {

__global__ void fill_i32_kernel(int* p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void fill_f32_kernel(float* p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// cnt[list[i]] = 0 for i < *count
__global__ void reset_listed_kernel(int* cnt, const int* list, const int* count, int max_n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < max_n && i < *count) cnt[list[i]] = 0;
}
// gather rows of a bf16 matrix and their norms: dst[i] = src[idx[i]], rows [m, m_pad) zero / +inf
__global__ void gather_rows_bf16_kernel(const bf16_t* src, const float* srcn, const int* idx,
                                        int m, int m_pad, int kp, bf16_t* dst, float* dstn) {
  int row = blockIdx.x;
  if (row >= m_pad) return;
  const bf16_t* s = row < m ? src + (int64_t)idx[row] * kp : nullptr;
  for (int c = threadIdx.x; c < kp; c += blockDim.x)
    dst[(int64_t)row * kp + c] = s ? s[c] : (bf16_t)0.0f;
  if (threadIdx.x == 0) dstn[row] = row < m ? srcn[idx[row]] : INFINITY;
}
__global__ void gather_rows_f32_kernel(const float* src, const int* idx, int m, int w, float* dst) {
  int row = blockIdx.x;
  if (row >= m) return;
  const float* s = src + (int64_t)idx[row] * w;
  for (int c = threadIdx.x; c < w; c += blockDim.x) dst[(int64_t)row * w + c] = s[c];
}
__global__ void gather_rows_i32_kernel(const int* src, const int* idx, int m, int w, int* dst) {
  int row = blockIdx.x;
  if (row >= m) return;
  const int* s = src + (int64_t)idx[row] * w;
  for (int c = threadIdx.x; c < w; c += blockDim.x) dst[(int64_t)row * w + c] = s[c];
}
__global__ void labels_to_i32_kernel(const int64_t* src, int64_t n, int* dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (int)(src[i] - 1);
}
__global__ void gather_f32_kernel(const float* src, const int* idx, int m, float* dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) dst[i] = src[idx[i]];
}

__global__ void ids_to_i64_kernel(const int* src, int64_t n, int64_t* dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (int64_t)src[i] + 1;
}
// copy rows between strides (rows of w floats)
__global__ void copy_rows_kernel(const float* src, int64_t lds, float* dst, int64_t ldd, int w,
                                 int64_t rows) {
  int64_t r = blockIdx.x;
  if (r >= rows) return;
  for (int c = threadIdx.x; c < ldd; c += blockDim.x) dst[r * ldd + c] = c < w ? src[r * lds + c] : 0.0f;
}

// dst[r][c] = float(src[r][c]) for c < w (exact: every binary16 is a binary32)
// one wavefront per row: dst = src / |src| (FP64 norm), zero rows stay zero
__global__ __launch_bounds__(256) void unit_rows_kernel(const float* src, float* dst, int64_t ld,
                                                        int w, int64_t rows) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  double ss = 0.0;
  for (int c = lane; c < w; c += 64) ss += (double)src[r * ld + c] * (double)src[r * ld + c];
  for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o);
  const double sc = ss > 0.0 ? 1.0 / sqrt(ss) : 0.0;
  for (int c = lane; c < ld; c += 64) dst[r * ld + c] = c < w ? (float)(src[r * ld + c] * sc) : 0.0f;
}

__global__ void f16_rows_to_f32_kernel(const _Float16* src, int64_t lds, float* dst, int64_t ldd,
                                       int w, int64_t rows) {
  const int64_t r = blockIdx.x;
  if (r >= rows) return;
  for (int c = threadIdx.x; c < w; c += blockDim.x) dst[r * ldd + c] = (float)src[r * lds + c];
}

void f16_rows_to_f32(const void* src, int64_t lds, float* dst, int64_t ldd, int w, int64_t rows,
                     hipStream_t st) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(f16_rows_to_f32_kernel, dim3((unsigned)rows), dim3(128), 0, st,
                     (const _Float16*)src, lds, dst, ldd, w, rows);
}

void ids_to_i64(const int* src, int64_t n, int64_t* dst, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(ids_to_i64_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, src, n,
                     dst);
}
void copy_rows(const float* src, int64_t lds, float* dst, int64_t ldd, int w, int64_t rows,
               hipStream_t st) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(copy_rows_kernel, dim3((unsigned)rows), dim3(64), 0, st, src, lds, dst, ldd, w,
                     rows);
}

void fill_i32(int* p, int64_t n, int v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, p, n, v);
}
void fill_f32(float* p, int64_t n, float v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, p, n, v);
}
void reset_listed(int* cnt, const int* list, const int* count, int max_n, hipStream_t st) {
  if (max_n <= 0) return;
  hipLaunchKernelGGL(reset_listed_kernel, dim3((unsigned)ceil_div(max_n, 256)), dim3(256), 0, st,
                     cnt, list, count, max_n);
}
void gather_rows_bf16(const bf16_t* src, const float* srcn, const int* idx, int m, int m_pad,
                      int kp, bf16_t* dst, float* dstn, hipStream_t st) {
  if (m_pad <= 0) return;
  hipLaunchKernelGGL(gather_rows_bf16_kernel, dim3(m_pad), dim3(64), 0, st, src, srcn, idx, m,
                     m_pad, kp, dst, dstn);
}
void gather_rows_f32(const float* src, const int* idx, int m, int w, float* dst, hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(gather_rows_f32_kernel, dim3(m), dim3(64), 0, st, src, idx, m, w, dst);
}
void gather_rows_i32(const int* src, const int* idx, int m, int w, int* dst, hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(gather_rows_i32_kernel, dim3(m), dim3(64), 0, st, src, idx, m, w, dst);
}
void labels_to_i32(const int64_t* src, int64_t n, int* dst, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(labels_to_i32_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, src,
                     n, dst);
}
void gather_f32(const float* src, const int* idx, int m, float* dst, hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(gather_f32_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, src,
                     idx, m, dst);
}

void unit_rows(const float* src, float* dst, int64_t ld, int w, int64_t rows, hipStream_t st) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(unit_rows_kernel, dim3((unsigned)ceil_div(rows, 4)), dim3(256), 0, st, src,
                     dst, ld, w, rows);
}

}  // namespace mann
