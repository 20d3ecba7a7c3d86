// Learned probe router kernels (BF16 MFMA GEMM, softmax cross-entropy, Adam).
#pragma once

#include "launch.h"

namespace mann {

struct GemmArgs {
  const float* A;       // A(m, k) = A[m * sam + k * sak]
  const float* B;       // B(n, k) = B[n * sbn + k * sbk]
  float* C;             // C[m * ldc + n]
  int64_t sam, sak, sbn, sbk, ldc;
  int M, N, K;
  const float* bias;    // optional, per n
  int relu;             // clamp at 0 after the bias
  int accumulate;       // C += result
  int split;            // FP32-accurate product (bf16 hi/lo split); needs gemm_tiled_ok
  int negate;           // store -result (after bias, ReLU and accumulate); tiled kernel only
};

// unit-stride K operands, 16-byte aligned rows: the tiled kernel applies
bool gemm_tiled_ok(const GemmArgs& g);
void launch_gemm(const GemmArgs& g, hipStream_t st);
void launch_softmax_xent(float* Z, int ncls, int batch, const int* labels, int L, float* loss,
                         hipStream_t st);
void launch_colsum(const float* G, int M, int N, float* out, hipStream_t st);
void launch_relu_mask(float* dH, const float* H, int64_t n, hipStream_t st);
void launch_adam(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t st);
void launch_neg(const float* s, float* d, int64_t n, hipStream_t st);
// out (cols x rows, row stride ldo) = transpose of in (rows x cols, row stride ldi)
void launch_transpose(const float* in, int64_t ldi, int rows, int cols, float* out, int64_t ldo,
                      hipStream_t st);
// Z[b][c] += gamma * (2 x_b.c_c - |c_c|^2) (FP32), d <= prior_dmax()
void launch_prior(const float* X, int64_t ldx, int B, int d, const float* C, int64_t ldc,
                  const float* cn2, int nlist, float gamma, float* Z, int64_t ldz,
                  hipStream_t st);
int prior_dmax();

}  // namespace mann
