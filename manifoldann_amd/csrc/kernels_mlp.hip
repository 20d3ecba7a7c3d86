// Learned probe router: BF16 MFMA GEMM with fused epilogues, softmax
// cross-entropy against soft list labels, column sums and Adam.
//
// gemm: C[m][n] = sum_k A(m,k) * B(n,k) (+ bias[n]) (ReLU) with arbitrary
// element strides for A and B, so one kernel serves the forward pass (X W^T),
// the weight gradients (G^T H) and the input gradients (G W).  Operands are
// FP32 in memory and rounded to bf16 (RNE) when staged into LDS; accumulation
// is FP32 in the MFMA.
#include "mlp.h"

namespace mann {

namespace {

constexpr int GT = 64;    // tile (m and n)
constexpr int GK = 32;    // k chunk
constexpr int GPAD = 8;

__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  __shared__ bf16_t As[GT][GK + GPAD];
  __shared__ bf16_t Bs[GT][GK + GPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;  // 2 x 2 waves of 32 x 32
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  floatx16 acc;
  for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
  for (int k0 = 0; k0 < g.K; k0 += GK) {
    // stage: 64 x 32 elements per operand, 8 per thread; consecutive threads
    // walk the operand's unit-stride axis so the loads coalesce
    for (int i = tid; i < GT * GK; i += 256) {
      int r, kk;
      if (g.sak == 1) {
        r = i / GK;
        kk = i - r * GK;
      } else {
        kk = i / GT;
        r = i - kk * GT;
      }
      const int m = m0 + r, k = k0 + kk;
      const float v = g.A[(int64_t)min(m, g.M - 1) * g.sam + (int64_t)min(k, g.K - 1) * g.sak];
      As[r][kk] = (bf16_t)((m < g.M && k < g.K) ? v : 0.0f);
    }
    for (int i = tid; i < GT * GK; i += 256) {
      int r, kk;
      if (g.sbk == 1) {
        r = i / GK;
        kk = i - r * GK;
      } else {
        kk = i / GT;
        r = i - kk * GT;
      }
      const int n = n0 + r, k = k0 + kk;
      const float v = g.B[(int64_t)min(n, g.N - 1) * g.sbn + (int64_t)min(k, g.K - 1) * g.sbk];
      Bs[r][kk] = (bf16_t)((n < g.N && k < g.K) ? v : 0.0f);
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < GK / 16; ++ks) {
      const int koff = ks * 16 + (lane >> 5) * 8;
      bf16x8 fa = *(const bf16x8*)&As[wm * 32 + (lane & 31)][koff];
      bf16x8 fb = *(const bf16x8*)&Bs[wn * 32 + (lane & 31)][koff];
      // output rows follow the A operand, columns the B operand
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int n = n0 + wn * 32 + (lane & 31);
  const int nc = min(n, g.N - 1);
  // every load before the first use (clamped addresses)
  const float bias = g.bias ? g.bias[nc] : 0.0f;
  float old[16];
  if (g.accumulate) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int m = min(m0 + wm * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3), g.M - 1);
      old[e] = g.C[(int64_t)m * g.ldc + nc];
    }
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    float v = acc[e];
    if (g.bias) v += bias;
    if (g.relu) v = v > 0.0f ? v : 0.0f;
    if (g.accumulate) v += old[e];
    acc[e] = v;
    asm volatile("" : "+v"(acc[e]));  // computed here, not sunk into the stores
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int m = m0 + wm * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
    if (m < g.M && n < g.N) g.C[(int64_t)m * g.ldc + n] = acc[e];
  }
}

// Tiled form for operands whose K axis is unit-stride (the forward pass and
// the centroid prior): 128 x 128 output tiles on 4 wavefronts (each 64 x 64,
// 2 x 2 MFMA blocks), K chunks of 64 (32 when split) read as float4 one chunk ahead into
// registers and rounded to bf16 (RNE) when written to LDS.  SPLIT: each
// operand also keeps its rounding remainder lo = bf16(x - hi), so x = hi + lo
// to 16 significant bits, and the product hi.hi + hi.lo + lo.hi is accurate
// to about 2^-15 relative: an FP32-grade dot product on the MFMA.
constexpr int TT = 128;
template <bool SPLIT>
__global__ __launch_bounds__(256) void gemm_tiled_kernel(GemmArgs g) {
  constexpr int NS = SPLIT ? 2 : 1, TK = SPLIT ? 32 : 64, TLD = TK + 8;
  constexpr int NP = TT * TK / 4 / 256;  // float4 per thread and operand
  constexpr int RS = TK / 4;             // float4 per tile row
  __shared__ __align__(16) bf16_t As[NS][TT][TLD];
  __shared__ __align__(16) bf16_t Bs[NS][TT][TLD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = blockIdx.y * TT, n0 = blockIdx.x * TT;
  // float4 f = tid + 256 p covers tile row f / RS, k offset 4 (f % RS)
  float4 ra[NP], rb[NP];
  auto load = [&](int k0) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int f = tid + 256 * p, r = f / RS, k = k0 + (f % RS) * 4;
      const int m = m0 + r, n = n0 + r;
      // unconditional loads at clamped addresses (K is a multiple of 4), so
      // the eight loads are in flight together; out-of-range values are zeroed
      const int kc = min(k, g.K - 4);
      const float4 va = *(const float4*)(g.A + (int64_t)min(m, g.M - 1) * g.sam + kc);
      const float4 vb = *(const float4*)(g.B + (int64_t)min(n, g.N - 1) * g.sbn + kc);
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      ra[p] = (m < g.M && k < g.K) ? va : z;
      rb[p] = (n < g.N && k < g.K) ? vb : z;
    }
  };
  auto put = [&](bf16_t (*L)[TT][TLD], int r, int kk, float4 v) {
    const float x[4] = {v.x, v.y, v.z, v.w};
    typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
    bf16x4 h, l;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      h[c] = (bf16_t)x[c];
      if (SPLIT) l[c] = (bf16_t)(x[c] - (float)h[c]);
    }
    *(bf16x4*)&L[0][r][kk] = h;
    if (SPLIT) *(bf16x4*)&L[1][r][kk] = l;
  };
  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
  load(0);
  for (int k0 = 0; k0 < g.K; k0 += TK) {
    __syncthreads();  // the previous chunk's fragments have been read
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int f = tid + 256 * p, r = f / RS, kk = (f % RS) * 4;
      put(As, r, kk, ra[p]);
      put(Bs, r, kk, rb[p]);
    }
    __syncthreads();
    if (k0 + TK < g.K) load(k0 + TK);
#pragma unroll
    for (int ks = 0; ks < TK / 16; ++ks) {
      const int koff = ks * 16 + (lane >> 5) * 8;
      bf16x8 fa[NS][2], fb[NS][2];
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          fa[s2][i] = *(const bf16x8*)&As[s2][wm * 64 + i * 32 + (lane & 31)][koff];
          fb[s2][i] = *(const bf16x8*)&Bs[s2][wn * 64 + i * 32 + (lane & 31)][koff];
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][i], fb[0][j], acc[i][j], 0, 0, 0);
          if (SPLIT) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[0][i], fb[NS - 1][j], acc[i][j],
                                                                 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[NS - 1][i], fb[0][j], acc[i][j],
                                                                 0, 0, 0);
          }
        }
    }
  }
  // epilogue: every load issued before the first use (clamped addresses),
  // so the latencies overlap instead of adding up output by output
  float bias[2] = {0.0f, 0.0f};
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = min(n0 + wn * 64 + j * 32 + (lane & 31), g.N - 1);
    if (g.bias) bias[j] = g.bias[n];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    float old[2][16];
    if (g.accumulate) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int n = min(n0 + wn * 64 + j * 32 + (lane & 31), g.N - 1);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int m = min(m0 + wm * 64 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3),
                            g.M - 1);
          old[j][e] = g.C[(int64_t)m * g.ldc + n];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        float v = acc[i][j][e];
        if (g.bias) v += bias[j];
        if (g.relu) v = v > 0.0f ? v : 0.0f;
        if (g.accumulate) v += old[j][e];
        if (g.negate) v = -v;
        acc[i][j][e] = v;
        asm volatile("" : "+v"(acc[i][j][e]));  // computed here, not sunk into the stores
      }
  }
  // all values final before the first store (a load used after a store would
  // make the compiler wait for that store)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 64 + j * 32 + (lane & 31);
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int m = m0 + wm * 64 + i * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        if (m < g.M && n < g.N) g.C[(int64_t)m * g.ldc + n] = acc[i][j][e];
      }
    }
}

// One workgroup per row: softmax over the logits, cross-entropy against the
// label distribution (counts / L of the row's list labels), gradient
// (softmax - target) / batch written in place of the logits.
__global__ __launch_bounds__(256) void softmax_xent_kernel(float* __restrict__ Z, int ncls, int batch,
                                                           const int* __restrict__ labels, int L,
                                                           float* __restrict__ loss) {
  __shared__ float red[256];
  __shared__ float s_max, s_sum;
  const int b = blockIdx.x;
  float* z = Z + (int64_t)b * ncls;
  float mx = -INFINITY;
  for (int c = threadIdx.x; c < ncls; c += 256) mx = fmaxf(mx, z[c]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) s_max = red[0];
  __syncthreads();
  float sum = 0.0f;
  for (int c = threadIdx.x; c < ncls; c += 256) sum += __expf(z[c] - s_max);
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) s_sum = red[0];
  __syncthreads();
  const float lse = s_max + __logf(s_sum);
  const int* lb = labels + (int64_t)b * L;
  int nl = 0;
  for (int t = 0; t < L; ++t) nl += lb[t] >= 0;
  const float w = nl > 0 ? 1.0f / (float)nl : 0.0f;
  float lsum = 0.0f;
  for (int c = threadIdx.x; c < ncls; c += 256) {
    float tc = 0.0f;
    for (int t = 0; t < L; ++t) tc += lb[t] == c ? w : 0.0f;
    const float p = __expf(z[c] - lse);
    lsum += tc > 0.0f ? -tc * (z[c] - lse) : 0.0f;
    z[c] = (p - tc) / (float)batch;
  }
  red[threadIdx.x] = lsum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)batch);
}

// out[n] = sum_m G[m][n]: a block owns 32 columns, 8 row groups sum rows
// m = g (mod 8) with coalesced loads, then the 8 partials add in fixed order
__global__ __launch_bounds__(256) void colsum_kernel(const float* __restrict__ G, int M, int N,
                                                     float* __restrict__ out) {
  __shared__ float part[8][32];
  const int c = threadIdx.x & 31, gr = threadIdx.x >> 5;
  const int n = blockIdx.x * 32 + c;
  float s = 0.0f;
  if (n < N)
    for (int m = gr; m < M; m += 8) s += G[(int64_t)m * N + n];
  part[gr][c] = s;
  __syncthreads();
  if (gr == 0 && n < N) {
    float t = part[0][c];
    for (int i = 1; i < 8; ++i) t += part[i][c];
    out[n] = t;
  }
}

// dH *= (H > 0)
__global__ void relu_mask_kernel(float* __restrict__ dH, const float* __restrict__ H, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(H[i] > 0.0f)) dH[i] = 0.0f;
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m1, float* __restrict__ m2, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// out[c][r] = in[r][c] for r < rows, c < cols (32 x 32 tiles through LDS)
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ in, int64_t ldi,
                                                        int rows, int cols,
                                                        float* __restrict__ out, int64_t ldo) {
  __shared__ float t[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    t[i][tx] = (r < rows && c < cols) ? in[(int64_t)r * ldi + c] : 0.0f;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;
    if (c < cols && r < rows) out[(int64_t)c * ldo + r] = t[tx][i];
  }
}

__global__ void neg_kernel(const float* __restrict__ s, float* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = -s[i];
}

}  // namespace

bool gemm_tiled_ok(const GemmArgs& g) {
  return g.sak == 1 && g.sbk == 1 && g.sam % 4 == 0 && g.sbn % 4 == 0 && g.K % 4 == 0 &&
         ((uintptr_t)g.A & 15) == 0 && ((uintptr_t)g.B & 15) == 0;
}

void launch_gemm(const GemmArgs& g, hipStream_t st) {
  if (g.M <= 0 || g.N <= 0) return;
  if (gemm_tiled_ok(g)) {
    dim3 grid((unsigned)ceil_div(g.N, TT), (unsigned)ceil_div(g.M, TT));
    if (g.split)
      hipLaunchKernelGGL(gemm_tiled_kernel<true>, grid, dim3(256), 0, st, g);
    else
      hipLaunchKernelGGL(gemm_tiled_kernel<false>, grid, dim3(256), 0, st, g);
    return;
  }
  MANN_REQUIRE(!g.split && !g.negate, ERR_UNSUPPORTED,
               "split or negated GEMM needs unit-stride K operands");
  dim3 grid((unsigned)ceil_div(g.N, GT), (unsigned)ceil_div(g.M, GT));
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, st, g);
}

void launch_softmax_xent(float* Z, int ncls, int batch, const int* labels, int L, float* loss,
                         hipStream_t st) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(softmax_xent_kernel, dim3(batch), dim3(256), 0, st, Z, ncls, batch, labels, L,
                     loss);
}

void launch_colsum(const float* G, int M, int N, float* out, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)ceil_div(N, 32)), dim3(256), 0, st, G, M, N,
                     out);
}

void launch_relu_mask(float* dH, const float* H, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(relu_mask_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, dH, H, n);
}

void launch_adam(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1,
                 float b2, float eps, int step, hipStream_t st) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, w, g, m1, m2,
                     n, lr, b1, b2, eps, c1, c2);
}

// Centroid prior of the router: Z[b][c] += gamma * (2 x_b.c_c - |c_c|^2), i.e.
// -gamma |x_b - c_c|^2 up to a per-row constant (softmax and ranking ignore
// it).  FP32; a block holds 8 rows in LDS, each thread walks one centroid.
constexpr int PRIOR_ROWS = 8;
constexpr int PRIOR_DMAX = 1024;
__global__ __launch_bounds__(256) void prior_kernel(const float* __restrict__ X, int64_t ldx, int B,
                                                    int d, const float* __restrict__ C,
                                                    int64_t ldc, const float* __restrict__ cn2,
                                                    int nlist, float gamma, float* __restrict__ Z,
                                                    int64_t ldz) {
  __shared__ float xs[PRIOR_ROWS][PRIOR_DMAX];
  const int b0 = blockIdx.x * PRIOR_ROWS;
  for (int i = threadIdx.x; i < PRIOR_ROWS * d; i += 256) {
    const int r = i / d, t = i - r * d;
    xs[r][t] = b0 + r < B ? X[(int64_t)(b0 + r) * ldx + t] : 0.0f;
  }
  __syncthreads();
  for (int c = blockIdx.y * 256 + threadIdx.x; c < nlist; c += 256 * gridDim.y) {
    float acc[PRIOR_ROWS];
#pragma unroll
    for (int r = 0; r < PRIOR_ROWS; ++r) acc[r] = 0.0f;
    const float* cr = C + (int64_t)c * ldc;
    for (int t = 0; t < d; ++t) {
      const float cv = cr[t];
#pragma unroll
      for (int r = 0; r < PRIOR_ROWS; ++r) acc[r] = __builtin_fmaf(xs[r][t], cv, acc[r]);
    }
    const float cc = cn2[c];
#pragma unroll
    for (int r = 0; r < PRIOR_ROWS; ++r)
      if (b0 + r < B) Z[(int64_t)(b0 + r) * ldz + c] += gamma * (2.0f * acc[r] - cc);
  }
}

int prior_dmax() { return PRIOR_DMAX; }

void launch_prior(const float* X, int64_t ldx, int B, int d, const float* C, int64_t ldc,
                  const float* cn2, int nlist, float gamma, float* Z, int64_t ldz,
                  hipStream_t st) {
  if (B <= 0 || nlist <= 0) return;
  dim3 grid((unsigned)ceil_div(B, PRIOR_ROWS), (unsigned)ceil_div(nlist, 256));
  hipLaunchKernelGGL(prior_kernel, grid, dim3(256), 0, st, X, ldx, B, d, C, ldc, cn2, nlist, gamma,
                     Z, ldz);
}

void launch_transpose(const float* in, int64_t ldi, int rows, int cols, float* out, int64_t ldo,
                      hipStream_t st) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((unsigned)ceil_div(cols, 32), (unsigned)ceil_div(rows, 32));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, st, in, ldi, rows, cols, out, ldo);
}

void launch_neg(const float* s, float* d, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(neg_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s, d, n);
}

}  // namespace mann
