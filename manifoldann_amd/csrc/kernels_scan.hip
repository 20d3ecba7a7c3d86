// First-pass distance scan on gfx950 MFMA (v_mfma_f32_32x32x16_bf16).
//
// A work item is (a tile of up to BN queries, a contiguous range of base rows).
// The workgroup streams the rows in BM-row tiles; per tile it computes the
// BM x BN block of  a' = |x~|^2 - 2 q~.x~  (bf16 operands, FP32 accumulate;
// a' + |q~|^2 is the squared distance of the bf16 vectors) with operands staged
// through double-buffered LDS, then runs the epilogue of its mode on the
// accumulators in registers:
//   FILTER : emit (a', row) to the query's candidate buffer when a' < tau[q]
//   DENSE  : store every a' (threshold estimation on a row sample)
//   TOP1   : FILTER whose per-lane threshold tightens to the certification band
//            of the smallest a' seen so far (k-means assignment)
// Only candidates cross HBM; the BM x BN block never leaves registers.
#include "launch.h"

namespace mann {

namespace {

constexpr int BK = 64;    // K chunk staged per LDS buffer
constexpr int LPAD = 8;   // bf16 padding per LDS row (kills ds_read_b128 bank conflicts)

// Exclusion threshold (a' domain) for the smallest norms-form distance, given
// the first-pass value v of an emitted candidate: any centroid whose a' is at
// least the result has an FP32 norms-form distance strictly above the
// candidate's.  Derivation: |a'+|q~|^2 - |q~-x~|^2| <= G, triangle inequality
// for the bf16 rounding (qe + ex), FP32 norms-form error H.
__device__ float top1_threshold(float v, float qn, float qe, float qr, float ex, float rx, int d) {
  const double u = 1.0 / 16777216.0;
  double R = (double)qr + rx;
  double G = (2.0 * d + 8.0) * u * R * R;
  double R32 = R + qe + ex;
  double H = (d + 8.0) * u * R32 * R32;
  double s = (double)v + qn + G;
  double dub = sqrt(s > 0.0 ? s : 0.0) + qe + ex;       // upper bound of |q - x|
  double fub = sqrt(dub * dub + H) * (1.0 + 8.0 * u);   // upper bound of FP32 F(v)
  double f2 = fub / (1.0 - 8.0 * u);
  double L = sqrt(f2 * f2 + H) + qe + ex;
  double t = (L * L - qn + G) * (1.0 + 1e-9) + 1e-30;
  return (float)t * (1.0f + 1e-6f);
}

template <int WM, int WN, int MODE>
struct ScanTile {
  static constexpr int BM = 64 * WM;
  static constexpr int BN = 64 * WN;
  static constexpr int A_CHUNKS = BM * (BK / 8) / 256;   // 16B chunks per thread
  static constexpr int B_CHUNKS = BN * (BK / 8) / 256;
};

template <int WM, int WN, int MODE>
__device__ __forceinline__ void scan_item(const ScanArgs& a, int qoff, int nqi, int row_begin,
                                          int row_end, bf16_t* lds) {
  using T = ScanTile<WM, WN, MODE>;
  constexpr int BM = T::BM, BN = T::BN, LROW = BK + LPAD;
  bf16_t* As = lds;                       // [2][BM][LROW]
  bf16_t* Bs = lds + 2 * BM * LROW;       // [2][BN][LROW]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  const int kp = a.kp, nk = kp / BK;
  const int ntiles = (row_end - row_begin + BM - 1) / BM;
  if (ntiles <= 0 || nqi <= 0) return;
  const int S = ntiles * nk;

  // query rows of the B tile (global query ids)
  const int* qidx = a.qidx;
  int brow[T::B_CHUNKS];
#pragma unroll
  for (int p = 0; p < T::B_CHUNKS; ++p) {
    int c = (tid + 256 * p) >> 3;
    int gq = c < nqi ? (qidx ? qidx[qoff + c] : qoff + c) : (qidx ? qidx[qoff] : qoff);
    brow[p] = gq;
  }
  // per-lane query columns and thresholds
  int gqc[WN];
  float tau_l[WN];
#pragma unroll
  for (int bn = 0; bn < WN; ++bn) {
    int c = wc * 32 * WN + bn * 32 + (lane & 31);
    int gq = c < nqi ? (qidx ? qidx[qoff + c] : qoff + c) : -1;
    gqc[bn] = gq;
    tau_l[bn] = (MODE != SCAN_DENSE && gq >= 0) ? a.tau[gq] : -INFINITY;
  }

  intx4 ra[T::A_CHUNKS], rb[T::B_CHUNKS];
  auto gload = [&](int s) {
    int t = s / nk, kc = s - t * nk;
    int r0 = row_begin + t * BM;
#pragma unroll
    for (int p = 0; p < T::A_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int r = c >> 3, cc = (c & 7) * 8;
      ra[p] = *(const intx4*)(a.xb + (int64_t)(r0 + r) * kp + kc * BK + cc);
    }
#pragma unroll
    for (int p = 0; p < T::B_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int cc = (c & 7) * 8;
      rb[p] = *(const intx4*)(a.qb + (int64_t)brow[p] * kp + kc * BK + cc);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int p = 0; p < T::A_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int r = c >> 3, cc = (c & 7) * 8;
      *(intx4*)(As + (buf * BM + r) * LROW + cc) = ra[p];
    }
#pragma unroll
    for (int p = 0; p < T::B_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int r = c >> 3, cc = (c & 7) * 8;
      *(intx4*)(Bs + (buf * BN + r) * LROW + cc) = rb[p];
    }
  };

  floatx16 acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  __syncthreads();  // LDS may still be read by a previous item of this workgroup
  gload(0);
  lstore(0);
  __syncthreads();

  for (int s = 0; s < S; ++s) {
    const int buf = s & 1;
    if (s + 1 < S) gload(s + 1);
    const bf16_t* Ab = As + buf * BM * LROW;
    const bf16_t* Bb = Bs + buf * BN * LROW;
#pragma unroll
    for (int ks = 0; ks < BK / 16; ++ks) {
      bf16x8 fa[WM], fb[WN];
      const int koff = ks * 16 + (lane >> 5) * 8;
#pragma unroll
      for (int i = 0; i < WM; ++i)
        fa[i] = *(const bf16x8*)(Ab + (wr * 32 * WM + i * 32 + (lane & 31)) * LROW + koff);
#pragma unroll
      for (int j = 0; j < WN; ++j)
        fb[j] = *(const bf16x8*)(Bb + (wc * 32 * WN + j * 32 + (lane & 31)) * LROW + koff);
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
    if (s + 1 < S) lstore(buf ^ 1);
    __syncthreads();

    const int t = s / nk;
    if (s - t * nk == nk - 1) {
      // ---------------- epilogue for row tile t ----------------
      const int r0 = row_begin + t * BM;
#pragma unroll
      for (int i = 0; i < WM; ++i) {
        const int rb0 = r0 + wr * 32 * WM + i * 32 + 4 * (lane >> 5);
        float nxv[16];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            int row = rb0 + 8 * g + e;
            nxv[4 * g + e] = row < row_end ? a.xn[row] : INFINITY;
          }
        }
#pragma unroll
        for (int j = 0; j < WN; ++j) {
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const int row = rb0 + 8 * (e >> 2) + (e & 3);
            const float v = __builtin_fmaf(-2.0f, acc[i][j][e], nxv[e]);
            if (MODE == SCAN_DENSE) {
              if (gqc[j] >= 0 && row < row_end)
                a.dense[(int64_t)gqc[j] * a.ldd + (row - row_begin)] = v;
            } else if (v < tau_l[j]) {
              const int gq = gqc[j];
              int pos = atomicAdd(a.cnt + gq, 1);
              if (pos < a.cap) {
                uint2 c;
                c.x = __float_as_uint(v);
                c.y = (unsigned)row;
                *(uint2*)(a.cand + ((int64_t)gq * a.cap + pos) * 2) = c;
              }
              if (MODE == SCAN_TOP1) {
                float th = top1_threshold(v, a.t1_qn[gq], a.t1_qe[gq], a.t1_qr[gq], a.t1_ex,
                                          a.t1_rx, a.d);
                tau_l[j] = fminf(tau_l[j], th);
              }
            }
            acc[i][j][e] = 0.0f;
          }
        }
      }
    }
  }
}

template <int WM, int WN, int MODE>
__global__ __launch_bounds__(256) void scan_grid_kernel(ScanArgs a) {
  using T = ScanTile<WM, WN, MODE>;
  extern __shared__ __align__(16) unsigned char smem[];
  bf16_t* lds = (bf16_t*)smem;
  int nq = a.nq_dev ? *a.nq_dev : a.nq;
  int qtile = blockIdx.x / a.parts, part = blockIdx.x - qtile * a.parts;
  int qoff = qtile * T::BN;
  if (qoff >= nq) return;
  int nqi = min(T::BN, nq - qoff);
  int rb = part * a.rows_per_part;
  int re = min(a.n_rows, rb + a.rows_per_part);
  scan_item<WM, WN, MODE>(a, qoff, nqi, rb, re, lds);
}

template <int WM, int WN, int MODE>
__global__ __launch_bounds__(256) void scan_items_kernel(ScanArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  bf16_t* lds = (bf16_t*)smem;
  __shared__ int s_item;
  const int n_items = *a.n_items_dev;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(a.item_counter, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= n_items) break;
    const ScanItem w = a.items[it];
    scan_item<WM, WN, MODE>(a, w.qoff, w.nq, w.row_begin, w.row_end, lds);
  }
}

template <int WM, int WN, int MODE>
void launch_t(const ScanArgs& a, int grid, bool persistent, hipStream_t st) {
  using T = ScanTile<WM, WN, MODE>;
  size_t lds = (size_t)2 * (T::BM + T::BN) * (BK + LPAD) * sizeof(bf16_t);
  static bool attr_set = false;  // dynamic LDS above 64 KiB must be opted into
  if (!attr_set) {
    MANN_HIP(hipFuncSetAttribute((const void*)scan_items_kernel<WM, WN, MODE>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    MANN_HIP(hipFuncSetAttribute((const void*)scan_grid_kernel<WM, WN, MODE>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  if (persistent)
    hipLaunchKernelGGL((scan_items_kernel<WM, WN, MODE>), dim3(grid), dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL((scan_grid_kernel<WM, WN, MODE>), dim3(grid), dim3(256), lds, st, a);
}

}  // namespace

int scan_tile_rows() { return ScanTile<2, 2, 0>::BM; }
int scan_tile_queries() { return ScanTile<2, 2, 0>::BN; }

void launch_scan(const ScanArgs& a, int mode, int grid, bool persistent, hipStream_t st) {
  if (grid <= 0) return;
  switch (mode) {
    case SCAN_FILTER: launch_t<2, 2, SCAN_FILTER>(a, grid, persistent, st); break;
    case SCAN_DENSE: launch_t<2, 2, SCAN_DENSE>(a, grid, persistent, st); break;
    case SCAN_TOP1: launch_t<2, 2, SCAN_TOP1>(a, grid, persistent, st); break;
    default: break;
  }
}

}  // namespace mann
