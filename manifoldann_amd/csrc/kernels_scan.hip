// First-pass distance scan on gfx950 MFMA (v_mfma_f32_32x32x16_bf16).
//
// A work item is (a tile of up to BN queries, a contiguous range of base rows).
// The workgroup streams the rows in BM-row tiles; per tile it computes the
// BM x BN block of  a' = |x~|^2 - 2 q~.x~  (bf16 operands, FP32 accumulate;
// a' + |q~|^2 is the squared distance of the bf16 vectors) with operands staged
// through double-buffered LDS, then runs the epilogue of its mode on the
// accumulators in registers:
//   FILTER : emit (a', row) when a' < tau[q] into the query's private region
//            of this item (no global atomics: counters live in LDS)
//   DENSE  : store every a' (threshold estimation on a row sample)
//   TOP1   : FILTER whose per-lane threshold tightens to the certification band
//            of the smallest a' seen so far (k-means assignment)
//   GATHER : FILTER for work items that share queries (IVF lists): emissions go
//            to an LDS buffer per query and reach the query's global buffer
//            with one atomic per (item, query) at the end of the item
// Only candidates cross HBM; the BM x BN block never leaves registers.
//
// Layout: 4 waves as 2 (rows) x 2 (queries); a wave owns 32*WM rows x 32*WN
// queries, i.e. WM x WN MFMA blocks.  MFMA 32x32 output: lane l holds query
// column l%32 and rows 8*(e/4) + 4*(l/32) + e%4 of register e.
#include "launch.h"

namespace mann {

namespace {

constexpr int BK = 32;    // K chunk staged per LDS buffer
constexpr int LPAD = 8;   // bf16 padding per LDS row (conflict-free ds_read_b128)

// Exclusion threshold (squared-distance domain) for the smallest norms-form
// distance, given the first-pass value v of an emitted candidate: any centroid
// whose first-pass value is at least the result has an FP32 norms-form distance
// strictly above the candidate's.  Derivation: |S^ - |q~ - x~|^2| <= G, triangle
// inequality for the bf16 rounding (qe + ex), FP32 norms-form error H.
__device__ float top1_threshold(float v, float qe, float qr, float ex, float rx, int d) {
  const double u = 1.0 / 16777216.0;
  double R = (double)qr + rx;
  double G = (2.0 * d + 8.0) * u * R * R;
  double R32 = R + qe + ex;
  double H = (d + 8.0) * u * R32 * R32;
  double s = (double)v + G;
  double dub = sqrt(s > 0.0 ? s : 0.0) + qe + ex;       // upper bound of |q - x|
  double fub = sqrt(dub * dub + H) * (1.0 + 8.0 * u);   // upper bound of FP32 F(v)
  double f2 = fub / (1.0 - 8.0 * u);
  double L = sqrt(f2 * f2 + H) + qe + ex;
  double t = (L * L + G) * (1.0 + 1e-9) + 1e-30;
  return (float)t * (1.0f + 1e-6f);
}

template <int WM, int WN, int PF>
struct Tile {
  static constexpr int BM = 64 * WM;
  static constexpr int BN = 64 * WN;
  static constexpr int A_CHUNKS = BM * (BK / 8) / 256;  // 16B chunks per thread
  static constexpr int B_CHUNKS = BN * (BK / 8) / 256;
  static constexpr int LROW = BK + LPAD;
  static constexpr size_t LDS_BYTES = (size_t)2 * (BM + BN) * LROW * sizeof(bf16_t) + BN * 4;
  // GATHER mode: per query column a small LDS candidate buffer
  static constexpr int LCAP = BN > 128 ? 16 : 32;
  static constexpr size_t LDS_GATHER = LDS_BYTES + (size_t)BN * LCAP * 8;
};

// One work item.  Values are in the squared-distance domain:
//   S^ = |x~|^2 - 2 q~.x~ + |q~|^2   (fma, then the query's norm added)
// Global loads run PF K-chunks ahead through registers; LDS is double-buffered.
template <int WM, int WN, int PF, int MODE>
__device__ __forceinline__ void scan_item(const ScanArgs& a, int qoff, int nqi, int row_begin,
                                          int row_end, int region, unsigned char* smem) {
  using T = Tile<WM, WN, PF>;
  constexpr int BM = T::BM, BN = T::BN, LROW = T::LROW;
  bf16_t* As = (bf16_t*)smem;                   // [2][BM][LROW]
  bf16_t* Bs = As + 2 * BM * LROW;              // [2][BN][LROW]
  int* lcnt = (int*)(Bs + 2 * BN * LROW);       // [BN] emitted count per query column
  uint2* lbuf = (uint2*)(lcnt + BN);            // [BN][LCAP] (GATHER mode)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  const int kp = a.kp, nk = kp / BK;
  const int ntiles = (row_end - row_begin + BM - 1) / BM;
  if (ntiles <= 0 || nqi <= 0) return;
  const int S = ntiles * nk;

  const int* qidx = a.qidx;
  int brow[T::B_CHUNKS];
#pragma unroll
  for (int p = 0; p < T::B_CHUNKS; ++p) {
    int c = (tid + 256 * p) >> 2;
    if (a.b_by_slot)
      brow[p] = qoff + (c < nqi ? c : 0);
    else
      brow[p] = c < nqi ? (qidx ? qidx[qoff + c] : qoff + c) : (qidx ? qidx[qoff] : qoff);
  }
  int gqc[WN], qcl[WN];
  float tau_l[WN], qo[WN];
#pragma unroll
  for (int bn = 0; bn < WN; ++bn) {
    int c = wc * 32 * WN + bn * 32 + (lane & 31);
    int gq = c < nqi ? (qidx ? qidx[qoff + c] : qoff + c) : -1;
    gqc[bn] = gq;
    qcl[bn] = c;
    tau_l[bn] = (MODE != SCAN_DENSE && gq >= 0) ? a.tau[gq] : -INFINITY;
    float o = 0.0f;
    if (gq >= 0) o = a.col_norm ? a.col_norm[qoff + c] : (a.qnorm ? a.qnorm[gq] : 0.0f);
    qo[bn] = o;
  }

  intx4 ra[PF][T::A_CHUNKS], rb[PF][T::B_CHUNKS];
  auto gload = [&](int R, int s) {
    int t = s / nk, kc = s - t * nk;
    int r0 = row_begin + t * BM;
#pragma unroll
    for (int p = 0; p < T::A_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int r = c >> 2, cc = (c & 3) * 8;
      ra[R][p] = *(const intx4*)(a.xb + (int64_t)(r0 + r) * kp + kc * BK + cc);
    }
#pragma unroll
    for (int p = 0; p < T::B_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int cc = (c & 3) * 8;
      rb[R][p] = *(const intx4*)(a.qb + (int64_t)brow[p] * kp + kc * BK + cc);
    }
  };
  auto lstore = [&](int R, int buf) {
#pragma unroll
    for (int p = 0; p < T::A_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int r = c >> 2, cc = (c & 3) * 8;
      *(intx4*)(As + (buf * BM + r) * LROW + cc) = ra[R][p];
    }
#pragma unroll
    for (int p = 0; p < T::B_CHUNKS; ++p) {
      int c = tid + 256 * p;
      int r = c >> 2, cc = (c & 3) * 8;
      *(intx4*)(Bs + (buf * BN + r) * LROW + cc) = rb[R][p];
    }
  };

  floatx16 acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  // Epilogue of one BM-row tile: per MFMA block, S^ from the accumulators
  // (not written back), threshold mask, then the rare emissions.
  auto epilogue = [&](int t) {
    const int r0 = row_begin + t * BM;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int rb0 = r0 + wr * 32 * WM + i * 32 + 4 * (lane >> 5);
      float nx[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int row = rb0 + 8 * (e >> 2) + (e & 3);
        nx[e] = row < row_end ? a.xn[row] : INFINITY;
      }
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        if (MODE == SCAN_DENSE) {
          if (gqc[j] >= 0) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
              const int row = rb0 + 8 * (e >> 2) + (e & 3);
              const float v = __builtin_fmaf(-2.0f, acc[i][j][e], nx[e]) + qo[j];
              if (row < row_end) a.dense[(int64_t)gqc[j] * a.ldd + (row - row_begin)] = v;
            }
          }
          continue;
        }
        unsigned m = 0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const float v = __builtin_fmaf(-2.0f, acc[i][j][e], nx[e]) + qo[j];
          m |= (v < tau_l[j] ? 1u : 0u) << e;
        }
        if (m == 0) continue;
        const int gq = gqc[j];
        int pos = atomicAdd(&lcnt[qcl[j]], __builtin_popcount(m));
        unsigned* reg = a.cand + ((int64_t)gq * a.parts + region) * (int64_t)a.cap * 2;
        while (m) {
          const int e = __builtin_ctz(m);
          m &= m - 1;
          const int row = rb0 + 8 * (e >> 2) + (e & 3);
          float av = 0.0f, xv = 0.0f;
#pragma unroll
          for (int ee = 0; ee < 16; ++ee) {
            av = ee == e ? acc[i][j][ee] : av;
            xv = ee == e ? nx[ee] : xv;
          }
          const float v = __builtin_fmaf(-2.0f, av, xv) + qo[j];
          uint2 cv;
          cv.x = __float_as_uint(v);
          cv.y = (unsigned)row;
          if (MODE == SCAN_GATHER) {
            if (pos < T::LCAP) {
              lbuf[qcl[j] * T::LCAP + pos] = cv;
            } else {  // spill straight to the query's global buffer
              int gp = atomicAdd(a.cnt + gq, 1);
              if (gp < a.cap) *(uint2*)(a.cand + ((int64_t)gq * a.cap + gp) * 2) = cv;
            }
          } else if (pos < a.cap) {
            *(uint2*)(reg + 2 * pos) = cv;
          }
          ++pos;
          if (MODE == SCAN_TOP1) {
            float th = top1_threshold(v, a.t1_qe[gq], a.t1_qr[gq], a.t1_ex, a.t1_rx, a.d);
            tau_l[j] = fminf(tau_l[j], th);
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
  };

  auto compute = [&](int buf) {
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
  };

  // step s uses register set R = s % PF
  auto step = [&](int R, int s) {
    if (s + PF < S) gload(R, s + PF);
    compute(s & 1);
    if (s + 1 < S) lstore((R + 1) % PF, (s + 1) & 1);
    __syncthreads();
    const int t = s / nk;
    if (s - t * nk == nk - 1) epilogue(t);
  };

  __syncthreads();  // LDS may still be in use by a previous item of this workgroup
  for (int c = tid; c < BN; c += 256) lcnt[c] = 0;
#pragma unroll
  for (int r = 0; r < PF; ++r)
    if (r < S) gload(r, r);
  lstore(0, 0);
  __syncthreads();
  for (int s = 0; s < S; s += PF) {
    step(0, s);
    if (PF > 1 && s + 1 < S) step(1 % PF, s + 1);
  }

  if (MODE == SCAN_GATHER) {
    __syncthreads();
    for (int c = tid; c < nqi; c += 256) {
      const int n = min(lcnt[c], T::LCAP);
      if (n == 0) continue;
      const int gq = qidx ? qidx[qoff + c] : qoff + c;
      const int base = atomicAdd(a.cnt + gq, n);
      unsigned* dst = a.cand + (int64_t)gq * a.cap * 2;
      for (int i = 0; i < n && base + i < a.cap; ++i)
        *(uint2*)(dst + 2 * (base + i)) = lbuf[c * T::LCAP + i];
    }
  } else if (MODE != SCAN_DENSE) {
    __syncthreads();
    for (int c = tid; c < nqi; c += 256) {
      int gq = qidx ? qidx[qoff + c] : qoff + c;
      a.cnt[(int64_t)gq * a.parts + region] = lcnt[c];
    }
  }
}

template <int WM, int WN, int PF, int MODE, int REFINE>
__global__ __launch_bounds__(256) void scan_grid_kernel(ScanArgs a) {
  using T = Tile<WM, WN, PF>;
  extern __shared__ __align__(16) unsigned char smem[];
  int nq = REFINE ? *a.nq_dev : a.nq;
  int bid = blockIdx.x;
  const int G = gridDim.x;
  // XCD-aware order: consecutive logical ids share an XCD (and its L2)
  if (a.xcd_map && (G & 7) == 0) bid = (bid & 7) * (G >> 3) + (bid >> 3);
  int qtile, part;
  if (a.xcd_map == 2) {  // part-major: an XCD streams one row partition
    part = bid / a.qtiles;
    qtile = bid - part * a.qtiles;
  } else {
    qtile = bid / a.parts;
    part = bid - qtile * a.parts;
  }
  int qoff = qtile * T::BN;
  if (qoff >= nq || part >= a.parts) return;
  int nqi = min(T::BN, nq - qoff);
  int rb = part * a.rows_per_part;
  int re = min(a.n_rows, rb + a.rows_per_part);
  scan_item<WM, WN, PF, MODE>(a, qoff, nqi, rb, re, part, smem);
}

template <int WM, int WN, int PF, int MODE>
__global__ __launch_bounds__(256) void scan_items_kernel(ScanArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_item;
  const int n_items = *a.n_items_dev;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(a.item_counter, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= n_items) break;
    const ScanItem w = a.items[it];
    scan_item<WM, WN, PF, MODE>(a, w.qoff, w.nq, w.row_begin, w.row_end, w.region, smem);
  }
}

template <class K>
void set_lds(K kern, size_t lds) {
  MANN_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
}

template <int WM, int WN, int PF, int MODE>
void launch_t(const ScanArgs& a, int grid, int kind, hipStream_t st) {
  using T = Tile<WM, WN, PF>;
  const size_t lds = MODE == SCAN_GATHER ? T::LDS_GATHER : T::LDS_BYTES;
  static bool attr_set = false;  // dynamic LDS above 64 KiB must be opted into
  if (!attr_set) {
    set_lds(scan_items_kernel<WM, WN, PF, MODE>, lds);
    set_lds(scan_grid_kernel<WM, WN, PF, MODE, 0>, lds);
    set_lds(scan_grid_kernel<WM, WN, PF, MODE, 1>, lds);
    attr_set = true;
  }
  if (kind == SCAN_KIND_ITEMS)
    hipLaunchKernelGGL((scan_items_kernel<WM, WN, PF, MODE>), dim3(grid), dim3(256), lds, st, a);
  else if (kind == SCAN_KIND_REFINE)
    hipLaunchKernelGGL((scan_grid_kernel<WM, WN, PF, MODE, 1>), dim3(grid), dim3(256), lds, st, a);
  else
    hipLaunchKernelGGL((scan_grid_kernel<WM, WN, PF, MODE, 0>), dim3(grid), dim3(256), lds, st, a);
}

}  // namespace

// Tile configurations: wave tile 32*WM rows x 32*WN queries, prefetch depth PF.
struct ScanCfg {
  int wm, wn, pf;
};
static const ScanCfg kCfg[] = {{2, 2, 1}, {4, 2, 1}, {2, 4, 1}, {2, 2, 2}, {4, 2, 2}, {4, 4, 1}};
constexpr int kNumCfg = sizeof(kCfg) / sizeof(kCfg[0]);

int scan_num_cfg() { return kNumCfg; }
int scan_tile_rows() { return 64 * kCfg[tuning().scan_cfg].wm; }
int scan_tile_queries() { return 64 * kCfg[tuning().scan_cfg].wn; }

template <int WM, int WN, int PF>
static void launch_cfg(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st) {
  switch (mode) {
    case SCAN_FILTER: launch_t<WM, WN, PF, SCAN_FILTER>(a, grid, kind, st); break;
    case SCAN_DENSE: launch_t<WM, WN, PF, SCAN_DENSE>(a, grid, kind, st); break;
    case SCAN_TOP1: launch_t<WM, WN, PF, SCAN_TOP1>(a, grid, kind, st); break;
    case SCAN_GATHER: launch_t<WM, WN, PF, SCAN_GATHER>(a, grid, kind, st); break;
    default: break;
  }
}

void launch_scan(const ScanArgs& a, int mode, int grid, int kind, hipStream_t st) {
  if (grid <= 0) return;
  switch (tuning().scan_cfg) {
    case 1: launch_cfg<4, 2, 1>(a, mode, grid, kind, st); break;
    case 2: launch_cfg<2, 4, 1>(a, mode, grid, kind, st); break;
    case 3: launch_cfg<2, 2, 2>(a, mode, grid, kind, st); break;
    case 4: launch_cfg<4, 2, 2>(a, mode, grid, kind, st); break;
    case 5: launch_cfg<4, 4, 1>(a, mode, grid, kind, st); break;
    default: launch_cfg<2, 2, 1>(a, mode, grid, kind, st); break;
  }
}

}  // namespace mann
