// FP8 (OCP E4M3FN) first pass on gfx950: v_mfma_scale_f32_32x32x64_f8f6f4 with
// unit block scales.
//
// The scan is the bf16 FILTER scan with 64-element K steps: an LDS row of a
// chunk is 64 bytes either way (32 bf16 or 64 e4m3 values), so the global->LDS
// ring, its XOR swizzle and the row-norm staging are the same, and one chunk is
// one MFMA K step.  The fragments of chunk s+1 are read while chunk s is
// multiplied out (register double buffer across chunks; an even number of
// chunks per row keeps the buffer index a compile-time constant).
//
// Operand lanes: lane l holds row (l % 32) of its block and the 16-byte slots
// (l / 32) and 2 + (l / 32) of the 64-byte chunk.  The MFMA sums A[r][k] B[c][k]
// over the 64 k of both operands with the same lane-to-k assignment, so any
// fixed assignment gives the dot product; this one reads LDS exactly like the
// bf16 scan's two K steps (bank-conflict free under its swizzle).
//
// Accumulation: the hardware does not sum the FP8 products with FP32
// rounding.  Measured on random operands (all-positive, mixed-sign, full
// exponent range, subnormals), the error of one 32x32x64 instruction stays
// below 2^-11.1 of its largest product; the search bounds it by 2^-9 of the sum
// of the products' magnitudes (kFp8G).
#include "bounds.h"
#include "fp8.h"
#include "radix.h"

namespace mann {

namespace {

typedef __attribute__((address_space(3))) void lds_void;
typedef int intx8 __attribute__((ext_vector_type(8)));

// ---- E4M3FN encoding -------------------------------------------------------------
// Codes are monotone in the magnitude: subnormals m 2^-9 have code m (0..7),
// normals (8 + m) 2^(e-10) have code 8 e + m, and code 8 (m = 8 of the
// subnormal form) is the smallest normal 2^-6, so rounding to the nearest
// multiple of the local spacing gives the code directly.
__device__ __forceinline__ unsigned e4m3_encode(double v) {
  const unsigned s = v < 0.0 ? 0x80u : 0u;
  const double a = fabs(v);
  unsigned code;
  if (a < 0x1p-6) {
    code = (unsigned)__builtin_rint(a * 512.0);  // round half to even
  } else {
    int e;
    (void)frexp(a, &e);                          // a = f 2^e, f in [0.5, 1)
    const int E = e - 1;                         // a / 2^E in [1, 2)
    const double qv = __builtin_rint(ldexp(a, 3 - E));  // in [8, 16]
    code = (unsigned)(8 * (E + 6) + (int)qv);
  }
  if (code > 0x7eu) code = 0x7eu;                // saturate at 448 (0x7f is NaN)
  return s | code;
}
__device__ __forceinline__ double e4m3_decode(unsigned b) {
  const unsigned e = (b >> 3) & 15u, m = b & 7u;
  const double v = e == 0 ? ldexp((double)m, -9) : ldexp((double)(8u + m), (int)e - 10);
  return (b & 0x80u) ? -v : v;
}

// ---- row preparation --------------------------------------------------------------
// One wavefront per row; lane l owns the 4-byte groups l, l + 64, ... of the row.
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// the row's t values (FP64): x / |x| when normalising, minus the centre
struct RowT {
  const float* x;
  const float* cen;
  double scale;
  int d;
  __device__ double operator()(int c) const {
    if (c >= d) return 0.0;
    double t = (double)x[c] * scale;
    if (cen) t -= (double)cen[c];
    return t;
  }
};

__device__ __forceinline__ RowT row_t(const Fp8PrepArgs& a, int64_t row, int lane) {
  RowT r;
  r.x = a.x32 + row * a.d16;
  r.cen = a.center;
  r.d = a.d;
  r.scale = 1.0;
  if (a.normalize) {
    double ss = 0.0;
    for (int c = lane; c < a.d; c += 64) ss += (double)r.x[c] * (double)r.x[c];
    ss = wave_sum(ss);
    r.scale = ss > 0.0 ? 1.0 / sqrt(ss) : 0.0;
  }
  return r;
}

__global__ __launch_bounds__(256) void fp8_maxabs_kernel(Fp8PrepArgs a, unsigned* out) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= a.n) return;
  const RowT t = row_t(a, row, lane);
  double m = 0.0;
  for (int c = lane; c < a.d; c += 64) m = fmax(m, fabs(t(c)));
  m = wave_max(m);
  if (lane == 0) atomicMax(out, __float_as_uint((float)(m * (1.0 + 1e-7))));
}

__device__ __forceinline__ int scale_exp_of(double m) {
  if (!(m > 0.0)) return 0;
  int e = ilogb(kE4m3Max / m);
  while (ldexp(m, e) > kE4m3Max) --e;        // (the division rounds)
  while (ldexp(m, e + 1) <= kE4m3Max) ++e;
  return e;
}

__global__ __launch_bounds__(256) void fp8_prep_kernel(Fp8PrepArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= a.n_pad) return;
  unsigned* dst = (unsigned*)(a.x8 + row * a.kp8);
  if (row >= a.n) {
    for (int g = lane; g * 4 < a.kp8; g += 64) dst[g] = 0u;
    if (lane == 0) {
      a.xn8[row] = INFINITY;
      if (a.rowmul) a.rowmul[row] = 0.0f;
    }
    return;
  }
  const RowT t = row_t(a, row, lane);
  int s = a.scale_exp;
  if (s < -1000) {
    double m = 0.0;
    for (int c = lane; c < a.d; c += 64) m = fmax(m, fabs(t(c)));
    s = scale_exp_of(wave_max(m));
  }
  double nb = 0.0, ee = 0.0;
  for (int g = lane; g * 4 < a.kp8; g += 64) {
    unsigned w = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = 4 * g + u;
      const double tv = t(c);
      const unsigned b = c < a.d ? e4m3_encode(ldexp(tv, s)) : 0u;
      const double dv = ldexp(e4m3_decode(b), -s);  // x~ (exact)
      nb += dv * dv;
      ee += (tv - dv) * (tv - dv);
      w |= b << (8 * u);
    }
    dst[g] = w;
  }
  nb = wave_sum(nb);
  ee = wave_sum(ee);
  if (lane == 0) {
    a.xn8[row] = (float)nb;
    const float e = (float)(sqrt(ee) * (1.0 + 1e-6)) + 1e-30f;
    const float r = (float)(sqrt(nb) * (1.0 + 1e-6)) + 1e-30f;
    a.err8[row] = e;
    a.nrm8[row] = r;
    if (a.rowmul) a.rowmul[row] = (float)(-2.0 * ldexp(1.0, -(s + a.other_exp)));
    if (a.max_err) atomicMax(a.max_err, __float_as_uint(e));
    if (a.max_nrm) atomicMax(a.max_nrm, __float_as_uint(r));
  }
}

// ---- FP8 FILTER scan ---------------------------------------------------------------

template <int N>
__device__ __forceinline__ void wait_vm_lgkm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | (((N >> 4) & 3) << 14) | (7 << 4));
}
__device__ __forceinline__ unsigned lds_addr(const void* p) {
  return (unsigned)(uintptr_t)(const lds_void*)p;
}
__device__ __forceinline__ int lds_fetch_add(int* p, int v) {
  int r;
  asm volatile("ds_add_rtn_u32 %0, %1, %2\n\ts_waitcnt lgkmcnt(0)"
               : "=v"(r)
               : "v"(lds_addr(p)), "v"(v)
               : "memory");
  return r;
}
__device__ __forceinline__ floatx4 lds_read_f4(const float* p) {
  floatx4 r;
  asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(r) : "v"(lds_addr(p)) : "memory");
  return r;
}
__device__ __forceinline__ void bare_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}
// One FP8 MFMA with unit block scales (E8M0 127 = 2^0).  The host pass of the
// compiler does not know the gfx950 builtin; it never runs device code.
__device__ __forceinline__ floatx16 mfma_f8(intx8 a, intx8 b, floatx16 c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 127, 0, 127);
#else
  return c;
#endif
}
// 16-byte global->LDS load per lane
__device__ __forceinline__ void gl_lds16(const void* g, lds_void* l) {
  __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
}
// chunk slot swizzle of 64-byte LDS rows (see the bf16 scan)
__device__ __forceinline__ int swz4(int r) { return (r >> 3) & 3; }

template <int WM, int WN, int WR, int WC, int ST>
struct Tile8 {
  static constexpr int NT = 64 * WR * WC;
  static constexpr int BM = 32 * WM * WR;
  static constexpr int BN = 32 * WN * WC;
  static constexpr int STAGES = ST;
  static constexpr int LROW = 64;                       // bytes per LDS row of a chunk
  static constexpr int A_CHUNKS = BM * 4 / NT;          // 16-byte DMA ops per thread
  static constexpr int B_CHUNKS = BN * 4 / NT;
  static constexpr int NXB = 4;                         // row-norm ring (tiles)
  static_assert(WR * WC >= BM / 64, "one wavefront stages the norms of each 64 rows");
  static_assert(A_CHUNKS >= 1 && B_CHUNKS >= 1, "tile too small for the workgroup");
  // + per query column {tau, |q~|^2, multiplier, query id}, read by the
  // epilogue (registers are scarce across the K loop)
  static constexpr size_t LDS_BYTES =
      (size_t)STAGES * (BM + BN) * LROW + BN * 4 + (size_t)NXB * BM * 4 + (size_t)BN * 16;
  static constexpr int NPC = A_CHUNKS + B_CHUNKS;
};

template <int WM, int WN, int WR, int WC, int ST>
__device__ __forceinline__ void scan8_item(const ScanArgs& a, int qoff, int nqi, int row_begin,
                                           int row_end, int region, unsigned char* smem) {
  using T = Tile8<WM, WN, WR, WC, ST>;
  constexpr int BM = T::BM, BN = T::BN, LROW = T::LROW, STAGES = T::STAGES, NT = T::NT;
  unsigned char* As = smem;                            // [STAGES][BM][64]
  unsigned char* Bs = As + STAGES * BM * LROW;         // [STAGES][BN][64]
  int* lcnt = (int*)(Bs + STAGES * BN * LROW);         // [BN]
  float* nxs = (float*)(lcnt + BN);                    // [NXB][BM]
  floatx4* colv = (floatx4*)(nxs + T::NXB * BM);       // [BN]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave % WR, wc = wave / WR;
  const int kp = a.kp, nk = kp / 64;
  const int ntiles = (row_end - row_begin + BM - 1) / BM;
  if (ntiles <= 0 || nqi <= 0) return;
  const int* qidx = a.qidx;

  unsigned offA[T::A_CHUNKS];
  int64_t offB[T::B_CHUNKS];
#pragma unroll
  for (int p = 0; p < T::A_CHUNKS; ++p) {
    const int c = tid + NT * p, r = c / 4;
    offA[p] = (unsigned)(r * kp + (((c % 4) ^ swz4(r)) * 16));
  }
#pragma unroll
  for (int p = 0; p < T::B_CHUNKS; ++p) {
    const int c = tid + NT * p, r = c / 4;
    const int brow = r < nqi ? (qidx ? qidx[qoff + r] : qoff + r) : (qidx ? qidx[qoff] : qoff);
    offB[p] = (int64_t)brow * kp + (((c % 4) ^ swz4(r)) * 16);
  }

  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const int nx_slot = (wv % (BM / 64)) * 64;
  const __amdgpu_buffer_rsrc_t nx_rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)a.xn, 0, 0xffffffffu, 0x00020000);
  auto dma = [&](int t, int kc, int st) {
    if (kc == 0)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(nx_rsrc,
                                               (lds_void*)(nxs + (t % T::NXB) * BM + nx_slot), 4,
                                               lane * 4,
                                               (int)((unsigned)(row_begin + t * BM + nx_slot) * 4u),
                                               0, 0);
    const unsigned char* baseA = a.x8 + (int64_t)(row_begin + t * BM) * kp + kc * 64;
    const unsigned char* baseB = a.q8 + kc * 64;
#pragma unroll
    for (int p = 0; p < T::A_CHUNKS; ++p)
      gl_lds16(baseA + offA[p], (lds_void*)(As + st * BM * LROW + (NT * p + 64 * wv) * 16));
#pragma unroll
    for (int p = 0; p < T::B_CHUNKS; ++p)
      gl_lds16(baseB + offB[p], (lds_void*)(Bs + st * BN * LROW + (NT * p + 64 * wv) * 16));
  };
  auto wait_chunks = [&](int later) {
    constexpr int N = T::NPC;
    static_assert(STAGES >= 3 && STAGES <= 5 && (STAGES - 2) * N <= 63, "vmcnt range");
    if (later <= 0) wait_vm_lgkm<0>();
    else if (later == 1 || STAGES == 3) wait_vm_lgkm<N>();
    else if (later == 2 || STAGES == 4) wait_vm_lgkm<2 * N>();
    else wait_vm_lgkm<3 * N>();
  };

  floatx16 acc[WM][WN];
  intx8 F[2][WM + WN];
  // fragments of the chunk in stage `stg` into buffer B
  auto rd = [&](auto BC, int stg) {
    constexpr int B = decltype(BC)::value;
    const unsigned char* Ab = As + stg * BM * LROW;
    const unsigned char* Bb = Bs + stg * BN * LROW;
    const int s0 = lane >> 5;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int r = wr * 32 * WM + i * 32 + (lane & 31);
      const intx4 lo = *(const intx4*)(Ab + r * LROW + ((s0 ^ swz4(r)) * 16));
      const intx4 hi = *(const intx4*)(Ab + r * LROW + (((2 + s0) ^ swz4(r)) * 16));
      F[B][i] = intx8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int r = wc * 32 * WN + j * 32 + (lane & 31);
      const intx4 lo = *(const intx4*)(Bb + r * LROW + ((s0 ^ swz4(r)) * 16));
      const intx4 hi = *(const intx4*)(Bb + r * LROW + (((2 + s0) ^ swz4(r)) * 16));
      F[B][WM + j] = intx8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
  };
  auto mm = [&](auto BC, bool zero) {
    constexpr int B = decltype(BC)::value;
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
        acc[i][j] = mfma_f8(F[B][i], F[B][WM + j], zero ? floatx16{} : acc[i][j]);
    __builtin_amdgcn_s_setprio(0);
  };

  auto epilogue = [&](int t) {
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int r0 = row_begin + t * BM;
    int gqc[WN];
    float tau_l[WN], qo[WN], qm[WN];
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const floatx4 v = lds_read_f4((const float*)&colv[wc * 32 * WN + j * 32 + (ln & 31)]);
      tau_l[j] = v[0];
      qo[j] = v[1];
      qm[j] = v[2];
      gqc[j] = __float_as_int(v[3]);
    }
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int rb0 = r0 + wr * 32 * WM + i * 32 + 4 * (ln >> 5);
      const float* nxt = nxs + (t % T::NXB) * BM + (rb0 - r0);
      float nx[16];
#pragma unroll
      for (int e4 = 0; e4 < 4; ++e4) {
        const floatx4 v4 = lds_read_f4(nxt + 8 * e4);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int row = rb0 + 8 * e4 + c;
          nx[4 * e4 + c] = row < row_end ? v4[c] : INFINITY;
        }
      }
      // a' = |x~|^2 - 2 q~.x~ (the multiplier is an exact power of two)
      auto val = [&](int j, int e) -> float { return __builtin_fmaf(qm[j], acc[i][j][e], nx[e]); };
      bool pass[WN];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        float m4[4];
#pragma unroll
        for (int g = 0; g < 4; ++g)
          m4[g] = fminf(fminf(val(j, 4 * g), val(j, 4 * g + 1)),
                        fminf(val(j, 4 * g + 2), val(j, 4 * g + 3)));
        const float mn = fminf(fminf(m4[0], m4[1]), fminf(m4[2], m4[3]));
        pass[j] = mn + qo[j] < tau_l[j];
      }
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        if (!pass[j]) continue;
        float w[16];
        unsigned m = 0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          w[e] = val(j, e) + qo[j];
          m |= (w[e] < tau_l[j] ? 1u : 0u) << e;
        }
        int gq = gqc[j];
        asm volatile("" : "+v"(gq));
        const int qc = wc * 32 * WN + j * 32 + (ln & 31);
        int pos = lds_fetch_add(&lcnt[qc], __builtin_popcount(m));
        unsigned* reg = a.cand + ((int64_t)gq * a.parts + region) * (int64_t)a.cap * 2;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          if (!((m >> e) & 1u)) continue;
          if (pos < a.cap) {
            uint2 cv;
            cv.x = __float_as_uint(w[e]);
            cv.y = (unsigned)(rb0 + 8 * (e >> 2) + (e & 3));
            *(uint2*)(reg + 2 * pos) = cv;
          }
          ++pos;
        }
      }
    }
  };

  __syncthreads();
  for (int c = tid; c < BN; c += NT) {
    lcnt[c] = 0;
    const int gq = c < nqi ? (qidx ? qidx[qoff + c] : qoff + c) : -1;
    floatx4 v;
    v[0] = gq >= 0 ? a.tau[gq] : -INFINITY;
    v[1] = gq >= 0 ? a.qnorm[gq] : 0.0f;
    v[2] = gq >= 0 ? a.qmul[gq] : 0.0f;
    v[3] = __int_as_float(gq);
    colv[c] = v;
  }
  const int S = ntiles * nk;
  constexpr int D = STAGES - 1;
  int ti = 0, ki = 0, issued = 0;
  for (int i = 0; i < D && i < S; ++i) {
    dma(ti, ki, i);
    ++issued;
    if (++ki == nk) {
      ki = 0;
      ++ti;
    }
  }
  wait_chunks(issued - 1);
  __syncthreads();
  int s = 0, st0 = 0, sti = D;
  using B0 = std::integral_constant<int, 0>;
  using B1 = std::integral_constant<int, 1>;
  // chunk s (fragments in buffer BC, read before): issue chunk s + D, wait for
  // chunk s+1, read its fragments into the other buffer (unless the tile ends
  // here: the epilogue needs the registers first), multiply chunk s out
  auto chunk = [&](auto BC, bool first, bool last) __attribute__((always_inline)) {
    constexpr int B = decltype(BC)::value;
    if (issued < S) {
      dma(ti, ki, sti);
      ++issued;
      if (++ki == nk) {
        ki = 0;
        ++ti;
      }
      sti = sti == STAGES - 1 ? 0 : sti + 1;
    }
    wait_chunks(issued - (s + 2));
    bare_barrier();
    st0 = st0 == STAGES - 1 ? 0 : st0 + 1;
    if (!last) rd(std::integral_constant<int, 1 - B>{}, st0);
    mm(BC, first);
    ++s;
  };
  rd(B0{}, st0);
  for (int t = 0; t < ntiles; ++t) {
    for (int kc = 0; kc < nk; kc += 2) {
      chunk(B0{}, kc == 0, false);
      chunk(B1{}, false, kc + 2 == nk);
    }
    epilogue(t);
    rd(B0{}, st0);
  }
  __syncthreads();
  for (int c = tid; c < nqi; c += NT) {
    const int gq = qidx ? qidx[qoff + c] : qoff + c;
    a.cnt[(int64_t)gq * a.parts + region] = lcnt[c];
  }
}

template <int WM, int WN, int WR, int WC, int ST, int REFINE>
__global__ __launch_bounds__(64 * WR * WC)
__attribute__((amdgpu_waves_per_eu(WR * WC >= 8 ? 2 : 1))) void scan8_grid_kernel(ScanArgs a) {
  using T = Tile8<WM, WN, WR, WC, ST>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nq = REFINE ? *a.nq_dev : a.nq;
  int bid = blockIdx.x;
  const int G = gridDim.x;
  if (a.xcd_map && (G & 7) == 0) bid = (bid & 7) * (G >> 3) + (bid >> 3);
  int qtile, part;
  if (a.xcd_map == 3) {
    const int SPB = a.xcd_band > 0 ? a.xcd_band : 8;
    const int g = bid / (SPB * a.qtiles);
    const int p0 = g * SPB;
    const int bp = min(SPB, a.parts - p0);
    const int idx = bid - g * SPB * a.qtiles;
    part = p0 + idx % bp;
    qtile = idx / bp;
  } else {
    qtile = bid / a.parts;
    part = bid - qtile * a.parts;
  }
  const int qoff = qtile * T::BN;
  if (qoff >= nq || part >= a.parts) return;
  const int nqi = min(T::BN, nq - qoff);
  const int rb = part * a.rows_per_part;
  const int re = min(a.n_rows, rb + a.rows_per_part);
  scan8_item<WM, WN, WR, WC, ST>(a, qoff, nqi, rb, re, part, smem);
}

template <int WM, int WN, int WR, int WC, int ST>
void launch8_t(const ScanArgs& a, int grid, int kind, hipStream_t st) {
  using T = Tile8<WM, WN, WR, WC, ST>;
  const size_t lds = T::LDS_BYTES;
  static bool attr_set = false;
  if (!attr_set) {
    MANN_HIP(hipFuncSetAttribute((const void*)scan8_grid_kernel<WM, WN, WR, WC, ST, 0>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    MANN_HIP(hipFuncSetAttribute((const void*)scan8_grid_kernel<WM, WN, WR, WC, ST, 1>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr_set = true;
  }
  if (kind == SCAN_KIND_REFINE)
    hipLaunchKernelGGL((scan8_grid_kernel<WM, WN, WR, WC, ST, 1>), dim3(grid), dim3(T::NT), lds,
                       st, a);
  else
    hipLaunchKernelGGL((scan8_grid_kernel<WM, WN, WR, WC, ST, 0>), dim3(grid), dim3(T::NT), lds,
                       st, a);
}

// ---- threshold conversion ---------------------------------------------------------
__global__ __launch_bounds__(256) void fp8_tau_kernel(Fp8TauArgs a) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= a.nq) return;
  const double u = 1.0 / 16777216.0;
  const float t = a.tau16[q];
  if (!(t < INFINITY)) {
    a.tau8[q] = INFINITY;
    return;
  }
  const double R16 = (double)a.qr16[q] + a.rx16;
  const double G16 = (2.0 * a.d + 8.0) * u * R16 * R16;
  const double R8 = (double)a.qr8[q] + a.rx8;
  const double G8 = kFp8G * R8 * R8;
  const double s = (double)t + G16;
  // |q - x| <= sqrt(S16 + G16) + qe16 + ex16 for every row at or below tau16,
  // and |q~8 - x~8| <= |q - x| + qe8 + ex8
  const double L = sqrt(s > 0.0 ? s : 0.0) + a.qe16[q] + a.ex16 + a.qe8[q] + a.ex8;
  const double v = (L * L + G8) * (1.0 + 1e-9) + 1e-30;
  a.tau8[q] = nextafterf((float)v, INFINITY);
}

// ---- bf16 re-score of the FP8 survivors -------------------------------------------
constexpr int RS_CAP = 8192;   // candidates per query held in LDS

struct RescoreShared {
  float vals[RS_CAP];
  uint32_t rows[RS_CAP];
  int pre[65];
  int total;
  int ovf;
  unsigned wpos, tpos;
  BlockRankScratch rank;
};

__global__ __launch_bounds__(256) void rescore_kernel(RescoreArgs a) {
  extern __shared__ __align__(16) unsigned char rs_smem[];
  RescoreShared& sm = *(RescoreShared*)rs_smem;
  const int tid = threadIdx.x;
  const int nq = a.nq_dev ? *a.nq_dev : a.nq;
  const int slot = blockIdx.x;
  if (slot >= nq) return;
  const int q = a.qlist ? a.qlist[slot] : slot;
  // region counts -> prefix (parts8 <= 64)
  if (tid < 64) {
    int cp = 0;
    bool ovf = false;
    if (tid < a.parts8) {
      const int v = a.cnt8[(int64_t)q * a.parts8 + tid];
      ovf = v > a.cap8;
      cp = v < a.cap8 ? v : a.cap8;
    }
    const bool any = __ballot(ovf) != 0;
    int inc = cp;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(inc, o);
      if (tid >= o) inc += t;
    }
    sm.pre[tid] = inc - cp;
    if (tid == 63) {
      sm.total = inc;
      sm.ovf = any ? 1 : 0;
    }
  }
  __syncthreads();
  const int total = sm.total;
  const bool ovf = sm.ovf != 0 || total > RS_CAP;
  const int c = total < RS_CAP ? total : RS_CAP;
  for (int i = tid; i < c; i += 256) {
    int lo = 0, hi = a.parts8 - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sm.pre[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const int64_t off = (((int64_t)q * a.parts8 + lo) * (int64_t)a.cap8 + (i - sm.pre[lo])) * 2;
    sm.rows[i] = a.cand8[off + 1];
  }
  __syncthreads();
  // S^ on the bf16 operands: a 16-lane group per candidate, four candidates
  // per step, lane j owning the 8-element groups j, j + 16, ... of the row
  const int g = tid >> 4, j = tid & 15;
  const bf16_t* qrow = a.qb + (int64_t)q * a.kp;
  const float qn = a.qn[q];
  for (int i0 = g * 4; i0 < c; i0 += 64) {
    const int nr = min(4, c - i0);
    const bf16_t* xr[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) xr[u] = a.xb + (int64_t)sm.rows[i0 + (u < nr ? u : 0)] * a.kp;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int t = 8 * j; t < a.kp; t += 128) {
      const bf16x8 qv = *(const bf16x8*)(qrow + t);
      bf16x8 xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) xv[u] = *(const bf16x8*)(xr[u] + t);
      // columns at or past d are zero in the queries, except in the dataset's
      // own rows, which may carry folded norms there (self-join)
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (t + e < a.d) acc[u] = __builtin_fmaf((float)xv[u][e], (float)qv[e], acc[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float dot = group16_tree(acc[u]);
      if (j == 0 && u < nr) {
        const uint32_t r = sm.rows[i0 + u];
        sm.vals[i0 + u] = __builtin_fmaf(-2.0f, dot, a.xn[r]) + qn;
      }
    }
  }
  __syncthreads();
  unsigned* out = a.cand + (int64_t)q * a.keep * 2;
  float dm = INFINITY;
  int kept;
  if (c <= a.keep) {
    for (int i = tid; i < c; i += 256) {
      uint2 cv;
      cv.x = __float_as_uint(sm.vals[i]);
      cv.y = sm.rows[i];
      *(uint2*)(out + 2 * i) = cv;
    }
    kept = c;
  } else {
    // the keep smallest: strictly below the keep-th smallest key, then ties
    const uint32_t key = block_rank_key(sm.vals, c, a.keep, sm.rank);
    if (tid == 0) {
      sm.wpos = 0;
      sm.tpos = 0;
    }
    __syncthreads();
    for (int i = tid; i < c; i += 256) {
      if (f2ord(sm.vals[i]) < key) {
        const unsigned p = atomicAdd(&sm.wpos, 1u);
        uint2 cv;
        cv.x = __float_as_uint(sm.vals[i]);
        cv.y = sm.rows[i];
        *(uint2*)(out + 2 * p) = cv;
      }
    }
    __syncthreads();
    const unsigned below = sm.wpos;
    for (int i = tid; i < c; i += 256) {
      if (f2ord(sm.vals[i]) == key) {
        const unsigned p = below + atomicAdd(&sm.tpos, 1u);
        if (p < (unsigned)a.keep) {
          uint2 cv;
          cv.x = __float_as_uint(sm.vals[i]);
          cv.y = sm.rows[i];
          *(uint2*)(out + 2 * p) = cv;
        }
      }
    }
    kept = a.keep;
    dm = ord2f(key);   // every value left out is at least the keep-th smallest
  }
  if (ovf) {
    // the selection reads all `keep` slots of an overflowed query (which it
    // never certifies): fill the unused ones with a valid row
    for (int i = kept + tid; i < a.keep; i += 256) {
      uint2 cv;
      cv.x = __float_as_uint(INFINITY);
      cv.y = 0u;
      *(uint2*)(out + 2 * i) = cv;
    }
  }
  if (tid == 0) {
    a.cnt[q] = ovf ? a.keep + 1 : kept;
    a.dropmin[q] = dm;
    if (a.stats) atomicAdd(a.stats + (slot % kStatSlots) * 4 + 2, total);
  }
}

}  // namespace

// ---- host side -----------------------------------------------------------------------

void launch_fp8_maxabs(const Fp8PrepArgs& a, unsigned* out, hipStream_t st) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(fp8_maxabs_kernel, dim3((unsigned)ceil_div(a.n, 4)), dim3(256), 0, st, a,
                     out);
}

void launch_fp8_prep(const Fp8PrepArgs& a, hipStream_t st) {
  MANN_REQUIRE(a.kp8 % 128 == 0 && a.kp8 >= a.d, ERR_ASSERTION, "FP8 row bytes");
  if (a.n_pad <= 0) return;
  hipLaunchKernelGGL(fp8_prep_kernel, dim3((unsigned)ceil_div(a.n_pad, 4)), dim3(256), 0, st, a);
}

int fp8_scale_exp(float maxabs) {
  const double m = maxabs;
  if (!(m > 0.0)) return 0;
  int e = std::ilogb(kE4m3Max / m);
  while (std::ldexp(m, e) > kE4m3Max) --e;
  while (std::ldexp(m, e + 1) <= kE4m3Max) ++e;
  return e;
}

// FP8 tile configurations: WR x WC wavefronts of 32*WM rows x 32*WN queries.
struct Scan8Cfg {
  int wm, wn, wr, wc, stages;
};
static const Scan8Cfg kCfg8[] = {{2, 2, 4, 2, 3}, {2, 4, 4, 2, 3}, {4, 2, 2, 2, 3}, {2, 2, 2, 2, 3},
                                 {2, 2, 4, 2, 4}};
constexpr int kNumCfg8 = sizeof(kCfg8) / sizeof(kCfg8[0]);

int scan8_num_cfg() { return kNumCfg8; }
int scan8_tile_rows(int cfg) { return 32 * kCfg8[cfg].wm * kCfg8[cfg].wr; }
int scan8_tile_queries(int cfg) { return 32 * kCfg8[cfg].wn * kCfg8[cfg].wc; }

void launch_scan8(const ScanArgs& a, int grid, int kind, hipStream_t st) {
  if (grid <= 0) return;
  MANN_REQUIRE(((int64_t)a.n_rows + 1024) * 4 < ((int64_t)1 << 32), ERR_UNSUPPORTED,
               "too many rows for one first-pass launch");
  MANN_REQUIRE(a.kp % 128 == 0, ERR_ASSERTION, "FP8 row bytes must be a multiple of 128");
  switch (a.cfg) {
    case 1: launch8_t<2, 4, 4, 2, 3>(a, grid, kind, st); break;
    case 2: launch8_t<4, 2, 2, 2, 3>(a, grid, kind, st); break;
    case 3: launch8_t<2, 2, 2, 2, 3>(a, grid, kind, st); break;
    case 4: launch8_t<2, 2, 4, 2, 4>(a, grid, kind, st); break;
    default: launch8_t<2, 2, 4, 2, 3>(a, grid, kind, st); break;
  }
}

void launch_fp8_tau(const Fp8TauArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  hipLaunchKernelGGL(fp8_tau_kernel, dim3((unsigned)ceil_div(a.nq, 256)), dim3(256), 0, st, a);
}

int rescore_cap() { return RS_CAP; }

void launch_rescore(const RescoreArgs& a, int nq_max, hipStream_t st) {
  if (nq_max <= 0) return;
  MANN_REQUIRE(a.parts8 >= 1 && a.parts8 <= 64, ERR_ASSERTION, "FP8 candidate regions");
  MANN_REQUIRE(a.kp % 8 == 0, ERR_ASSERTION, "bf16 row length");
  static bool attr_set = false;  // dynamic LDS above 64 KiB must be opted into
  if (!attr_set) {
    MANN_HIP(hipFuncSetAttribute((const void*)rescore_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)sizeof(RescoreShared)));
    attr_set = true;
  }
  hipLaunchKernelGGL(rescore_kernel, dim3(nq_max), dim3(256), sizeof(RescoreShared), st, a);
}

}  // namespace mann
