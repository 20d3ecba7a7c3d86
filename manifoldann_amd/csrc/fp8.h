// FP8 (OCP E4M3FN) first pass of the exact search: operand preparation, the
// MFMA distance scan, the threshold conversion from the bf16 sample and the
// bf16 re-score of the FP8 survivors.
//
// Operands: x^ = e4m3(x * 2^s) with a power-of-two scale (one per dataset,
// one per query), so that x~ = x^ * 2^-s exactly and the MFMA products of the
// scaled operands are exact; |x~|^2, |x - x~| and |x~| come from FP64 sums like
// the bf16 copy's.
#pragma once

#include "launch.h"

namespace mann {

// Largest finite E4M3FN value.
constexpr double kE4m3Max = 448.0;

// Error coefficient of the FP8 first pass in the squared-distance domain:
// |S^ - |q~ - x~|^2| <= kFp8G * (|q~| + |x~|)^2 (MFMA accumulation of the FP8
// products plus the FP32 roundings of the epilogue).
constexpr double kFp8G = 1.0 / 512.0;

// Row preparation.  scale_exp >= -1000: every row uses 2^scale_exp (a
// dataset); else each row picks the largest s with max|t| 2^s <= 448 (queries)
// and rowmul[row] = -2 * 2^-(s + other_exp) multiplies its accumulators.
// t = x (normalised when normalize, minus center when given), from the FP32
// rows x32 (d16 floats each).  Outputs: x8 (kp8 bytes per row, zero padded),
// xn8 = |x~|^2 (+inf on padding rows n .. n_pad), err8 = |t - x~| and nrm8 =
// |x~| rounded up, their maxima by atomicMax on the float bits.
struct Fp8PrepArgs {
  const float* x32;
  int d16, d, kp8;
  int64_t n, n_pad;
  int normalize;
  const float* center;
  int scale_exp;      // dataset scale, or -1001 for per-row scales
  int other_exp;      // the other operand's scale (per-row mode)
  unsigned char* x8;
  float* xn8;
  float* err8;
  float* nrm8;
  float* rowmul;      // per-row mode: -2 * 2^-(s + other_exp)
  unsigned* max_err;
  unsigned* max_nrm;
};
// max |t| over the rows (atomicMax of the float bits into *out)
void launch_fp8_maxabs(const Fp8PrepArgs& a, unsigned* out, hipStream_t st);
void launch_fp8_prep(const Fp8PrepArgs& a, hipStream_t st);
// the scale exponent s for max|t| = m: largest s with m 2^s <= 448
int fp8_scale_exp(float maxabs);

// First pass in FILTER mode over the FP8 operands (ScanArgs: x8, q8, qmul,
// xn = xn8, qnorm = qn8, kp = row bytes, a multiple of 128).
int scan8_num_cfg();
int scan8_tile_rows(int cfg);
int scan8_tile_queries(int cfg);
void launch_scan8(const ScanArgs& a, int grid, int kind, hipStream_t st);

// tau8[q] = (sqrt(max(tau16[q] + G16, 0)) + qe16[q] + ex16 + qe8[q] + ex8)^2 + G8:
// every row whose bf16 first-pass value is at most tau16 has an FP8
// first-pass value below tau8 (an emission target carried across domains).
struct Fp8TauArgs {
  int nq, d;
  const float* tau16;
  const float* qe16;
  const float* qr16;
  float ex16, rx16;
  const float* qe8;
  const float* qr8;
  float ex8, rx8;
  float* tau8;
};
void launch_fp8_tau(const Fp8TauArgs& a, hipStream_t st);

// Re-score of the FP8 survivors on the bf16 operands: for each query the
// candidates of all regions get S^ of the bf16 vectors (FP32 sums of exact
// products, any order: within the bf16 bound), and the `keep` smallest are
// written to one region per query for the selection; dropmin[q] is a lower
// bound of the dropped values (+inf when none was dropped), and cnt[q] = keep
// + 1 marks an overflow of the FP8 regions.
struct RescoreArgs {
  const int* qlist;
  const int* nq_dev;
  int nq;
  const unsigned* cand8;
  const int* cnt8;
  int cap8, parts8;
  const bf16_t* xb;
  const float* xn;
  const bf16_t* qb;
  const float* qn;
  int kp, d;
  unsigned* cand;
  int* cnt;
  int keep;
  float* dropmin;
  int* stats;         // [kStatSlots][4]: slot 2 += FP8 candidates read
};
int rescore_cap();    // FP8 candidates per query the re-score holds
void launch_rescore(const RescoreArgs& a, int nq_max, hipStream_t st);

}  // namespace mann
