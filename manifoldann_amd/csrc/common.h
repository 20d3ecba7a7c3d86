// Shared device/host definitions for the ManifoldANN MI355X (gfx950) search core.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

namespace mann {

// ---- vector types used by the gfx950 MFMA builtins --------------------------
typedef __bf16 bf16_t;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int intx4;

// Status codes of the C ABI (mirrors the reference's exception classes).
enum Status : int {
  OK = 0,
  ERR_DIMENSION = -1,  // DimensionMismatch
  ERR_ARGUMENT = -2,   // ArgumentError
  ERR_ASSERTION = -3,  // AssertionError
  ERR_DEVICE = -4,     // HIP runtime failure
  ERR_ALLOC = -5,      // device allocation failure
  ERR_NULL = -6,       // null handle / pointer
  ERR_UNSUPPORTED = -7
};

// Metric ids (utils/distances.jl:15-18).
enum Metric : int { EUCLIDEAN = 0, SQEUCLIDEAN = 1, COSINE = 2 };

void set_error(const std::string& msg);

struct Error {
  int code;
  std::string msg;
};

#define MANN_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::mann::Error{::mann::ERR_DEVICE, std::string(#call) + ": " +               \
                                                  hipGetErrorString(e_)};               \
  } while (0)

#define MANN_REQUIRE(cond, code, msg)                  \
  do {                                                 \
    if (!(cond)) throw ::mann::Error{(code), (msg)};   \
  } while (0)

static inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
static inline int64_t ceil_div(int64_t x, int64_t m) { return (x + m - 1) / m; }

// ---- device helpers ------------------------------------------------------------

// Monotone map float -> uint32 (total order consistent with '<' for non-NaN).
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(u);
}
__device__ __forceinline__ uint64_t pack_key(float f, uint32_t r) {
  return ((uint64_t)f2ord(f) << 32) | (uint64_t)r;
}

// Sum over a 16-lane group with the fixed tree (xor 8, 4, 2, 1): lane j of the
// group ends with ((a0+a8)+(a4+a12))+... exactly the A16 tree of the oracle.
// LDS writes of a wavefront visible to its own later reads (other lanes)
__device__ __forceinline__ void wave_fence_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP row rotate right by S within each 16-lane row: lane i reads lane (i - S) mod 16
template <int S>
__device__ __forceinline__ float row_ror(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + S, 0xF, 0xF,
                                                    false));
}
// Steps +8, +4, +2, +1 by row rotations: after the step of S the values are
// S-periodic within the row, so lane i adds the same pair as lane i ^ S would
// (a + b == b + a exactly) and every lane ends with the tree's sum.
__device__ __forceinline__ float group16_tree(float v) {
  v = v + row_ror<8>(v);
  v = v + row_ror<4>(v);
  v = v + row_ror<2>(v);
  v = v + row_ror<1>(v);
  return v;
}

}  // namespace mann
