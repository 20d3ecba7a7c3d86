// Declarations of the small device utilities in kernels_util.hip.
#pragma once

#include "common.h"

namespace mann {

void fill_i32(int* p, int64_t n, int v, hipStream_t st);
void fill_f32(float* p, int64_t n, float v, hipStream_t st);
void reset_listed(int* cnt, const int* list, const int* count, int max_n, hipStream_t st);
void gather_rows_bf16(const bf16_t* src, const float* srcn, const int* idx, int m, int m_pad,
                      int kp, bf16_t* dst, float* dstn, hipStream_t st);
void gather_rows_f32(const float* src, const int* idx, int m, int w, float* dst, hipStream_t st);
void gather_f32(const float* src, const int* idx, int m, float* dst, hipStream_t st);
void gather_rows_i32(const int* src, const int* idx, int m, int w, int* dst, hipStream_t st);
// 1-based int64 labels (0 = none) -> 0-based int32 (-1 = none)
void labels_to_i32(const int64_t* src, int64_t n, int* dst, hipStream_t st);

void ids_to_i64(const int* src, int64_t n, int64_t* dst, hipStream_t st);
// binary16 rows (stride lds) -> binary32 rows (stride ldd), first w columns
void f16_rows_to_f32(const void* src, int64_t lds, float* dst, int64_t ldd, int w, int64_t rows,
                     hipStream_t st);
// dst[r][c] = src[r][c] for c < w, 0 for w <= c < ldd
void copy_rows(const float* src, int64_t lds, float* dst, int64_t ldd, int w, int64_t rows,
               hipStream_t st);

// dst row r = src row r / |src row r| (first w of ld columns; the rest zero)
void unit_rows(const float* src, float* dst, int64_t ld, int w, int64_t rows, hipStream_t st);

int select_cap_max();
int select_exmax();   // exactly re-ranked candidates per query
int exact_kmax();

}  // namespace mann
