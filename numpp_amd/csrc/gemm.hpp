// Internal GEMM launcher shared by matmul and the blocked LU trailing update.
#pragma once

#include <cstring>

#include "common.hpp"

namespace npph {

// Row indirection of a GEMM's operands: row r of A (B, C) in matrix z is the
// physical row a[z * stride + r] (b, c) of that operand; a null map is the
// identity. The recursive LU, whose rows never move, reaches position ranges
// through its permutation this way.
struct RowMaps {
  const int32_t* a = nullptr;
  const int32_t* b = nullptr;
  const int32_t* c = nullptr;
  int64_t stride = 0;
};

// Row-major C = alpha * A(MxK) * B(KxN) + beta * C over `batch` matrices with
// element strides sA / sB / sC between consecutive matrices.
template <class T>
struct GemmArgs {
  int64_t M, N, K;
  T alpha, beta;
  const T* A;
  int64_t lda, sA;
  const T* B;
  int64_t ldb, sB;
  T* C;
  int64_t ldc, sC;
  int tiles_m, tiles_n, vec;  // filled by gemm_launch
  RowMaps maps{};
};

// prof: the npph_prof class the launch is timed under; -1 = by element type,
// -2 = not timed on its own (the caller's scope covers it).
template <class T>
int gemm_launch(GemmArgs<T> p, int64_t batch, hipStream_t s, int prof = -1);

// Split-BF16 update C -= A * B for FP32 matrices (the low-precision LU's
// trailing update): each operand is split into one, two or three BF16 pieces
// after a power-of-two scaling per panel and the partial products are
// accumulated in FP32 on the BF16 MFMA. terms: 1 (plain BF16), 3 (two-piece
// split, 3 products) or 6 (three-piece split, 6 products).
int gemm_bf16split_update(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, int64_t sA,
                          const float* B, int64_t ldb, int64_t sB, float* C, int64_t ldc,
                          int64_t sC, int64_t batch, int terms, hipStream_t s, const RowMaps& maps = RowMaps{});

}  // namespace npph
