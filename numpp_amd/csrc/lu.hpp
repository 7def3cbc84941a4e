// Internal interfaces of the LU / solve kernels.
#pragma once

#include "common.hpp"

namespace npph {

// One small system for the wave-per-matrix kernel (row-major, leading dims).
struct SmallSys {
  int n;
  const double* A;
  int64_t lda;
  double* LU;      // nullable: factors out (may alias A)
  int64_t ldlu;
  int32_t* ipiv;   // nullable
  int32_t* info;   // nullable
  int nrhs;
  const double* B; // nullable when nrhs == 0
  int64_t ldb;
  double* X;       // nullable: solution out (may alias B)
  int64_t ldx;
};

// A uniform stack of count systems of order n.
struct UniformBatch {
  int64_t count;
  int n;
  const double* A;
  int64_t strideA;
  double* LU;
  int32_t* ipiv;   // count x n
  int32_t* info;   // count
  int nrhs;
  const double* B;
  int64_t strideB;
  double* X;
};

// n <= 64: factor (and solve when nrhs > 0) one matrix per wavefront.
int small_lu_launch(const UniformBatch& u, hipStream_t s);

// Largest order the blocked LU supports (panel rows held in registers).
int max_blocked_n();

// Blocked right-looking LU in working precision T (double or float) of a
// uniform stack stored in place (count x n x n, element stride strideA).
// ipiv: count x n (0-based pivot positions), perm: count x n (original row at
// each final position), info: count (0 or first zero-pivot step + 1).
// bf16_terms > 0 selects the split-BF16 MFMA trailing update (T = float only).
// mask (nullable, count): matrices with mask == 0 are skipped.
template <class T>
int blocked_getrf_perm(int64_t count, int n, T* A, int64_t strideA, int32_t* ipiv, int32_t* perm,
                       int32_t* info, int bf16_terms, const int32_t* mask, hipStream_t s);

// X = U^-1 L^-1 P B with packed factors in T and the permutation from the
// factorisation. B and X must not alias. Systems with info != 0 (when info is
// non-null) or mask == 0 (when mask is non-null) are skipped.
template <class T>
int blocked_getrs_perm(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm,
                       const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info,
                       const int32_t* mask, hipStream_t s);

// perm from LAPACK-style pivots (replays the swaps).
int ipiv_to_perm(int64_t count, int n, const int32_t* ipiv, int32_t* perm, hipStream_t s);

}  // namespace npph
