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
// Bytes of workspace one blocked factorisation of count matrices of order n
// with elem-byte entries requests, each request rounded up to 256 bytes.
size_t blocked_getrf_scratch_bytes(int64_t count, int n, size_t elem);

// Blocked right-looking LU in working precision T (double or float) of a
// uniform stack stored in place (count x n x n, element stride strideA).
// Rows never move: on return logical row p of L\U is physical row perm[p]
// (perm: count x n, the original row at each final position); ipiv: count x n
// (0-based pivot positions), info: count (0 or first zero-pivot step + 1).
// bf16_terms > 0 selects the split-BF16 MFMA trailing update (T = float only).
// mask (nullable, count): matrices with mask == 0 are skipped.
template <class T>
int blocked_getrf_perm(int64_t count, int n, T* A, int64_t strideA, int32_t* ipiv, int32_t* perm,
                       int32_t* info, int bf16_terms, const int32_t* mask, hipStream_t s);
// Packed row order (row p = physical row perm[p]) of factors from
// blocked_getrf_perm, in place.
template <class T>
int unscramble_rows(int64_t count, int n, T* A, int64_t strideA, const int32_t* perm, hipStream_t s);
// Factors of the refined solver keep the blocked LU's row map, except FP64
// factors of order <= 64 (the one-wave kernel writes them packed).
inline bool factor_rows_mapped(int precision, int n) { return !(precision == NPPH_PREC_FP64 && n <= 64); }

// X = U^-1 L^-1 P B with factors in T and the permutation from the
// factorisation; logical factor row p is physical row rowmap[p] (rowmap =
// perm for blocked_getrf_perm's output) or row p when rowmap is null (packed
// factors). B and X must not alias. Systems with info != 0 (when info is
// non-null) or mask == 0 (when mask is non-null) are skipped.
template <class T>
int blocked_getrs_perm(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm,
                       const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info,
                       const int32_t* mask, hipStream_t s, const int32_t* rowmap = nullptr);

// One refinement correction with FP32 factors: D = U^-1 L^-1 P fp32(R * scale)
// (per-column scale), then X = D / scale (first) or X += D / scale, in two
// sweeps; Y (count x n x nrhs FP32) holds the forward sweep's result. Only
// systems with mask == 1 (still iterating) are corrected. live, when given,
// is a device count of those systems: at zero the sweeps return at once.
int getrs_refine(int64_t count, int n, int nrhs, const float* LU, int64_t strideA, const int32_t* perm,
                 const double* R, const double* scale, float* Y, double* X, int first, const int32_t* mask,
                 hipStream_t s, const int32_t* rowmap, const int32_t* live = nullptr);

// Sweeps with many workgroups per system (getrs_mw.hip): wanted for stacks of
// few systems of order >= 512, where one workgroup per system and column chunk
// would leave most CUs idle. Same contract as blocked_getrs_perm (REF: the
// refinement correction of getrs_refine, with R, scale and the FP64 iterate
// X64; X then holds the FP32 forward result and correction).
bool getrs_mw_wanted(int64_t count, int n, int nrhs);
template <class T, bool REF>
int getrs_mw(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm, const int32_t* rowmap,
             const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info, const int32_t* mask,
             const double* R, const double* scale, double* X64, int first, const int32_t* live, hipStream_t s);

// perm from LAPACK-style pivots (replays the swaps).
int ipiv_to_perm(int64_t count, int n, const int32_t* ipiv, int32_t* perm, hipStream_t s);

// Factors of a uniform stack for the refined solver (device memory carved from
// one buffer by refine_factor_carve). lu64 is used for FP64 precision, lu32
// for the FP32 / split-BF16 precisions.
struct RefineFactors {
  int precision = NPPH_PREC_FP64;
  int64_t count = 0;
  int n = 0;
  double* lu64 = nullptr;
  float* lu32 = nullptr;
  int32_t* ipiv = nullptr;
  int32_t* perm = nullptr;
  int32_t* info = nullptr;
  double* anorm = nullptr;
};
size_t refine_factor_bytes(int precision, int64_t count, int n);
void refine_factor_carve(int precision, int64_t count, int n, void* buf, RefineFactors& f);
// Factorises A (count x n x n FP64, unchanged) into f in f.precision.
// prepped: the caller has already written A's working copy (f.lu64 or the
// FP32 rounding in f.lu32) and ||A||_inf per system into f.anorm.
int refine_factor(const npph_refine_opts& o, const double* A, RefineFactors& f, hipStream_t s,
                  bool prepped = false);
// Solves with resident factors, refining in FP64 against A (needed for the
// residual) until every column meets o.tol; B and X must not alias.
int refine_solve(const npph_refine_opts& o, const double* A, const RefineFactors& f, int nrhs, const double* B,
                 double* X, double* berr, int32_t* iters_out, int32_t* info_out, hipStream_t s);
// One refined solve as resumable steps, for callers that issue the sweeps of
// several stacks before waiting on any (refine_solve = begin, loop, end):
//   refine_begin  binds the bookkeeping (the bound workspace set's kWsNorm,
//                 kWsR, kWsD slots) and issues sweep 0 (FP64: the whole solve);
//   refine_steps  issues `budget` test + correction iterations and the next
//                 test, with no host read-back (capturable in a graph);
//   refine_loop   continues with a host read-back per iteration until no
//                 system iterates; n_fb = systems stopped without converging;
//   refine_end    FP64 fallback of those (when enabled and n_fb > 0) and the
//                 outputs (with n_fb = 0: the outputs only; callable again).
// The device state lives in workspace slots: nothing else may use the same
// set between the steps of one run.
// Matrices of a bucket kept at their own orders (n <= 256): slot k holds
// system ids[k], whose nn[id] x nn[id] matrix starts at A + aoff[id]; the
// padded system is diag(A, I).
struct PackedA {
  const int32_t* ids = nullptr;
  const int32_t* nn = nullptr;
  const int64_t* aoff = nullptr;
  const double* A = nullptr;
};

struct RefineRun {
  npph_refine_opts o{};
  const double* A = nullptr;  // uniform padded stack, unless var.A is set
  PackedA var;
  RefineFactors f;
  int nrhs = 0;
  const double* B = nullptr;
  double* X = nullptr;
  double* berr_out = nullptr;
  int32_t* iters_out = nullptr;
  int32_t* info_out = nullptr;
  // set by refine_begin
  int64_t count = 0;
  int n = 0;
  double *bmax = nullptr, *rmax = nullptr, *xmax = nullptr, *scale = nullptr, *berr_p = nullptr, *prev = nullptr;
  double* R = nullptr;
  float* Dlo = nullptr;
  int32_t *info = nullptr, *active = nullptr, *iters = nullptr, *mask = nullptr, *ctr = nullptr;
  int it = 0;               // next loop iteration
  bool eta_issued = false;  // the test of iteration `it` is already queued
  bool done = false;
};
int refine_begin(RefineRun& r, hipStream_t s);
int refine_steps(RefineRun& r, int budget, hipStream_t s);
int refine_loop(RefineRun& r, int32_t* n_fb, hipStream_t s);
int refine_end(RefineRun& r, int32_t n_fb, hipStream_t s);
int refined_impl(const npph_refine_opts& o, int64_t count, int n, int nrhs, const double* A, const double* B,
                 double* X, double* berr, int32_t* iters_out, int32_t* info_out, hipStream_t s);
int gesv_dev_impl(int64_t count, int n, int nrhs, double* A, int64_t strideA, double* B, int64_t strideB,
                  int32_t* ipiv, int32_t* info, hipStream_t s);

}  // namespace npph
