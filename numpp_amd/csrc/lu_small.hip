// Batched LU with partial pivoting + getrs for n <= 64: one wavefront per
// matrix, the matrix resident in registers.
//
// Lane r owns original row r (N doubles, N = n rounded up to 8/16/32/64; the
// padding is the identity, which leaves the leading n x n factorisation, its
// pivots and its exact-zero singularity test unchanged). Rows are never moved
// between lanes: each lane tracks the logical position `pos` of its row, and a
// step-k swap of positions k and p (linalg.cpp:58-62) just exchanges two
// lanes' `pos` values. The pivot search is a wave-wide argmax over
// (|a[k]|, pos) with NumPP's tie rule (largest magnitude, then smallest current
// position: linalg.cpp:55-56; a NaN at position k is kept, a NaN below it is
// never chosen), the pivot row is broadcast with v_readlane into
// scalar registers, and each lane applies the rank-1 update to its own row
// (linalg.cpp:64-68, division by the pivot as in the reference). The solve
// walks the same lanes: forward substitution with the unit-lower L, then back
// substitution dividing by U's diagonal (linalg.cpp:73-90).
#include "common.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int NR = 8;  // right-hand-side columns held in registers per pass

template <int N>
__device__ void lu_wave(const SmallSys& s) {
  const int lane = threadIdx.x & 63;
  const int n = s.n;
  const bool act = lane < N;
  double a[N];
  const double* Ar = s.A + (int64_t)lane * s.lda;
#pragma unroll
  for (int j = 0; j < N; ++j) a[j] = (lane < n && j < n) ? Ar[j] : (lane == j ? 1.0 : 0.0);

  int pos = act ? lane : 1 << 20;
  int sing = -1;
  // plane[k]: lane holding logical row k (kept in LDS so the solve can index it).
  __shared__ int plane[64];

  for (int k = 0; k < N; ++k) {
    double ak = a[0];
#pragma unroll
    for (int j = 1; j < N; ++j) ak = (j == k) ? a[j] : ak;
    double v = (act && pos >= k) ? piv_key(ak, pos, k) : -1.0;
    int p = pos;
    wave_argmax(v, p);
    p = __builtin_amdgcn_readfirstlane(p);
    if (v == 0.0 && k < n) { sing = k; break; }
    const uint64_t m = __ballot(pos == p);
    const int wl = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(m));
    if (s.ipiv && lane == 0 && k < n) s.ipiv[k] = p;
    if (lane == 0) plane[k] = wl;
    const bool is_k = (pos == k);
    if (lane == wl) pos = k;
    else if (is_k) pos = p;
    const double piv = readlane_d(ak, wl);
    const double l = (pos > k) ? ak / piv : 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double u = readlane_d(a[j], wl);
      u = (j > k) ? u : 0.0;
      a[j] = fma(-l, u, a[j]);
      if (j == k && pos > k) a[j] = l;
    }
  }
  __syncthreads();  // one-wave block: orders lane 0's plane[] writes before the solve reads
  if (s.info && lane == 0) *s.info = sing < 0 ? 0 : sing + 1;
  if (s.LU && lane < n && pos < n) {
    double* o = s.LU + (int64_t)pos * s.ldlu;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (j < n) o[j] = a[j];
  }
  if (sing >= 0 || s.nrhs <= 0 || !s.X) return;

  for (int c0 = 0; c0 < s.nrhs; c0 += NR) {
    double b[NR];
    const double* Br = s.B + (int64_t)lane * s.ldb + c0;
#pragma unroll
    for (int q = 0; q < NR; ++q) b[q] = (lane < n && c0 + q < s.nrhs) ? Br[q] : 0.0;
    // forward: unit-lower L, logical row order 0..N-1
    for (int i = 0; i < N; ++i) {
      const int wl = plane[i];
      double li = a[0];
#pragma unroll
      for (int j = 1; j < N; ++j) li = (j == i) ? a[j] : li;
      li = (pos > i) ? li : 0.0;
#pragma unroll
      for (int q = 0; q < NR; ++q) b[q] = fma(-li, readlane_d(b[q], wl), b[q]);
    }
    // back: upper U with division by the diagonal, rows N-1..0
    for (int i = N - 1; i >= 0; --i) {
      const int wl = plane[i];
      double ui = a[0];
#pragma unroll
      for (int j = 1; j < N; ++j) ui = (j == i) ? a[j] : ui;
      if (pos == i) {
#pragma unroll
        for (int q = 0; q < NR; ++q) b[q] = b[q] / ui;
      }
      const double f = (pos < i) ? ui : 0.0;
#pragma unroll
      for (int q = 0; q < NR; ++q) b[q] = fma(-f, readlane_d(b[q], wl), b[q]);
    }
    if (lane < n && pos < n) {
      double* Xr = s.X + (int64_t)pos * s.ldx + c0;
#pragma unroll
      for (int q = 0; q < NR; ++q)
        if (c0 + q < s.nrhs) Xr[q] = b[q];
    }
  }
}

// Uniform stack: block = 1 wave, blockIdx.x = matrix.
template <int N>
__global__ __launch_bounds__(64) void lu_wave_uniform(UniformBatch u) {
  const int64_t i = blockIdx.x;
  SmallSys s;
  s.n = u.n;
  s.A = u.A + i * u.strideA;
  s.lda = u.n;
  s.LU = u.LU ? u.LU + i * u.strideA : nullptr;
  s.ldlu = u.n;
  s.ipiv = u.ipiv ? u.ipiv + i * u.n : nullptr;
  s.info = u.info ? u.info + i : nullptr;
  s.nrhs = u.nrhs;
  s.B = u.B ? u.B + i * u.strideB : nullptr;
  s.ldb = u.nrhs;
  s.X = u.X ? u.X + i * u.strideB : nullptr;
  s.ldx = u.nrhs;
  lu_wave<N>(s);
}

}  // namespace

int small_lu_launch(const UniformBatch& u, hipStream_t st) {
  if (u.count <= 0 || u.n <= 0) return NPPH_OK;
  dim3 grid((unsigned)u.count), block(64);
  if (u.count > 0x7fffffff) { set_error("batch too large"); return NPPH_ERR_INVALID; }
  const double n = u.n, flu = n * (n - 1) / 2 + (n - 1) * n * (2 * n - 1) / 3;
  ProfScope ps(kProfSmallLu, u.count * (flu + (2 * n * n - n) * u.nrhs),
               8.0 * u.count * (n * n + 2 * n * u.nrhs + (u.LU ? n * n : 0)), st);
  if (u.n <= 8) hipLaunchKernelGGL(lu_wave_uniform<8>, grid, block, 0, st, u);
  else if (u.n <= 16) hipLaunchKernelGGL(lu_wave_uniform<16>, grid, block, 0, st, u);
  else if (u.n <= 32) hipLaunchKernelGGL(lu_wave_uniform<32>, grid, block, 0, st, u);
  else if (u.n <= 64) hipLaunchKernelGGL(lu_wave_uniform<64>, grid, block, 0, st, u);
  else { set_error("small_lu: n=%d > 64", u.n); return NPPH_ERR_INVALID; }
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace npph
