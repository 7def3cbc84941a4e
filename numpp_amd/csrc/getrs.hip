// Triangular solves with packed LU factors (getrs) and pivot conversion.
//
// X = U^-1 L^-1 P B, following lu_solve (linalg.cpp:73-90): gather B by the
// row permutation, forward substitution with the unit-lower L, back
// substitution with U dividing by its diagonal.
//
// Each sweep is one launch with one 256-thread workgroup per (matrix, 64-row
// block, 8-column RHS chunk). Block i first accumulates the contributions of
// the blocks it depends on (L_ij X_j for j < i forward, U_ij X_j for j > i
// backward) as soon as each X_j is published, reading its 64-row strip of the
// factor coalesced, then solves its 64 x 64 diagonal block with one wave
// (lane = row, v_readlane broadcasts; U's diagonal is applied as a multiply by
// its reciprocal), publishes X_i and raises its flag
// (release at device scope; readers acquire). Workgroups are dispatched in
// index order and a block only waits for lower-indexed ones, so every wait
// ends; a bounded spin reports an error instead of hanging.
#include "common.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int GT = 256;
constexpr int GB = 64;   // rows per block
constexpr int GNR = 8;   // RHS columns per chunk
constexpr unsigned kSpinLimit = 1u << 22;

// Waits for a published block; gives up (and raises *err) after a bounded
// number of polls or as soon as another block has given up.
__device__ __forceinline__ bool wait_flag(const int* f, int* err) {
  unsigned spins = 0;
  // Poll relaxed (an acquire per poll would invalidate caches on every spin),
  // then acquire once the flag is seen.
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    __builtin_amdgcn_s_sleep(2);
    if (++spins > kSpinLimit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      atomicExch(err, 1);
      return false;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// Forward (LOWER = true, unit-lower L, blocks in increasing order, B gathered
// by perm) or backward (LOWER = false, upper U with division by the diagonal,
// blocks in decreasing order) sweep. X holds the forward result between sweeps.
template <class T, bool LOWER>
__global__ __launch_bounds__(GT) void sweep_kernel(int n, int nrhs, const T* LUb, int64_t strideA,
                                                   const int32_t* permb, const T* Bb, int64_t strideB, T* Xb,
                                                   int64_t strideX, int* flags, int* err, const int32_t* info,
                                                   const int32_t* mask) {
  const int nblk = (n + GB - 1) / GB;
  // x = matrix (fastest), y = block: the resident workgroups cover the early
  // blocks of every matrix, so all chains advance together.
  const int64_t mat = blockIdx.x;
  const int z = blockIdx.z;
  int* fl = flags + (mat * gridDim.z + z) * (int64_t)nblk;
  if ((info && info[mat] != 0) || (mask && mask[mat] == 0)) return;
  const int bi = LOWER ? (int)blockIdx.y : nblk - 1 - (int)blockIdx.y;
  const int r0 = bi * GB, bs = min(GB, n - r0);
  const int c0 = z * GNR, nc = min(GNR, nrhs - c0);
  const T* __restrict__ LU = LUb + mat * strideA;
  T* __restrict__ X = Xb + mat * strideX;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r = t >> 2, qq = t & 3;  // row in block, quarter of the 64 columns
  __shared__ T xs[GB][GNR];
  __shared__ T part[GB][GNR];
  __shared__ int ok_s;

  T acc[GNR];
#pragma unroll
  for (int c = 0; c < GNR; ++c) acc[c] = T(0);
  if (t == 0) ok_s = 1;

  for (int step = 0; step < nblk; ++step) {
    const int bj = LOWER ? step : nblk - 1 - step;
    if (LOWER ? bj >= bi : bj <= bi) break;
    if (t == 0 && !wait_flag(fl + bj, err)) ok_s = 0;
    __syncthreads();
    if (!ok_s) return;
    const int j0 = bj * GB, js = min(GB, n - j0);
    for (int e = t; e < GB * GNR; e += GT) {
      const int q = e / GNR, c = e - q * GNR;
      xs[q][c] = (q < js && c < nc) ? X[(int64_t)(j0 + q) * nrhs + c0 + c] : T(0);
    }
    __syncthreads();
    if (r < bs) {
      const T* lr = LU + (int64_t)(r0 + r) * n + j0 + qq * 16;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int qi = qq * 16 + q;
        const T l = qi < js ? lr[q] : T(0);
#pragma unroll
        for (int c = 0; c < GNR; ++c) acc[c] = acc[c] + l * xs[qi][c];
      }
    }
    __syncthreads();
  }
  // Combine the four quarter sums of each row.
#pragma unroll
  for (int c = 0; c < GNR; ++c) {
    acc[c] += __shfl_xor(acc[c], 1, 64);
    acc[c] += __shfl_xor(acc[c], 2, 64);
  }
  if (qq == 0) {
#pragma unroll
    for (int c = 0; c < GNR; ++c) part[r][c] = acc[c];
  }
  __syncthreads();
  // Diagonal block: wave 0, lane = row.
  if (wave == 0) {
    const int i = r0 + lane;
    T x[GNR];
    if (LOWER) {
      const int src = lane < bs ? permb[mat * n + i] : 0;
      const T* Bm = Bb + mat * strideB;
#pragma unroll
      for (int c = 0; c < GNR; ++c) x[c] = (lane < bs && c < nc) ? Bm[(int64_t)src * nrhs + c0 + c] - part[lane][c] : T(0);
    } else {
#pragma unroll
      for (int c = 0; c < GNR; ++c) x[c] = (lane < bs && c < nc) ? X[(int64_t)i * nrhs + c0 + c] - part[lane][c] : T(0);
    }
    T row[GB];
    const T* lr = LU + (int64_t)(lane < bs ? i : r0) * n + r0;
#pragma unroll
    for (int q = 0; q < GB; ++q) row[q] = (lane < bs && q < bs) ? lr[q] : T(0);
    if (LOWER) {
#pragma unroll
      for (int q = 0; q < GB; ++q) {
        const T f = lane > q ? row[q] : T(0);
#pragma unroll
        for (int c = 0; c < GNR; ++c) x[c] = x[c] - f * readlane<T>(x[c], q);
      }
    } else {
      T d = T(1);
#pragma unroll
      for (int q = 0; q < GB; ++q) d = (q == lane && lane < bs) ? row[q] : d;
      // one division per lane, in parallel; the row's scaling at its step is
      // then a multiply (within one rounding of lu_solve's division)
      const T dinv = T(1) / d;
#pragma unroll
      for (int q = GB - 1; q >= 0; --q) {
        if (lane == q) {
#pragma unroll
          for (int c = 0; c < GNR; ++c) x[c] = x[c] * dinv;
        }
        const T f = lane < q ? row[q] : T(0);
#pragma unroll
        for (int c = 0; c < GNR; ++c) x[c] = x[c] - f * readlane<T>(x[c], q);
      }
    }
    if (lane < bs) {
#pragma unroll
      for (int c = 0; c < GNR; ++c)
        if (c < nc) X[(int64_t)i * nrhs + c0 + c] = x[c];
    }
  }
  __syncthreads();
  if (t == 0) {
    __threadfence();
    __hip_atomic_store(fl + bi, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ void clear_flags_kernel(int* f, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = 0;
}

// Permutation from LAPACK-style pivots: one wave per matrix replays the swaps
// in LDS (used only by the standalone getrs entry point).
__global__ __launch_bounds__(64) void ipiv_to_perm_kernel(int n, const int32_t* ipivb, int32_t* permb) {
  const int64_t mat = blockIdx.x;
  const int32_t* ipiv = ipivb + mat * n;
  int32_t* perm = permb + mat * n;
  extern __shared__ int p_s[];
  for (int i = threadIdx.x; i < n; i += 64) p_s[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < n; ++k) {
      const int p = ipiv[k];
      if (p != k && p >= 0 && p < n) { const int tmp = p_s[k]; p_s[k] = p_s[p]; p_s[p] = tmp; }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 64) perm[i] = p_s[i];
}

}  // namespace

template <class T>
int blocked_getrs_perm(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm,
                       const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info, const int32_t* mask,
                       hipStream_t s) {
  if (count <= 0 || n <= 0 || nrhs <= 0) return NPPH_OK;
  if (B == X) { set_error("getrs: B and X must not alias"); return NPPH_ERR_INVALID; }
  if (count > 65535) { set_error("getrs: count too large"); return NPPH_ERR_INVALID; }
  const int nblk = (int)ceil_div(n, GB), nz = (int)ceil_div(nrhs, GNR);
  const int64_t nflag = count * nz * (int64_t)nblk;
  int* flags;
  NPPH_TRY(workspace(kWsFlags, sizeof(int) * (size_t)(2 * nflag + 1), (void**)&flags));
  int* err = flags + 2 * nflag;
  hipLaunchKernelGGL(clear_flags_kernel, dim3((unsigned)ceil_div(2 * nflag + 1, 256)), dim3(256), 0, s, flags,
                     2 * nflag + 1);
  dim3 grid((unsigned)count, (unsigned)nblk, (unsigned)nz);
  ProfScope ps(kProfGetrs, 2.0 * count * n * n * nrhs,
               (double)count * (sizeof(T) * (double)n * n + 3.0 * sizeof(T) * n * nrhs), s);
  hipLaunchKernelGGL((sweep_kernel<T, true>), grid, dim3(GT), 0, s, n, nrhs, LU, strideA, perm, B, strideB, X,
                     strideX, flags, err, info, mask);
  hipLaunchKernelGGL((sweep_kernel<T, false>), grid, dim3(GT), 0, s, n, nrhs, LU, strideA, perm, B, strideB, X,
                     strideX, flags + nflag, err, info, mask);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

template int blocked_getrs_perm<double>(int64_t, int, int, const double*, int64_t, const int32_t*, const double*,
                                        int64_t, double*, int64_t, const int32_t*, const int32_t*, hipStream_t);
template int blocked_getrs_perm<float>(int64_t, int, int, const float*, int64_t, const int32_t*, const float*,
                                       int64_t, float*, int64_t, const int32_t*, const int32_t*, hipStream_t);

int ipiv_to_perm(int64_t count, int n, const int32_t* ipiv, int32_t* perm, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  hipLaunchKernelGGL(ipiv_to_perm_kernel, dim3((unsigned)count), dim3(64), (size_t)n * sizeof(int), s, n, ipiv,
                     perm);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace npph
