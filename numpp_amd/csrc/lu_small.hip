// Batched LU with partial pivoting + getrs for n <= 64: one workgroup per
// matrix, the matrix resident in registers.
//
// The matrix is padded to N = 8, 16, 32 or 64 with the identity (which leaves
// the leading n x n factorisation, its pivots and its exact-zero singularity
// test unchanged) and split by columns over NW = N/16 waves (one wave for
// N <= 16; eight waves for N = 64 with one right-hand side): lane r of every
// wave holds original row r, wave w the columns [w CPW, (w+1) CPW). Rows are never moved: each wave tracks the logical
// position `pos` of every lane's row, and the step-k swap of positions k and p
// (linalg.cpp:58-62) exchanges two lanes' `pos` values.
//
// The k loop is unrolled at compile time, so column k is a static register and
// every wave knows statically whether it owns column k and which of its columns
// lie right of k. Step k:
//   - the owner wave of column k finds the pivot with a wave-wide argmax over
//     (|a_k|, pos) with NumPP's rule (largest magnitude, then smallest current
//     position, linalg.cpp:55-56; a NaN at position k is kept, a NaN below it is
//     never chosen), divides the column by the pivot (linalg.cpp:65-66) and
//     publishes the multipliers, the pivot position and the pivot lane in LDS
//     (lane -1 for an exact-zero pivot: the factorisation stops there);
//   - one workgroup barrier (the LDS slots alternate between even and odd
//     steps, so the next owner never overwrites what a slower wave still reads);
//   - every wave with columns right of k updates them (linalg.cpp:67), reading
//     the pivot row from its own lane `wl` with v_readlane.
// The factors go to HBM (row order = logical position) and to LDS, where the
// solve reads them: forward substitution with the unit-lower L, then back
// substitution with U (linalg.cpp:73-90), lane = logical row, waves taking
// right-hand-side chunks in turn.
#include <cstdlib>
#include <utility>

#include "common.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

namespace {


template <int N, int WV = (N >= 32 ? N / 16 : 1)>
struct Small {
  static constexpr int NW = WV;                    // waves per matrix
  static constexpr int CPW = N / NW;               // columns per wave
  static constexpr int NT = 64 * NW;
};

// Shared state of one workgroup.
template <int N>
struct SmallShared {
  double l[2][64];       // multipliers of the step (by lane), even / odd steps
  int p[2], wl[2];       // pivot position and pivot lane of the step
  int piv[N];            // pivot rows, written to HBM after the loop (a global
                         // store inside it would make each step's barrier wait)
  int sing;              // first step with an exact-zero pivot, or -1
  double L[N][N + 1];    // strictly lower factor by logical row (zeros elsewhere)
  double U[N][N + 1];    // strictly upper factor by logical row, column k / u_kk
  double D[N];           // diagonal of U
  double Dr[N];          // its reciprocals
  double y[N];           // L^-1 P b by logical row (single right-hand side)
  int perm[N];           // original row at each logical position
};

// Factorisation, as seen by wave W (all arguments uniform across the wave).
// AUG: the last wave also carries b as an extra column (ab, by physical row),
// so the forward substitution happens inside the elimination.
template <int N, int WV, int W, bool AUG>
__device__ __forceinline__ void factor(SmallShared<N>& sh, double (&a)[Small<N, WV>::CPW], int& pos, int n, int lane,
                                       double& ab) {
  constexpr int CPW = Small<N, WV>::CPW;
  bool stop = false;
  auto step = [&]<int K>() __attribute__((always_inline)) {
    if (stop || K >= n) return;  // uniform over the workgroup (every wave stops at the same step)
    constexpr int O = K / CPW, CO = K % CPW, BUF = K & 1;
    double l = 0.0;
    int wl = 0, p = K;
    if constexpr (W == O) {
      double v = (pos >= K && pos < N) ? piv_key(a[CO], pos, K) : -1.0;
      p = pos;
      wave_argmax(v, p);
      p = __builtin_amdgcn_readfirstlane(p);
      const uint64_t m = __ballot(pos == p);
      wl = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(m));
      const double piv = readlane_d(a[CO], wl);
      if (pos == p) pos = K;
      else if (pos == K) pos = p;
      if (pos > K && pos < N) {
        l = a[CO] / piv;
        a[CO] = l;
      }
      sh.l[BUF][lane] = l;
      if (lane == 0) {
        sh.p[BUF] = p;
        sh.wl[BUF] = v == 0.0 ? -1 : wl;
        sh.piv[K] = p;
        if (v == 0.0) sh.sing = K;
      }
      __syncthreads();
      if (v == 0.0) {
        stop = true;
        return;
      }
    } else {
      __syncthreads();
      p = sh.p[BUF];
      wl = sh.wl[BUF];
      if (wl < 0) {
        stop = true;
        return;
      }
      if (pos == p) pos = K;
      else if (pos == K) pos = p;
      l = sh.l[BUF][lane];
    }
    // columns of this wave right of the pivot column
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      if (W * CPW + c > K) {
        const double u = readlane_d(a[c], wl);
        a[c] = fma(-l, u, a[c]);
      }
    }
    if constexpr (AUG && W == Small<N, WV>::NW - 1) ab = fma(-l, readlane_d(ab, wl), ab);
  };
  [&]<int... K>(std::integer_sequence<int, K...>) __attribute__((always_inline)) {
    (step.template operator()<K>(), ...);
  }(std::make_integer_sequence<int, N>{});
}

// NR: right-hand-side columns held in registers per pass (1 for a single
// right-hand side, which then rides along the elimination as an extra column,
// else 8).
template <int N, int WV, int W, int NR>
__device__ __forceinline__ void wave_body(const UniformBatch& u, SmallShared<N>& sh, int64_t mat, int lane) {
  constexpr int CPW = Small<N, WV>::CPW, NW = Small<N, WV>::NW;
  const int n = u.n;
  double a[CPW];
  {
    const double* Ar = u.A + mat * u.strideA + (int64_t)(lane < n ? lane : 0) * n;
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const int j = W * CPW + c;
      const double v = Ar[j < n ? j : 0];
      a[c] = (lane < n && j < n) ? v : (lane == j ? 1.0 : 0.0);
    }
  }
  constexpr bool AUG = NR == 1;
  double ab = 0.0;
  if (AUG && W == NW - 1 && u.nrhs > 0 && u.X && lane < n) ab = u.B[mat * u.strideB + lane];
  int pos = lane < N ? lane : (1 << 20);
  factor<N, WV, W, AUG>(sh, a, pos, n, lane, ab);
  const bool singular = sh.sing >= 0;
  if (W == 0 && lane == 0 && u.info) u.info[mat] = singular ? sh.sing + 1 : 0;
  if (W == 0 && u.ipiv) {
    // steps after an exact-zero pivot were not run: their pivot rows stay
    // unspecified, as the reference stops there
    const int ks = singular ? sh.sing + 1 : n;
    if (lane < ks) u.ipiv[mat * n + lane] = sh.piv[lane];
  }
  if (u.LU && lane < n && pos < n) {
    double* o = u.LU + mat * u.strideA + (int64_t)pos * n + W * CPW;
    if (n == N && ((reinterpret_cast<uintptr_t>(u.LU) | (uintptr_t)u.strideA * 8) & 15) == 0) {
      // full rows: this lane's CPW columns as 16-byte vectors
#pragma unroll
      for (int c = 0; c < CPW; c += 2) *reinterpret_cast<double2*>(o + c) = make_double2(a[c], a[c + 1]);
    } else {
#pragma unroll
      for (int c = 0; c < CPW; ++c)
        if (W * CPW + c < n) o[c] = a[c];
    }
  }
  if (singular || u.nrhs <= 0 || !u.X) return;  // uniform over the workgroup
  // Solve operands in LDS by logical row, shaped so the per-step reads need no
  // lane masks: L strictly lower (zeros elsewhere) and V = U D^-1 strictly
  // upper (column k of U times 1 / u_kk, one division per column). Back
  // substitution then runs on z = D x: z_i -= V[i][k] z_k, and x_i = z_i / u_ii
  // at the end. (linalg.cpp:83-88 divides x_k by u_kk before it is used;
  // scaling U's column instead gives the same solution within rounding.)
  if (lane < N) {
    // the lane holding its row's diagonal entry (at most one column here)
    double dv = 1.0;
    bool has = false;
#pragma unroll
    for (int c = 0; c < CPW; ++c)
      if (W * CPW + c == pos) { dv = a[c]; has = true; }
    if (has) {
      sh.D[pos] = dv;
      sh.Dr[pos] = 1.0 / dv;
    }
    if (W == 0) sh.perm[pos] = lane;
    if (AUG && W == NW - 1) sh.y[pos] = ab;
  }
  __syncthreads();
  if (lane < N) {
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const int j = W * CPW + c;
      if constexpr (!AUG) sh.L[pos][j] = j < pos ? a[c] : 0.0;
      sh.U[pos][j] = j > pos ? a[c] * sh.Dr[j] : 0.0;
    }
  }
  __syncthreads();
  const double* B = u.B + mat * u.strideB;
  double* X = u.X + mat * u.strideB;
  const int nrhs = u.nrhs;
  const int src = lane < n ? sh.perm[lane] : 0;
  const int row = lane < N ? lane : 0;
  const double dl = sh.D[row];
  // waves take chunks of NR right-hand-side columns in turn
  for (int c0 = W * NR; c0 < nrhs; c0 += NW * NR) {
    // n through an opaque per-chunk zero: the 2N step guards below are then
    // not hoisted out of the chunk loop as 2N saved masks
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    const int nn = n + z;
    double y[NR];
    if constexpr (AUG) {
      y[0] = lane < n ? sh.y[lane] : 0.0;  // already L^-1 P b
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q) y[q] = (lane < n && c0 + q < nrhs) ? B[(int64_t)src * nrhs + c0 + q] : 0.0;
      // forward: y_i -= L[i][k] y_k (row order = logical position)
#pragma unroll
      for (int k = 0; k < N - 1; ++k) {
        if (k + 1 >= nn) break;
        const double f = sh.L[row][k];
#pragma unroll
        for (int q = 0; q < NR; ++q) y[q] = fma(-f, readlane_d(y[q], k), y[q]);
      }
    }
    // back: z_i -= V[i][k] z_k for i < k
#pragma unroll
    for (int k = N - 1; k >= 1; --k) {
      if (k >= nn) continue;
      const double f = sh.U[row][k];
#pragma unroll
      for (int q = 0; q < NR; ++q) y[q] = fma(-f, readlane_d(y[q], k), y[q]);
    }
    if (lane < n) {
      double* xr = X + (int64_t)lane * nrhs + c0;
#pragma unroll
      for (int q = 0; q < NR; ++q)
        if (c0 + q < nrhs) xr[q] = y[q] / dl;
    }
  }
}

// Uniform stack: one workgroup of NW waves per matrix.
template <int N, int NR, int WV>
__global__ __launch_bounds__((Small<N, WV>::NT)) void lu_small_kernel(UniformBatch u) {
  __shared__ SmallShared<N> sh;
  const int64_t mat = blockIdx.x;
  // wave index made provably uniform, so the per-wave bodies are scalar
  // branches (no exec-mask state kept live across them)
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (threadIdx.x == 0) sh.sing = -1;
  __syncthreads();
  [&]<int... W>(std::integer_sequence<int, W...>) __attribute__((always_inline)) {
    ((wave == W ? (wave_body<N, WV, W, NR>(u, sh, mat, lane), 0) : 0), ...);
  }(std::make_integer_sequence<int, Small<N, WV>::NW>{});
}

}  // namespace

int small_lu_launch(const UniformBatch& u, hipStream_t st) {
  if (u.count <= 0 || u.n <= 0) return NPPH_OK;
  if (u.count > 0x7fffffff) { set_error("batch too large"); return NPPH_ERR_INVALID; }
  if (u.nrhs > 0 && u.X && u.B != u.X) {
    set_error("small_lu: the solve works in place (X must alias B)");
    return NPPH_ERR_INVALID;
  }
  dim3 grid((unsigned)u.count);
  const double n = u.n, flu = n * (n - 1) / 2 + (n - 1) * n * (2 * n - 1) / 3;
  ProfScope ps(kProfSmallLu, u.count * (flu + (2 * n * n - n) * u.nrhs),
               8.0 * u.count * (n * n + 2 * n * u.nrhs + (u.LU ? n * n : 0)), st);
  // a single right-hand side (or none) solves one register column per pass
  const bool one = u.nrhs <= 1;
  // order 64 with one right-hand side: the columns split over eight waves (the
  // per-step update of each wave's columns, which every step waits for, is
  // half of what four waves do; 256 x 64^2: 51.6 -> 49.3 us per solve with
  // the input copies); NUMPP_HIP_SMALL_WAVES=4 or 16 selects another split
  static const int waves64 = [] {
    const char* e = std::getenv("NUMPP_HIP_SMALL_WAVES");
    const int v = e ? std::atoi(e) : 8;
    return v == 4 || v == 16 ? v : 8;
  }();
  auto launch = [&]<int NV>() {
    constexpr int W0 = Small<NV>::NW;
    if constexpr (NV == 64) {
      if (one && waves64 == 8) {
        hipLaunchKernelGGL((lu_small_kernel<NV, 1, 8>), grid, dim3(Small<NV, 8>::NT), 0, st, u);
        return;
      }
      if (one && waves64 == 16) {
        hipLaunchKernelGGL((lu_small_kernel<NV, 1, 16>), grid, dim3(Small<NV, 16>::NT), 0, st, u);
        return;
      }
    }
    if (one) hipLaunchKernelGGL((lu_small_kernel<NV, 1, W0>), grid, dim3(Small<NV>::NT), 0, st, u);
    else hipLaunchKernelGGL((lu_small_kernel<NV, 8, W0>), grid, dim3(Small<NV>::NT), 0, st, u);
  };
  if (u.n <= 8) launch.template operator()<8>();
  else if (u.n <= 16) launch.template operator()<16>();
  else if (u.n <= 32) launch.template operator()<32>();
  else if (u.n <= 64) launch.template operator()<64>();
  else { set_error("small_lu: n=%d > 64", u.n); return NPPH_ERR_INVALID; }
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace npph
