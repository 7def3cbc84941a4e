// Triangular solves with packed LU factors (getrs) and pivot conversion.
//
// X = U^-1 L^-1 P B, following lu_solve (linalg.cpp:73-90): gather B by the
// row permutation, forward substitution with the unit-lower L, back
// substitution with U dividing by its diagonal.
//
// Each sweep is one launch with one workgroup per (matrix, chunk of CW
// right-hand-side columns). The workgroup keeps its n x CW slice of the
// right-hand side in LDS and walks the 64-row blocks in order: it first
// subtracts the contributions of the blocks already solved (L_ij X_j for
// j < i forward, U_ij X_j for j > i backward), streaming its 64-row strip of
// the factor with S threads per row, then solves the 64 x 64 diagonal block
// with one wave per column (lane = row, v_readlane broadcasts; U's diagonal is
// applied as a multiply by its reciprocal). Nothing is shared between
// workgroups, so the sweep needs no inter-workgroup flags (which on this
// device would need L2 write-back and invalidation at device scope).
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "common.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int GB = 64;  // rows per block

// FP64 side of a refinement correction (unused by plain solves)
struct RefIO {
  const double* R = nullptr;      // residual (the forward sweep's right-hand side)
  const double* scale = nullptr;  // per (system, column) power-of-two scale
  double* X = nullptr;            // iterate: X (+)= correction / scale
  int first = 0;                  // 1: X = correction / scale
  const int32_t* live = nullptr;  // device count of systems still iterating
};

__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// output row of accumulator element v for lane group lk (16x16x4 layouts)
template <class T>
__device__ __forceinline__ int mfma_row(int lk, int v) { return sizeof(T) == 4 ? 4 * lk + v : lk + 4 * v; }

// REF: one refinement correction (FP32 factors): the forward sweep reads the
// FP64 residual, scaled per column and rounded (in place of B), and the
// backward sweep adds x / scale to the FP64 iterate (in place of storing X).
template <class T, bool LOWER, int NT, int CW, bool REF>
__global__ __launch_bounds__(NT) void sweep_kernel(int n, int nrhs, const T* LUb, int64_t strideA,
                                                   const int32_t* permb, const int32_t* rowmapb, const T* Bb,
                                                   int64_t strideB, T* Xb, int64_t strideX, const int32_t* info,
                                                   const int32_t* mask, RefIO ref) {
  constexpr int S = NT / GB;       // threads per row in the off-diagonal pass
  constexpr int NWV = NT / 64;     // waves
  constexpr int CPW = (CW + NWV - 1) / NWV;  // columns per wave in the diagonal solve
  // factor entries per thread and step: deep for long rows (n > 256), short
  // for small orders, where more resident workgroups hide the latency better
  // than deeper batches (measured)
  constexpr int M = NT > 256 ? (sizeof(T) == 8 ? 16 : 32) : 8;
  const int64_t mat = blockIdx.x;
  // a refinement correction touches only systems still iterating (state 1);
  // converged (0) and stopped (2) systems keep their iterate
  if ((info && info[mat] != 0) || (mask && (REF ? mask[mat] != 1 : mask[mat] == 0))) return;
  const int c0 = (int)blockIdx.y * CW, nc = min(CW, nrhs - c0);
  const T* __restrict__ LU = LUb + mat * strideA;
  T* __restrict__ X = Xb + mat * strideX;
  // factor row of position p (factors whose rows never moved carry a row map;
  // orders up to 256 keep it in LDS, the sweeps' chains read it at every block)
  const int32_t* rowmap = rowmapb ? rowmapb + mat * n : nullptr;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  __shared__ int rmap_s[NT == 256 ? 256 : 1];
  if constexpr (NT == 256) {
    if (rowmap)
      for (int i = t; i < n; i += NT) rmap_s[i] = rowmap[i];
  }
  auto rowp = [&](int p) -> int64_t {
    if constexpr (NT == 256) return rowmap ? rmap_s[p] : p;
    return rowmap ? rowmap[p] : p;
  };
  extern __shared__ unsigned char smem_raw[];
  T* xs = reinterpret_cast<T*>(smem_raw);  // n x CW
  __shared__ T part[GB][CW];
  __shared__ T dg[GB][GB + 1];  // the current diagonal block (padded rows)
  constexpr int DPT = GB * GB / NT;  // diagonal-block entries per thread

  // per-column scales of a refinement correction
  double sc[CW];
#pragma unroll
  for (int c = 0; c < CW; ++c) sc[c] = 1.0;
  if constexpr (REF) {
#pragma unroll
    for (int c = 0; c < CW; ++c) sc[c] = c < nc ? ref.scale[mat * nrhs + c0 + c] : 1.0;
  }
  // right-hand side slice: B gathered by perm (forward) or X (backward); a
  // thread takes whole rows, RP rows per pass with their loads in flight
  {
    constexpr int RP = CW >= 16 ? 2 : 4;
    const T* Bm = REF ? nullptr : Bb + mat * strideB;
    const int32_t* perm = permb + mat * n;
    for (int i0 = t; i0 < n; i0 += RP * NT) {
      int src[RP];
#pragma unroll
      for (int u = 0; u < RP; ++u) {
        const int i = min(i0 + u * NT, n - 1);
        src[u] = LOWER ? perm[i] : i;
      }
      T v[RP][CW];
#pragma unroll
      for (int u = 0; u < RP; ++u) {
        if constexpr (REF && LOWER) {
          const double* rr = ref.R + mat * strideB + (int64_t)src[u] * nrhs + c0;
#pragma unroll
          for (int c = 0; c < CW; ++c) v[u][c] = c < nc ? (T)(rr[c] * sc[c]) : T(0);
        } else {
          const T* br = (LOWER ? Bm : X) + (int64_t)src[u] * nrhs + c0;
#pragma unroll
          for (int c = 0; c < CW; ++c) v[u][c] = c < nc ? br[c] : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < RP; ++u) {
        const int i = i0 + u * NT;
        if (i < n) {
#pragma unroll
          for (int c = 0; c < CW; ++c) xs[(int64_t)i * CW + c] = v[u][c];
        }
      }
    }
  }
  __syncthreads();

  const int nblk = (n + GB - 1) / GB;
  const int r = t / S, sg = t - r * S;
  // A step's diagonal block, loaded by all threads (coalesced, clamped
  // indices): the first before the loop, each next one as soon as the current
  // one is in LDS, so its latency hides behind the diagonal solve and the next
  // off-diagonal pass.
  T dpre[DPT];
  auto load_diag = [&](int st) {
    const int b = LOWER ? st : nblk - 1 - st;
    const int q0 = b * GB, qs = min(GB, n - q0);
#pragma unroll
    for (int u = 0; u < DPT; ++u) {
      const int e = t + u * NT, rr = e / GB, cc = e - rr * GB;
      dpre[u] = LU[rowp(q0 + min(rr, qs - 1)) * n + q0 + min(cc, qs - 1)];
    }
  };
  load_diag(0);
  for (int step = 0; step < nblk; ++step) {
    const int bi = LOWER ? step : nblk - 1 - step;
    const int r0 = bi * GB, bs = min(GB, n - r0);
    const int klo = LOWER ? 0 : r0 + bs, khi = LOWER ? r0 : n;
    if constexpr (NT == 256) {
      // ---- off-diagonal contributions as a (64 x K) x (K x 16) product on
      // MFMA 16x16x4: wave w takes block rows 16 w .. 16 w + 15; a lane loads
      // four consecutive factor entries of its row per 16-column chunk (MFMA q
      // of the chunk contracts over k = 4 lk + q in both operands), four
      // chunks in flight; the X operand comes from the LDS slice (columns
      // past CW are zero)
      typedef std::conditional_t<sizeof(T) == 4, v4f, v4d> acc_t;
      constexpr int KU = 4;
      const int lr = lane & 15, lk = lane >> 4, rw = 16 * wave;
      acc_t acc = {};
      if (rw < bs && klo < khi) {
        const T* __restrict__ Lr = LU + rowp(r0 + min(rw + lr, bs - 1)) * n;
        // aligned vector loads when every row starts on 16 bytes
        const bool vec = n % (16 / (int)sizeof(T)) == 0 && n >= 4 && reinterpret_cast<uintptr_t>(LUb) % 16 == 0 &&
                         strideA % (16 / (int)sizeof(T)) == 0;
        for (int kb = klo; kb < khi; kb += 16 * KU) {
          T a[KU][4];
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            const int k = kb + 16 * u + 4 * lk;
            if (vec) {
              // k and n are multiples of 4 (FP32) or 2 (FP64): the clamped
              // vectors lie inside the row
              if constexpr (sizeof(T) == 4) {
                const float4 p = *reinterpret_cast<const float4*>(Lr + min(k, n - 4));
                a[u][0] = p.x; a[u][1] = p.y; a[u][2] = p.z; a[u][3] = p.w;
              } else {
                const double2 p0 = *reinterpret_cast<const double2*>(Lr + min(k, n - 2));
                const double2 p1 = *reinterpret_cast<const double2*>(Lr + min(k + 2, n - 2));
                a[u][0] = p0.x; a[u][1] = p0.y; a[u][2] = p1.x; a[u][3] = p1.y;
              }
            } else {
#pragma unroll
              for (int q = 0; q < 4; ++q) a[u][q] = Lr[min(k + q, khi - 1)];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) a[u][q] = k + q < khi ? a[u][q] : T(0);
          }
#pragma unroll
          for (int u = 0; u < KU; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int kx = min(kb + 16 * u + 4 * lk + q, khi - 1);
              const T b = lr < CW ? xs[(int64_t)kx * CW + (lr < CW ? lr : 0)] : T(0);
              acc = mfma16(a[u][q], b, acc);
            }
        }
      }
      if (lr < CW) {
#pragma unroll
        for (int v = 0; v < 4; ++v) part[rw + mfma_row<T>(lk, v)][lr] = acc[v];
      }
    } else {
    // ---- off-diagonal contributions: row r of the block; its S threads take
    // interleaved columns (klo + sg + S m), so a row's threads read
    // consecutive factor entries and consecutive rows of the LDS slice
    T acc[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) acc[c] = T(0);
    if (r < bs && klo < khi) {
      const T* __restrict__ Lr = LU + rowp(r0 + r) * n;
      // M entries per thread in flight. Indices past the range are clamped
      // and their factor entry zeroed, so neither the loads nor the FMAs need
      // a branch (a clamped slice row is a solved, finite value).
      for (int kb = klo + sg; kb < khi; kb += S * M) {
        T l[M];
#pragma unroll
        for (int m = 0; m < M; ++m) l[m] = Lr[min(kb + S * m, khi - 1)];
#pragma unroll
        for (int m = 0; m < M; ++m) l[m] = kb + S * m < khi ? l[m] : T(0);
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const T* xr = xs + (int64_t)min(kb + S * m, khi - 1) * CW;
#pragma unroll
          for (int c = 0; c < CW; ++c) acc[c] = acc[c] + l[m] * xr[c];
        }
      }
    }
    // the S threads of a row are consecutive lanes of one wave
#pragma unroll
    for (int m = 1; m < S; m <<= 1)
#pragma unroll
      for (int c = 0; c < CW; ++c) acc[c] += __shfl_xor(acc[c], m, 64);
    if (sg == 0) {
#pragma unroll
      for (int c = 0; c < CW; ++c) part[r][c] = acc[c];
    }
    }
#pragma unroll
    for (int u = 0; u < DPT; ++u) {
      const int e = t + u * NT, rr = e / GB, cc = e - rr * GB;
      dg[rr][cc] = (rr < bs && cc < bs) ? dpre[u] : T(0);
    }
    if (step + 1 < nblk) load_diag(step + 1);
    __syncthreads();
    // ---- diagonal block: wave w solves columns w, w + NWV, ... (lane = row)
    if (wave < CW) {
      const int i = r0 + lane;
      T x[CPW];
#pragma unroll
      for (int j = 0; j < CPW; ++j) {
        const int c = wave + j * NWV;
        x[j] = (lane < bs && c < CW) ? xs[(int64_t)i * CW + c] - part[lane][c] : T(0);
      }
      // The lane's factor entries come from LDS DCH steps at a time, all
      // loads of a group issued together (one LDS round trip per group, not
      // per step); groups past the block's bs rows are skipped (bs is
      // uniform) and the steps of a group past them are masked to no-ops.
      constexpr int DCH = sizeof(T) == 4 ? 8 : 4;
      if (LOWER) {
#pragma unroll
        for (int q0 = 0; q0 < GB; q0 += DCH) {
          if (q0 + 1 >= bs) break;
          T fv[DCH];
#pragma unroll
          for (int u = 0; u < DCH; ++u) fv[u] = dg[lane][q0 + u];
#pragma unroll
          for (int u = 0; u < DCH; ++u) {
            const int q = q0 + u;
            const T f = (lane > q && q + 1 < bs) ? fv[u] : T(0);
#pragma unroll
            for (int j = 0; j < CPW; ++j) x[j] = x[j] - f * readlane<T>(x[j], q);
          }
        }
      } else {
        const T d = lane < bs ? dg[lane][lane] : T(1);
        // one division per lane, in parallel; the row's scaling at its step is
        // then a multiply (within one rounding of lu_solve's division)
        const T dinv = T(1) / d;
#pragma unroll
        for (int q0 = GB - DCH; q0 >= 0; q0 -= DCH) {
          if (q0 >= bs) continue;
          T fv[DCH];
#pragma unroll
          for (int u = 0; u < DCH; ++u) fv[u] = dg[lane][q0 + u];
#pragma unroll
          for (int u = DCH - 1; u >= 0; --u) {
            const int q = q0 + u;
            const bool live = q < bs;
            if (lane == q && live) {
#pragma unroll
              for (int j = 0; j < CPW; ++j) x[j] = x[j] * dinv;
            }
            const T f = (lane < q && live) ? fv[u] : T(0);
#pragma unroll
            for (int j = 0; j < CPW; ++j) x[j] = x[j] - f * readlane<T>(x[j], q);
          }
        }
      }
      if (lane < bs) {
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
          const int c = wave + j * NWV;
          if (c < CW) xs[(int64_t)i * CW + c] = x[j];
        }
      }
    }
    __syncthreads();
  }
  if constexpr (REF && !LOWER) {
    for (int i = t; i < n; i += NT) {
      double* xr = ref.X + mat * strideX + (int64_t)i * nrhs + c0;
#pragma unroll
      for (int c = 0; c < CW; ++c)
        if (c < nc) {
          const double d = (double)xs[(int64_t)i * CW + c] / sc[c];
          xr[c] = ref.first ? d : xr[c] + d;
        }
    }
  } else {
    for (int i = t; i < n; i += NT) {
      T* xr = X + (int64_t)i * nrhs + c0;
#pragma unroll
      for (int c = 0; c < CW; ++c)
        if (c < nc) xr[c] = xs[(int64_t)i * CW + c];
    }
  }
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

// Dynamic LDS a sweep may use: the device's per-workgroup maximum less 48 KiB
// for the static arrays (the diagonal block, up to 64 x 65 doubles, and the
// partials), at most 128 KiB.
size_t lds_budget() {
  static const size_t b = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      v = 64 * 1024;
    const size_t avail = v > 64 * 1024 ? (size_t)v - 48 * 1024 : (size_t)16 * 1024;
    return std::min<size_t>(avail, 128 * 1024);
  }();
  return b;
}

// Right-hand-side columns per workgroup: the widest of 16, 8, 4, 2, 1 whose
// n x CW slice fits in LDS, that still gives at least one workgroup per CU
// (each workgroup streams its matrix's whole triangle, so few wide workgroups
// leave most CUs idle) and that is less than twice nrhs.
// NUMPP_HIP_GETRS_CW = 1, 2, 4 or 8 caps it.
template <class T>
int chunk_width(int64_t count, int n, int nrhs) {
  static const int cap_env = [] {
    const char* e = std::getenv("NUMPP_HIP_GETRS_CW");
    const int v = e ? std::atoi(e) : 16;
    return (v == 1 || v == 2 || v == 4 || v == 8) ? v : 16;
  }();
  const size_t cap = lds_budget();
  const int64_t want = cu_count() > 0 ? cu_count() : 256;
  // FP64 slices of 16 columns hold too much LDS for enough workgroups per CU
  int cw = sizeof(T) == 8 ? std::min(cap_env, 8) : cap_env;
  while (cw > 1 && ((size_t)n * cw * sizeof(T) > cap || count * ceil_div(nrhs, cw) < want || cw >= 2 * nrhs))
    cw /= 2;
  return cw;
}

// Launches one sweep; dynamic LDS above the default 64 KiB is enabled once per
// instantiation.
template <class T, bool LOWER, int NT, int CW, bool REF>
int launch_sweep_cw(dim3 grid, size_t shm, hipStream_t s, int n, int nrhs, const T* LU, int64_t strideA,
                    const int32_t* perm, const int32_t* rowmap, const T* B, int64_t strideB, T* X, int64_t strideX,
                    const int32_t* info, const int32_t* mask, const RefIO& ref) {
  auto* k = sweep_kernel<T, LOWER, NT, CW, REF>;
  // static LDS (diagonal block, partials) plus the slice can pass 64 KiB:
  // opt in once per instantiation to the full dynamic budget
  static const hipError_t once = hipFuncSetAttribute(
      reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_budget());
  NPPH_HIP(once);
  (void)shm;
  hipLaunchKernelGGL(k, grid, dim3(NT), shm, s, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask,
                     ref);
  return NPPH_OK;
}

template <class T, bool LOWER, int NT, bool REF>
int launch_sweep(int cw, dim3 grid, size_t shm, hipStream_t s, int n, int nrhs, const T* LU, int64_t strideA,
                 const int32_t* perm, const int32_t* rowmap, const T* B, int64_t strideB, T* X, int64_t strideX,
                 const int32_t* info, const int32_t* mask, const RefIO& ref) {
#define NPPH_SWEEP(CWV)                                                                                        \
  return launch_sweep_cw<T, LOWER, NT, CWV, REF>(grid, shm, s, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X, \
                                                 strideX, info, mask, ref)
  if constexpr (sizeof(T) == 4) {
    if (cw == 16) NPPH_SWEEP(16);
  }
  if (cw == 8) NPPH_SWEEP(8);
  if (cw == 4) NPPH_SWEEP(4);
  if (cw == 2) NPPH_SWEEP(2);
  NPPH_SWEEP(1);
#undef NPPH_SWEEP
}

// NUMPP_HIP_GETRS_MW=0 keeps the single-workgroup sweeps for every shape.
bool getrs_mw_off() {
  static const bool off = [] {
    const char* e = std::getenv("NUMPP_HIP_GETRS_MW");
    return e && e[0] == '0';
  }();
  return off;
}

template <class T, bool REF>
int getrs_impl(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm,
               const int32_t* rowmap, const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info,
               const int32_t* mask, const RefIO& ref, hipStream_t s) {
  if (count <= 0 || n <= 0 || nrhs <= 0) return NPPH_OK;
  if (!REF && B == X) { set_error("getrs: B and X must not alias"); return NPPH_ERR_INVALID; }
  if (count > 2147483647LL) { set_error("getrs: count too large"); return NPPH_ERR_INVALID; }
  if (getrs_mw_wanted(count, n, nrhs) && !getrs_mw_off()) {
    ProfScope ps(kProfGetrs, 2.0 * count * n * n * nrhs,
                 (double)count * (sizeof(T) * (double)n * n + 3.0 * sizeof(T) * n * nrhs), s);
    return getrs_mw<T, REF>(count, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask, ref.R,
                            ref.scale, ref.X, ref.first, ref.live, s);
  }
  const int cw = chunk_width<T>(count, n, nrhs);
  const int nz = (int)ceil_div(nrhs, cw);
  if (nz > 65535) { set_error("getrs: nrhs too large"); return NPPH_ERR_INVALID; }
  const size_t shm = (size_t)n * cw * sizeof(T);
  dim3 grid((unsigned)count, (unsigned)nz);
  ProfScope ps(kProfGetrs, 2.0 * count * n * n * nrhs,
               (double)count * (sizeof(T) * (double)n * n + 3.0 * sizeof(T) * n * nrhs), s);
  if (n <= 256) {
    NPPH_TRY((launch_sweep<T, true, 256, REF>(cw, grid, shm, s, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X,
                                              strideX, info, mask, ref)));
    NPPH_TRY((launch_sweep<T, false, 256, REF>(cw, grid, shm, s, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X,
                                               strideX, info, mask, ref)));
  } else {
    NPPH_TRY((launch_sweep<T, true, 512, REF>(cw, grid, shm, s, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X,
                                              strideX, info, mask, ref)));
    NPPH_TRY((launch_sweep<T, false, 512, REF>(cw, grid, shm, s, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X,
                                               strideX, info, mask, ref)));
  }
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

template <class T>
int blocked_getrs_perm(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm,
                       const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info, const int32_t* mask,
                       hipStream_t s, const int32_t* rowmap) {
  return getrs_impl<T, false>(count, n, nrhs, LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask, RefIO{},
                              s);
}

int getrs_refine(int64_t count, int n, int nrhs, const float* LU, int64_t strideA, const int32_t* perm,
                 const double* R, const double* scale, float* Y, double* X, int first, const int32_t* mask,
                 hipStream_t s, const int32_t* rowmap, const int32_t* live) {
  const int64_t per = (int64_t)n * nrhs;
  return getrs_impl<float, true>(count, n, nrhs, LU, strideA, perm, rowmap, nullptr, per, Y, per, nullptr, mask,
                                 RefIO{R, scale, X, first, live}, s);
}

template int blocked_getrs_perm<double>(int64_t, int, int, const double*, int64_t, const int32_t*, const double*,
                                        int64_t, double*, int64_t, const int32_t*, const int32_t*, hipStream_t,
                                        const int32_t*);
template int blocked_getrs_perm<float>(int64_t, int, int, const float*, int64_t, const int32_t*, const float*,
                                       int64_t, float*, int64_t, const int32_t*, const int32_t*, hipStream_t,
                                       const int32_t*);

int ipiv_to_perm(int64_t count, int n, const int32_t* ipiv, int32_t* perm, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  hipLaunchKernelGGL(ipiv_to_perm_kernel, dim3((unsigned)count), dim3(64), (size_t)n * sizeof(int), s, n, ipiv,
                     perm);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace npph
