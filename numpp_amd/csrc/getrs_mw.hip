// Triangular sweeps spread over many workgroups per system, for stacks of few
// systems of large order (the single-workgroup sweeps of getrs.hip would leave
// most CUs idle and stream each factor with one CU).
//
// X = U^-1 L^-1 P B as in lu_solve (linalg.cpp:73-90), in two launches. In
// each, one workgroup owns one 64-row block of one system and one chunk of
// right-hand-side columns: it accumulates the products of its factor strip
// with the blocks solved before it (L_bj Y_j for j < b forward, U_bj X_j for
// j > b backward) on MFMA 16x16x4 as their results appear, solves its 64 x 64
// diagonal block and writes its rows, to X and, tagged, to a staging area that
// the later blocks poll.
//
// Workgroups take their (step, system, chunk) from a ticket counter in step
// order, so a workgroup only ever waits for blocks whose workgroups took
// earlier tickets and are therefore already running: the lowest unfinished
// ticket never waits, and the sweep always drains. Waits are bounded (a
// workgroup that gives up records it in an error word and goes on).
//
// Publication without flags or fences: every solved entry goes to the staging
// area as 64-bit words that carry the value (FP32: one word; FP64: two words,
// one per half) in their low 32 bits and the sweep's tag (1 forward, 2
// backward) in their high 32 bits, written and read with device-scope (agent)
// relaxed atomic stores and loads, which the hardware keeps coherent across
// the XCDs' L2s word by word. A reader polls the words it needs until every
// tag is the sweep's, so the poll is the data read: one trip through the
// memory side per dependency instead of flag-then-data, and the writer never
// waits for its stores to be acknowledged. The staging area is cleared before
// each forward sweep, so no word carries a tag it was not written with. An
// acquire/release pair would instead write back or invalidate the whole L2 of
// the XCD on every poll and publish.
#include <type_traits>

#include "common.hpp"
#include "lu.hpp"

namespace npph {

namespace {

constexpr int MB = 64;   // rows per block
constexpr int MT = 256;  // threads per workgroup (4 waves: 16 block rows each)
constexpr uint32_t kSpinLimit = 1u << 20;  // polls (about 0.5 s) before a wait gives up
// sync words: [0] forward tickets, [1] backward tickets, [2] error, [3] unused,
// then the staging area: per (system, chunk, block), MB x CW entries of
// WPV(T) tagged 64-bit words
constexpr int kSyncHead = 4;
template <class T>
constexpr int WPV = sizeof(T) / 4;  // tagged words per value

struct MwRef {
  const double* R = nullptr;      // residual (forward right-hand side)
  const double* scale = nullptr;  // per (system, column) scale of the residual
  double* X = nullptr;            // FP64 iterate: X (+)= correction / scale
  int first = 0;
  const int32_t* live = nullptr;  // device count of systems still iterating
};

__device__ __forceinline__ v4f mfma_mw(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4d mfma_mw(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// accumulator element v of lane group lk holds output row ... (16x16x4 layouts)
template <class T>
__device__ __forceinline__ int mw_row(int lk, int v) { return sizeof(T) == 4 ? 4 * lk + v : lk + 4 * v; }

// Device-coherent (agent scope) element access through the integer of the same width.
template <class T>
using bits_t = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
template <class T>
__device__ __forceinline__ T coherent_load(const T* p) {
  return __builtin_bit_cast(T, __hip_atomic_load(reinterpret_cast<const bits_t<T>*>(p), __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT));
}

// A value's tagged words: (32 value bits) | tag << 32, one word per 32 bits.
template <class T>
__device__ __forceinline__ void tagged_store(uint64_t* w, T v, uint32_t tag) {
  const uint64_t hi = (uint64_t)tag << 32;
  if constexpr (sizeof(T) == 4) {
    __hip_atomic_store(w, hi | __builtin_bit_cast(uint32_t, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    __hip_atomic_store(w, hi | (b & 0xffffffffu), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 1, hi | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}
// Loads a value's tagged words; false while any of them lacks the tag.
template <class T>
__device__ __forceinline__ bool tagged_load(const uint64_t* w, uint32_t tag, T& v) {
  const uint64_t a = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if constexpr (sizeof(T) == 4) {
    v = __builtin_bit_cast(float, (uint32_t)a);
    return (uint32_t)(a >> 32) == tag;
  } else {
    const uint64_t b = __hip_atomic_load(w + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = __builtin_bit_cast(double, (a & 0xffffffffu) | (b << 32));
    return (uint32_t)(a >> 32) == tag && (uint32_t)(b >> 32) == tag;
  }
}

__global__ void mw_clear_kernel(uint32_t* sync, int64_t words) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) sync[i] = 0u;
}

// CW: right-hand-side columns per workgroup (4, 8 or 16); wave w solves
// columns w, w + 4, ... of the diagonal block (lane = row).
template <class T, bool LOWER, int CW, bool REF>
__global__ __launch_bounds__(MT) void sweep_mw_kernel(int64_t count, int n, int nrhs, int nz,
                                                      const T* __restrict__ LUb, int64_t strideA,
                                                      const int32_t* __restrict__ permb,
                                                      const int32_t* __restrict__ rowmapb, const T* __restrict__ Bb,
                                                      int64_t strideB, T* Xb, int64_t strideX,
                                                      const int32_t* __restrict__ info,
                                                      const int32_t* __restrict__ mask, MwRef ref, uint32_t* sync) {
  static_assert(CW == 4 || CW == 8 || CW == 16, "chunk width");
  constexpr int CPW = CW / 4;
  typedef std::conditional_t<sizeof(T) == 4, v4f, v4d> acc_t;
  __shared__ T dg[MB][MB + 1];
  __shared__ T xs[2][MB][16];
  __shared__ T part[MB][16];
  __shared__ uint32_t tk_s;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // a refinement sweep queued after the last system stopped: every workgroup
  // leaves before taking a ticket (thousands of tickets on one counter)
  if (REF && ref.live && *ref.live == 0) return;
  if (t == 0) tk_s = atomicAdd(&sync[LOWER ? 0 : 1], 1u);
  __syncthreads();
  const int64_t per = count * nz;
  const int nblk = (n + MB - 1) / MB;
  const int64_t tk = tk_s;
  const int st = (int)(tk / per);
  const int64_t rem = tk - (int64_t)st * per;
  const int64_t mat = rem / nz;
  const int z = (int)(rem - mat * nz);
  if (st >= nblk) return;
  // a refinement correction touches only systems still iterating (mask 1)
  if ((info && info[mat] != 0) || (mask && (REF ? mask[mat] != 1 : mask[mat] == 0))) return;
  const int b = LOWER ? st : nblk - 1 - st;
  const int r0 = b * MB, bs = min(MB, n - r0);
  const int c0 = z * CW, nc = min(CW, nrhs - c0);
  const T* __restrict__ LU = LUb + mat * strideA;
  T* X = Xb + mat * strideX;
  // factor row of position p (factors whose rows never moved carry a row map)
  const int32_t* rowmap = rowmapb ? rowmapb + mat * n : nullptr;
  auto rowp = [&](int p) -> int64_t { return rowmap ? rowmap[p] : p; };
  // this (system, chunk)'s staging blocks
  uint64_t* stg = reinterpret_cast<uint64_t*>(sync + kSyncHead) + (mat * nz + z) * nblk * (MB * CW * WPV<T>);
  const uint32_t tag = LOWER ? 1u : 2u;

  // ---- diagonal block into LDS (clamped, coalesced; entries past bs zeroed)
  {
    constexpr int DPT = MB * MB / MT;
    T d[DPT];
#pragma unroll
    for (int u = 0; u < DPT; ++u) {
      const int e = t + u * MT, rr = e / MB, cc = e - rr * MB;
      d[u] = LU[rowp(r0 + min(rr, bs - 1)) * n + r0 + min(cc, bs - 1)];
    }
#pragma unroll
    for (int u = 0; u < DPT; ++u) {
      const int e = t + u * MT, rr = e / MB, cc = e - rr * MB;
      dg[rr][cc] = (rr < bs && cc < bs) ? d[u] : T(0);
    }
  }
  // ---- this block's right-hand side in the solve layout (lane = row)
  T x[CPW];
  {
    const int i = r0 + min(lane, bs - 1);
    if constexpr (LOWER) {
      const int src = permb[mat * n + i];
#pragma unroll
      for (int j = 0; j < CPW; ++j) {
        const int c = wave + 4 * j;
        if constexpr (REF) {
          x[j] = (lane < bs && c < nc)
                     ? (T)(ref.R[mat * strideB + (int64_t)src * nrhs + c0 + c] * ref.scale[mat * nrhs + c0 + c])
                     : T(0);
        } else {
          x[j] = (lane < bs && c < nc) ? Bb[mat * strideB + (int64_t)src * nrhs + c0 + c] : T(0);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < CPW; ++j) {
        const int c = wave + 4 * j;
        x[j] = (lane < bs && c < nc) ? X[(int64_t)i * nrhs + c0 + c] : T(0);
      }
    }
  }

  // ---- products with the blocks solved before this one
  const int lr = lane & 15, lk = lane >> 4, rw = 16 * wave;
  const T* __restrict__ Lrow = LU + rowp(r0 + min(rw + lr, bs - 1)) * n;
  const bool vec = (n % 4 == 0) && (reinterpret_cast<uintptr_t>(LUb) % 16 == 0) && (strideA % 4 == 0);
  // a[u][q] = factor entry (row rw + lr, column q0 + 16 u + 4 lk + q) of
  // dependency block j; columns past the block are zeroed
  auto load_tile = [&](int j, T (&a)[4][4]) {
    const int q0 = j * MB, qs = min(MB, n - q0);
    if (vec && qs == MB) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const T* p = Lrow + q0 + 16 * u + 4 * lk;
        if constexpr (sizeof(T) == 4) {
          const float4 v = *reinterpret_cast<const float4*>(p);
          a[u][0] = v.x; a[u][1] = v.y; a[u][2] = v.z; a[u][3] = v.w;
        } else {
          const double2 v0 = *reinterpret_cast<const double2*>(p);
          const double2 v1 = *reinterpret_cast<const double2*>(p + 2);
          a[u][0] = v0.x; a[u][1] = v0.y; a[u][2] = v1.x; a[u][3] = v1.y;
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int k = 16 * u + 4 * lk + q;
          a[u][q] = Lrow[q0 + min(k, qs - 1)];
          a[u][q] = k < qs ? a[u][q] : T(0);
        }
    }
  };
  acc_t acc = {};
  // block j's solved entries from the staging area: every thread polls its
  // entries until the whole workgroup holds tagged ones (bounded)
  constexpr int EPT = MB * CW / MT;  // entries per thread
  auto consume = [&](const T (&a)[4][4], int j, int buf) {
    const uint64_t* sj = stg + (int64_t)j * (MB * CW * WPV<T>);
    T val[EPT];
    uint32_t spins = 0;
    for (;;) {
      bool ok = true;
#pragma unroll
      for (int u = 0; u < EPT; ++u) ok = tagged_load<T>(sj + (t + MT * u) * WPV<T>, tag, val[u]) && ok;
      if (__syncthreads_and(ok)) break;
      if (++spins > kSpinLimit) {
        if (t == 0) __hip_atomic_store(&sync[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      // about 1000 cycles between polls: a thousand waiting workgroups
      // polling back to back slow the memory side that carries the
      // publications (64 x 2048 FP32 sweeps: 1.54 -> 1.50 ms per four calls)
      __builtin_amdgcn_s_sleep(16);
    }
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = t + MT * u, rr = e / CW, cc = e % CW;
      xs[buf][rr][cc] = val[u];
    }
    if constexpr (CW < 16) {
      // columns CW..15 of the MFMA operand stay zero
      for (int e = t; e < MB * (16 - CW); e += MT) xs[buf][e / (16 - CW)][CW + e % (16 - CW)] = T(0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc = mfma_mw(a[u][q], xs[buf][16 * u + 4 * lk + q][lr], acc);
  };
  const int ndep = LOWER ? b : nblk - 1 - b;
  auto dep = [&](int d) { return LOWER ? d : nblk - 1 - d; };
  if (ndep > 0) {
    T a0[4][4], a1[4][4];
    load_tile(dep(0), a0);
    for (int d = 0; d < ndep; d += 2) {
      if (d + 1 < ndep) load_tile(dep(d + 1), a1);
      consume(a0, dep(d), 0);
      if (d + 1 < ndep) {
        if (d + 2 < ndep) load_tile(dep(d + 2), a0);
        consume(a1, dep(d + 1), 1);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) part[rw + mw_row<T>(lk, v)][lr] = acc[v];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < CPW; ++j) x[j] = x[j] - part[lane][wave + 4 * j];

  // ---- diagonal block (lane = row); factor entries read from LDS DCH at a time
  constexpr int DCH = sizeof(T) == 4 ? 8 : 4;
  if constexpr (LOWER) {
#pragma unroll
    for (int q0 = 0; q0 < MB; q0 += DCH) {
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
    // one division per lane; the row's scaling at its step is a multiply
    const T dinv = T(1) / (lane < bs ? dg[lane][lane] : T(1));
#pragma unroll
    for (int q0 = MB - DCH; q0 >= 0; q0 -= DCH) {
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

  // ---- rows out: published (tagged) first, for the blocks that wait on them
  if (st + 1 < nblk) {
    uint64_t* sb = stg + (int64_t)b * (MB * CW * WPV<T>);
#pragma unroll
    for (int j = 0; j < CPW; ++j) {
      const int c = wave + 4 * j;
      // rows past bs and columns past nc publish zeros
      tagged_store<T>(sb + (lane * CW + c) * WPV<T>, (lane < bs && c < nc) ? x[j] : T(0), tag);
    }
  }
  if (lane < bs) {
    const int i = r0 + lane;
#pragma unroll
    for (int j = 0; j < CPW; ++j) {
      const int c = wave + 4 * j;
      if (c < nc) {
        X[(int64_t)i * nrhs + c0 + c] = x[j];
        if constexpr (REF && !LOWER) {
          const double dd = (double)x[j] / ref.scale[mat * nrhs + c0 + c];
          double* xr = ref.X + mat * strideX + (int64_t)i * nrhs + c0 + c;
          *xr = ref.first ? dd : *xr + dd;
        }
      }
    }
  }
}

template <class T, int CW, bool REF>
void launch_mw(int64_t count, int n, int nrhs, int nz, const T* LU, int64_t strideA, const int32_t* perm,
               const int32_t* rowmap, const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info,
               const int32_t* mask, const MwRef& ref, uint32_t* sync, hipStream_t s) {
  const int64_t wgs = count * nz * ceil_div(n, MB);
  hipLaunchKernelGGL((sweep_mw_kernel<T, true, CW, REF>), dim3((unsigned)wgs), dim3(MT), 0, s, count, n, nrhs, nz,
                     LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask, ref, sync);
  hipLaunchKernelGGL((sweep_mw_kernel<T, false, CW, REF>), dim3((unsigned)wgs), dim3(MT), 0, s, count, n, nrhs, nz,
                     LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask, ref, sync);
}

}  // namespace

bool getrs_mw_wanted(int64_t count, int n, int nrhs) {
  if (n < 512 || nrhs <= 0) return false;
  const int64_t cus = cu_count() > 0 ? cu_count() : 256;
  // one single-workgroup sweep per (system, 16 columns) fills the device only
  // with several workgroups per CU
  return count * ceil_div(nrhs, 16) < 4 * cus && count * ceil_div(nrhs, 16) * ceil_div(n, MB) < (1LL << 31);
}

template <class T, bool REF>
int getrs_mw(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm, const int32_t* rowmap,
             const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info, const int32_t* mask,
             const double* R, const double* scale, double* X64, int first, const int32_t* live, hipStream_t s) {
  const int cw = nrhs <= 4 ? 4 : (nrhs <= 8 ? 8 : 16);
  const int nz = (int)ceil_div(nrhs, cw);
  const int64_t words = kSyncHead + 2 * count * nz * ceil_div(n, MB) * (MB * cw * (int64_t)(sizeof(T) / 4));
  uint32_t* sync = nullptr;
  NPPH_TRY(workspace(kWsSync, sizeof(uint32_t) * (size_t)words, (void**)&sync));
  hipLaunchKernelGGL(mw_clear_kernel, dim3((unsigned)ceil_div(words, 256)), dim3(256), 0, s, sync, words);
  MwRef ref{R, scale, X64, first, live};
  if (cw == 4)
    launch_mw<T, 4, REF>(count, n, nrhs, nz, LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask, ref, sync, s);
  else if (cw == 8)
    launch_mw<T, 8, REF>(count, n, nrhs, nz, LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask, ref, sync, s);
  else
    launch_mw<T, 16, REF>(count, n, nrhs, nz, LU, strideA, perm, rowmap, B, strideB, X, strideX, info, mask, ref, sync,
                          s);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

template int getrs_mw<double, false>(int64_t, int, int, const double*, int64_t, const int32_t*, const int32_t*,
                                     const double*, int64_t, double*, int64_t, const int32_t*, const int32_t*,
                                     const double*, const double*, double*, int, const int32_t*, hipStream_t);
template int getrs_mw<float, false>(int64_t, int, int, const float*, int64_t, const int32_t*, const int32_t*,
                                    const float*, int64_t, float*, int64_t, const int32_t*, const int32_t*,
                                    const double*, const double*, double*, int, const int32_t*, hipStream_t);
template int getrs_mw<float, true>(int64_t, int, int, const float*, int64_t, const int32_t*, const int32_t*,
                                   const float*, int64_t, float*, int64_t, const int32_t*, const int32_t*,
                                   const double*, const double*, double*, int, const int32_t*, hipStream_t);

}  // namespace npph
