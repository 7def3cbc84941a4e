// Mixed-precision stacked solve with FP64 iterative refinement.
//
// The LU factorisation runs in FP32 (FP32 MFMA trailing updates) or with
// split-BF16 MFMA trailing updates; FP64 is used on the device only for the
// residual r = b - A x, the solution update x += d and the backward-error test
//   eta_j = ||r_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
// per right-hand-side column. Each sweep: residual (reads A once in FP64, the
// HBM-bound kernel), eta + convergence per system, power-of-two scaling of r
// into FP32, getrs with the low-precision factors, x += d. A system leaves the
// loop when every column meets tol; one that stagnates (eta not halving) or
// reaches max_iter is re-solved with the plain FP64 LU when fallback is on.
// Per-system arithmetic never depends on the other systems in the stack.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

int gesv_dev_impl(int64_t count, int n, int nrhs, double* A, int64_t strideA, double* B, int64_t strideB,
                  int32_t* ipiv, int32_t* info, hipStream_t s);

namespace {

__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// A (FP64) -> Alo (FP32, or an FP64 copy) and ||A||_inf per matrix. Each wave
// converts kDemoteRows consecutive rows, eight 16-byte loads per lane in
// flight, and folds their absolute row sums into one atomic maximum per
// matrix it touches (the atomics are few: on one address they serialise).
constexpr int kDemoteRows = 4;
template <class To>
__global__ __launch_bounds__(256) void demote_kernel(int64_t count, int n, const double* A, To* Alo,
                                                     double* anorm) {
  typedef To pair_t __attribute__((ext_vector_type(2)));
  const int64_t row0 = (((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6) * kDemoteRows;
  const int lane = threadIdx.x & 63;
  const int64_t rows = count * n;
  int64_t cur = -1;  // matrix of the running maximum
  double best = 0.0;
  for (int r = 0; r < kDemoteRows && row0 + r < rows; ++r) {
    const int64_t row = row0 + r;
    const int64_t mat = row / n;
    const double* a = A + row * n;
    To* o = Alo + row * n;
    double s = 0.0;
    if ((n & 1) == 0) {
      int j = 2 * lane;
      for (; j + 896 < n; j += 1024) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const double2*>(a + j + 128 * u);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          *reinterpret_cast<pair_t*>(o + j + 128 * u) = pair_t{(To)v[u].x, (To)v[u].y};
          s += fabs(v[u].x) + fabs(v[u].y);
        }
      }
      for (; j < n; j += 128) {
        const double2 v = *reinterpret_cast<const double2*>(a + j);
        *reinterpret_cast<pair_t*>(o + j) = pair_t{(To)v.x, (To)v.y};
        s += fabs(v.x) + fabs(v.y);
      }
    } else {
      for (int j = lane; j < n; j += 64) {
        const double v = a[j];
        o[j] = (To)v;
        s += fabs(v);
      }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (mat != cur) {
      if (cur >= 0 && lane == 0) atomic_max_nonneg(anorm + cur, best);
      cur = mat;
      best = 0.0;
    }
    best = fmax(best, s);
  }
  if (cur >= 0 && lane == 0) atomic_max_nonneg(anorm + cur, best);
}

// ||b_j||_inf per (matrix, column): one workgroup per matrix, CPW columns
// (nrhs rounded up to a power of two, at most 64) by 256 / CPW rows per pass,
// reduced across a wave's row groups with xor shuffles and across the four
// waves in LDS; every maximum is written once (no atomics).
// (ctr, when given: the run's two counters, zeroed by workgroup 0.)
__global__ __launch_bounds__(256) void bnorm_kernel(int n, int nrhs, const double* Bb, double* bmax, int32_t* ctr) {
  const int64_t mat = blockIdx.x;
  if (ctr && mat == 0 && threadIdx.x < 2) ctr[threadIdx.x] = 0;
  const double* __restrict__ B = Bb + mat * (int64_t)n * nrhs;
  __shared__ double red[4][64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int cpw = 1;
  while (cpw < nrhs && cpw < 64) cpw <<= 1;
  const int rp = 256 / cpw, r = t / cpw, cl = t % cpw;
  for (int c0 = 0; c0 < nrhs; c0 += cpw) {
    const int c = c0 + cl;
    double m = 0.0;
    // eight loads in flight per thread (one per pass would leave the pass
    // loop bound by load latency: a workgroup per matrix is few workgroups)
    if (c < nrhs) {
#pragma unroll 8
      for (int i = r; i < n; i += rp) m = fmax(m, fabs(B[(int64_t)i * nrhs + c]));
    }
    for (int o = cpw; o < 64; o <<= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane < cpw) red[wave][lane] = m;
    __syncthreads();
    if (t < cpw && c0 + t < nrhs)
      bmax[mat * nrhs + c0 + t] = fmax(fmax(red[0][t], red[1][t]), fmax(red[2][t], red[3][t]));
    __syncthreads();
  }
}

__global__ void zero_int_kernel(int32_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void copy_masked_int_kernel(int64_t count, const int32_t* src, int32_t* dst, const int32_t* mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count && mask[i]) dst[i] = src[i];
}

__global__ void fill_kernel(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ double pow2_scale(double m) {
  // 2^-e with m in [2^e, 2^(e+1)); 1 for zero / non-finite.
  if (!(m > 0.0) || !isfinite(m)) return 1.0;
  return ldexp(1.0, -ilogb(m));
}

// Adds the wave's count of `flag` to *ctr with one atomic (every lane of the
// wave must call it).
__device__ __forceinline__ void wave_count(bool flag, int32_t* ctr) {
  const uint64_t m = __ballot(flag);
  if (ctr && m && (threadIdx.x & 63) == 0) atomicAdd(ctr, (int32_t)__popcll(m));
}

// State of a refined solve after bnorm_kernel (one thread per system): the
// column maxima of r and x cleared, berr = NaN, the first sweep's residual
// scales from ||b||, prev = +inf, info from the factorisation, active = 1 (a
// low-precision zero pivot: 2, for the FP64 fallback, counted in n_stop) and
// iters = 0.
__global__ void begin_state_kernel(int64_t count, int nrhs, const double* bmax, double* rmax, double* xmax,
                                   double* scale, double* berr, double* prev, const int32_t* finfo, int32_t* info,
                                   int32_t* active, int32_t* iters, int32_t* n_stop) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool stop = false;
  if (i < count) {
    for (int c = 0; c < nrhs; ++c) {
      const int64_t k = i * nrhs + c;
      rmax[k] = 0.0;
      xmax[k] = 0.0;
      berr[k] = NAN;
      scale[k] = pow2_scale(bmax[k]);
    }
    prev[i] = INFINITY;
    const int32_t f = finfo[i];
    info[i] = f;
    active[i] = f == 0 ? 1 : 2;
    iters[i] = 0;
    stop = f != 0;
  }
  wave_count(stop, n_stop);
}

// One more sweep for the systems still iterating; ctr (when given) is the
// active-system counter of the next test, cleared here (its last value has
// already been copied out on this stream).
__global__ void count_sweep_kernel(int64_t count, const int32_t* active, int32_t* iters, int32_t* ctr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ctr && i == 0) *ctr = 0;
  if (i < count && active[i] == 1) iters[i] += 1;
}

// R = B - A X in two passes. Pass 1 (residual_partial_kernel): one
// workgroup per (row block, matrix, 512-column chunk x 8-RHS chunk) stages
// its X chunk in LDS once and walks its rows in npass 16-row passes (4 rows per
// wave, two adjacent columns per lane in each of four 128-column slices,
// 16-byte loads); the first pass's A loads are issued before the X staging, so
// that traffic is in flight while X arrives; its 32 partial sums per lane are
// combined with an xor step and a halving butterfly and written as one
// partial per (row, column, chunk). The staged X keeps the two rows a lane
// reads (16 doubles) together in a 144-byte slot, so the lanes' 16-byte LDS
// reads start 36 banks apart and never collide. Pass 2 (residual_finish_kernel) adds the
// partials in chunk order (deterministic), forms r = b - sum and the per-column
// max |r|. The block of matrix rows with index 0 also reports max |x| per column.
constexpr int RES_RPW = 4, RES_ROWS = 4 * RES_RPW, RES_JC = 512, RES_NC = 8;
constexpr int RES_PASS = 4;  // least 16-row passes per workgroup over one staged X chunk
__global__ __launch_bounds__(256) void residual_partial_kernel(int n, int nrhs, int nchunk, int npass, const double* Ab,
                                                               const double* Xb, double* part, double* xmax,
                                                               const int32_t* active) {
  const int64_t mat = blockIdx.y;
  if (active && active[mat] != 1) return;
  const int ch = blockIdx.z % nchunk, zc = blockIdx.z / nchunk;
  const int c0 = zc * RES_NC;
  const int nc = min(RES_NC, nrhs - c0);
  const int jc = ch * RES_JC;
  const double* __restrict__ A = Ab + mat * (int64_t)n * n;
  const double* __restrict__ X = Xb + mat * (int64_t)n * nrhs;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int XP = 2 * RES_NC + 2;  // doubles per row-pair slot
  __shared__ double xs[RES_JC / 2][XP];
  const bool even = (n & 1) == 0;
  double a0[RES_JC / 128][RES_RPW], a1[RES_JC / 128][RES_RPW];
  auto load_a = [&](int row0) {
#pragma unroll
    for (int sub = 0; sub < RES_JC / 128; ++sub) {
      const int j0 = jc + sub * 128 + 2 * lane;
#pragma unroll
      for (int r = 0; r < RES_RPW; ++r) {
        const int row = row0 + r;
        a0[sub][r] = 0.0;
        a1[sub][r] = 0.0;
        if (row < n && j0 < n) {
          const double* ar = A + (int64_t)row * n;
          if (even) {
            const double2 v = *reinterpret_cast<const double2*>(ar + j0);
            a0[sub][r] = v.x;
            a1[sub][r] = v.y;
          } else {
            a0[sub][r] = ar[j0];
            a1[sub][r] = j0 + 1 < n ? ar[j0 + 1] : 0.0;
          }
        }
      }
    }
  };
  const int rowbase = blockIdx.x * (RES_ROWS * npass) + wave * RES_RPW;
  load_a(rowbase);
  double xm = 0.0;
#pragma unroll
  for (int q = 0; q < (RES_JC * RES_NC) / 256; ++q) {
    const int e = t + 256 * q;
    const int j = e / RES_NC, c = e % RES_NC;
    const double v = (jc + j < n && c < nc) ? X[(int64_t)(jc + j) * nrhs + c0 + c] : 0.0;
    xs[j >> 1][(j & 1) * RES_NC + c] = v;
    xm = fmax(xm, fabs(v));
  }
  if (blockIdx.x == 0) {
    const int cc = t % RES_NC;
    if (cc < nc) atomic_max_nonneg(xmax + mat * nrhs + c0 + cc, xm);
  }
  __syncthreads();
  for (int pass = 0; pass < npass; ++pass) {
    const int row0 = rowbase + pass * RES_ROWS;
    if (pass > 0) {
      if (row0 - wave * RES_RPW >= n) break;  // uniform over the workgroup
      load_a(row0);
    }
    double acc[RES_RPW][RES_NC];
#pragma unroll
    for (int r = 0; r < RES_RPW; ++r)
#pragma unroll
      for (int c = 0; c < RES_NC; ++c) acc[r][c] = 0.0;
#pragma unroll
    for (int sub = 0; sub < RES_JC / 128; ++sub) {
      const int jp = sub * 64 + lane;
#pragma unroll
      for (int c = 0; c < RES_NC; ++c) {
        const double x0 = xs[jp][c], x1 = xs[jp][RES_NC + c];
#pragma unroll
        for (int r = 0; r < RES_RPW; ++r) acc[r][c] = fma(a1[sub][r], x1, fma(a0[sub][r], x0, acc[r][c]));
      }
    }
    // 32 values per lane: fold the two half-waves, then a halving butterfly.
    double v[RES_RPW * RES_NC];
#pragma unroll
    for (int r = 0; r < RES_RPW; ++r)
#pragma unroll
      for (int c = 0; c < RES_NC; ++c) v[r * RES_NC + c] = acc[r][c] + __shfl_xor(acc[r][c], 32, 64);
#pragma unroll
    for (int h = 16; h >= 1; h >>= 1) {
      const bool up = (lane & h) != 0;
#pragma unroll
      for (int i = 0; i < h; ++i) {
        const double send = up ? v[i] : v[i + h];
        const double keep = up ? v[i + h] : v[i];
        v[i] = keep + __shfl_xor(send, h, 64);
      }
    }
    const int r = (lane & 31) / RES_NC, c = lane % RES_NC;
    const int row = row0 + r;
    if (lane < 32 && row < n && c < nc)
      part[((mat * nchunk + ch) * (int64_t)n + row) * nrhs + c0 + c] = v[0];
  }
}

// One thread per element. When nrhs divides 64 (and so the per-system
// element count), a wave covers whole rows of one system: its lanes of one
// column first reduce |r| with xor shuffles, so each column takes one atomic
// per wave instead of one per element.
__global__ void residual_finish_kernel(int64_t count, int n, int nrhs, int nchunk, const double* part,
                                       const double* Bb, double* Rb, double* rmax, const int32_t* active) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)n * nrhs;
  const int64_t mat = i / per, e = i - mat * per;
  const bool live = i < count * per && !(active && active[mat] != 1);
  double m = 0.0;
  if (live) {
    double s = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) s += part[(mat * nchunk + ch) * per + e];
    const double res = Bb[i] - s;
    Rb[i] = res;
    m = fabs(res);
  }
  const int lane = threadIdx.x & 63;
  if (64 % nrhs == 0 && per % 64 == 0) {
    // the wave is one system's (live is uniform over it); lanes l and
    // l + k * nrhs share a column
    for (int o = nrhs; o < 64; o <<= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (live && lane < nrhs) atomic_max_nonneg(rmax + mat * nrhs + (int)(e % nrhs), m);
  } else if (live) {
    atomic_max_nonneg(rmax + mat * nrhs + (int)(e % nrhs), m);
  }
}

// R = B - A X with the per-column maxima of |r| and |x|, for n <= 256: one
// workgroup per (system, 16 right-hand-side columns). The X chunk is staged in
// LDS (zero rows up to a multiple of 16, zero columns up to 16); each wave takes
// 16-row tiles of A and accumulates them with FP64 MFMA 16x16x4, loading A
// straight from global memory: a lane takes four consecutive entries of its row
// per 16-column chunk (MFMA q of the chunk contracts over k = 4 lk + q, the
// same k in the A and the X operand), four chunks in flight. A is read once;
// R, the maxima and B's loads are fused into the same pass.
constexpr int RSM_N = 256;
template <bool VEC>
__global__ __launch_bounds__(256) void residual_small_kernel(int n, int nrhs, const double* Ab, const double* Xb,
                                                             const double* Bb, double* Rb, double* rmax,
                                                             double* xmax, const int32_t* active) {
  constexpr int KU = 4;  // 16-column chunks per step
  const int64_t mat = blockIdx.x;
  if (active && active[mat] != 1) return;
  const int c0 = blockIdx.y * 16, nc = min(16, nrhs - c0);
  const int64_t per = (int64_t)n * nrhs;
  const double* __restrict__ A = Ab + mat * (int64_t)n * n;
  const double* __restrict__ X = Xb + mat * per;
  const double* __restrict__ B = Bb + mat * per;
  double* __restrict__ R = Rb + mat * per;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, lr = lane & 15, lk = lane >> 4;
  __shared__ double xs[RSM_N][16];
  __shared__ double red_r[4][16], red_x[4][16];
  const int np = (n + 15) & ~15;
  double xm = 0.0;  // column t & 15
  for (int e = t; e < np * 16; e += 256) {
    const int k = e >> 4, j = e & 15;
    const double v = (k < n && j < nc) ? X[(int64_t)k * nrhs + c0 + j] : 0.0;
    xs[k][j] = v;
    xm = fmax(xm, fabs(v));
  }
  xm = fmax(xm, __shfl_xor(xm, 16, 64));
  xm = fmax(xm, __shfl_xor(xm, 32, 64));
  if (lane < 16) red_x[wave][lane] = xm;
  __syncthreads();
  double rm = 0.0;  // column lr
  for (int r0 = wave * 16; r0 < n; r0 += 64) {
    const int row = min(r0 + lr, n - 1);
    const double* __restrict__ ar = A + (int64_t)row * n;
    // this lane's outputs: rows r0 + lk + 4 v, column lr
    double bv[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = min(r0 + lk + 4 * v, n - 1);
      bv[v] = lr < nc ? B[(int64_t)rr * nrhs + c0 + lr] : 0.0;
    }
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < n; k0 += 16 * KU) {
      double a[KU][4];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int k = k0 + 16 * u + 4 * lk;
        if constexpr (VEC) {
          // n even: a pair starting at an even index <= n - 2 lies inside the
          // row (and is 16-byte aligned); pairs past the row are clamped and
          // their values zeroed
          const double2 p0 = *reinterpret_cast<const double2*>(ar + min(k, n - 2));
          const double2 p1 = *reinterpret_cast<const double2*>(ar + min(k + 2, n - 2));
          a[u][0] = k < n ? p0.x : 0.0;
          a[u][1] = k + 1 < n ? p0.y : 0.0;
          a[u][2] = k + 2 < n ? p1.x : 0.0;
          a[u][3] = k + 3 < n ? p1.y : 0.0;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) a[u][q] = k + q < n ? ar[min(k + q, n - 1)] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int kb = k0 + 16 * u + 4 * lk;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          // chunks past np hold no staged rows: their A values are zero, and
          // the X operand is read from row 0 instead (finite)
          const int kx = kb + q < np ? kb + q : 0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][q], xs[kx][lr], acc, 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = r0 + lk + 4 * v;
      if (rr < n && lr < nc) {
        const double r = bv[v] - acc[v];
        R[(int64_t)rr * nrhs + c0 + lr] = r;
        rm = fmax(rm, fabs(r));
      }
    }
  }
  rm = fmax(rm, __shfl_xor(rm, 16, 64));
  rm = fmax(rm, __shfl_xor(rm, 32, 64));
  if (lane < 16) red_r[wave][lane] = rm;
  __syncthreads();
  if (t < nc) {
    double r = red_r[0][t], x = red_x[0][t];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      r = fmax(r, red_r[w][t]);
      x = fmax(x, red_x[w][t]);
    }
    rmax[mat * nrhs + c0 + t] = r;
    xmax[mat * nrhs + c0 + t] = x;
  }
}

// residual_small_kernel for a bucket of systems kept at their own order: the
// workgroup's system (bucket slot `mat`, system ids[mat]) has its n x n matrix
// packed at A + aoff[id] (row pitch n), while X, B and R are the bucket's
// padded N x nrhs blocks. The padded system is diag(A, I), so rows i >= n give
// r_i = b_i - x_i and rows i < n sum over k < n only.
__global__ __launch_bounds__(256) void residual_var_kernel(int N, int nrhs, const int32_t* ids, const int32_t* nn,
                                                           const int64_t* aoff, const double* Apk, const double* Xb,
                                                           const double* Bb, double* Rb, double* rmax, double* xmax,
                                                           const int32_t* active) {
  const int64_t mat = blockIdx.x;
  if (active && active[mat] != 1) return;
  const int id = ids[mat], n = nn[id];
  const int c0 = blockIdx.y * 16, nc = min(16, nrhs - c0);
  const int64_t per = (int64_t)N * nrhs;
  const double* __restrict__ A = Apk + aoff[id];
  const double* __restrict__ X = Xb + mat * per;
  const double* __restrict__ B = Bb + mat * per;
  double* __restrict__ R = Rb + mat * per;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, lr = lane & 15, lk = lane >> 4;
  __shared__ double xs[RSM_N][16];
  __shared__ double red_r[4][16], red_x[4][16];
  const int np = (N + 15) & ~15;
  double xm = 0.0;
  for (int e = t; e < np * 16; e += 256) {
    const int k = e >> 4, j = e & 15;
    const double v = (k < N && j < nc) ? X[(int64_t)k * nrhs + c0 + j] : 0.0;
    xs[k][j] = v;
    xm = fmax(xm, fabs(v));
  }
  xm = fmax(xm, __shfl_xor(xm, 16, 64));
  xm = fmax(xm, __shfl_xor(xm, 32, 64));
  if (lane < 16) red_x[wave][lane] = xm;
  __syncthreads();
  // pairs of doubles per load: a packed matrix starts on 8 bytes only, so
  // the pair type is 8-byte aligned (the device supports such vector loads)
  typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));
  constexpr int KV = 8;  // 16-column chunks in flight per step
  double rm = 0.0;
  for (int r0 = wave * 16; r0 < N; r0 += 64) {
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    if (r0 < n) {
      // rows past n read row n - 1: output row i depends on operand row i
      // only, and rows i >= n take the identity below
      const int row = min(r0 + lr, n - 1);
      const double* __restrict__ ar = A + (int64_t)row * n;
      for (int k0 = 0; k0 < n; k0 += 16 * KV) {
        // pairs (k, k + 1) and (k + 2, k + 3), each clamped to start at most
        // at n - 2 so it stays inside the row (n >= 2; n = 1 clamps to 0 and
        // reads the row's single entry twice); an index past the clamp
        // (k = n - 1 for odd n) takes the pair's second entry, and entries
        // past the row are zero in the operand
        d2u pr[KV][2];
#pragma unroll
        for (int u = 0; u < KV; ++u) {
          const int k = k0 + 16 * u + 4 * lk;
          if (n >= 2) {
            pr[u][0] = *reinterpret_cast<const d2u*>(ar + min(k, n - 2));
            pr[u][1] = *reinterpret_cast<const d2u*>(ar + min(k + 2, n - 2));
          } else {
            pr[u][0] = d2u{ar[0], 0.0};
            pr[u][1] = d2u{0.0, 0.0};
          }
        }
#pragma unroll
        for (int u = 0; u < KV; ++u) {
          const int k = k0 + 16 * u + 4 * lk;
          const int b0 = min(k, n - 2), b1 = min(k + 2, n - 2);
          const double a0 = k < n ? (k == b0 || n < 2 ? pr[u][0].x : pr[u][0].y) : 0.0;
          const double a1 = k + 1 < n ? pr[u][0].y : 0.0;
          const double a2 = k + 2 < n ? (k + 2 == b1 ? pr[u][1].x : pr[u][1].y) : 0.0;
          const double a3 = k + 3 < n ? pr[u][1].y : 0.0;
          const int kb = k0 + 16 * u + 4 * lk;
          const double av[4] = {a0, a1, a2, a3};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int kx = kb + q < np ? kb + q : 0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], xs[kx][lr], acc, 0, 0, 0);
          }
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = r0 + lk + 4 * v;
      if (rr < N && lr < nc) {
        const double ax = rr < n ? acc[v] : xs[rr][lr];
        const double r = B[(int64_t)rr * nrhs + c0 + lr] - ax;
        R[(int64_t)rr * nrhs + c0 + lr] = r;
        rm = fmax(rm, fabs(r));
      }
    }
  }
  rm = fmax(rm, __shfl_xor(rm, 16, 64));
  rm = fmax(rm, __shfl_xor(rm, 32, 64));
  if (lane < 16) red_r[wave][lane] = rm;
  __syncthreads();
  if (t < nc) {
    double r = red_r[0][t], x = red_x[0][t];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      r = fmax(r, red_r[w][t]);
      x = fmax(x, red_x[w][t]);
    }
    rmax[mat * nrhs + c0 + t] = r;
    xmax[mat * nrhs + c0 + t] = x;
  }
}

// diag(A_id, I) of order N for each bucket slot, from the packed matrices.
__global__ void pad_var_kernel(int64_t count, int N, const int32_t* ids, const int32_t* nn, const int64_t* aoff,
                               const double* Apk, double* out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)N * N;
  if (e >= count * per) return;
  const int64_t mat = e / per, w = e - mat * per;
  const int i = (int)(w / N), j = (int)(w - (int64_t)i * N);
  const int id = ids[mat], n = nn[id];
  out[e] = (i < n && j < n) ? Apk[aoff[id] + (int64_t)i * n + j] : (i == j ? 1.0 : 0.0);
}

// Convergence test per system (one thread per system). Records eta per column,
// decides active (1 = continue, 0 = converged, 2 = stopped without
// convergence), computes the next sweep's scales, clears the max buffers and
// counts systems still active (n_active) and newly stopped (n_stop).
__global__ void eta_kernel(int64_t count, int nrhs, double tol, int max_iter, const double* anorm,
                           const double* bmax, double* rmax, double* xmax, double* scale, double* prev,
                           double* berr, int32_t* active, const int32_t* iters, int32_t* n_active,
                           int32_t* n_stop) {
  const int64_t mat = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int state = 0;  // 1: iterates on, 2: stopped here
  if (mat < count && active[mat] == 1) {
    double worst = 0.0;
    for (int c = 0; c < nrhs; ++c) {
      const int64_t k = mat * nrhs + c;
      const double rm = rmax[k], xm = xmax[k];
      const double den = anorm[mat] * xm + bmax[k];
      const double eta = den > 0.0 ? rm / den : (rm > 0.0 ? INFINITY : 0.0);
      if (berr) berr[k] = eta;
      worst = (eta > worst || isnan(eta)) ? eta : worst;
      scale[k] = pow2_scale(rm);
      rmax[k] = 0.0;
      xmax[k] = 0.0;
    }
    if (worst <= tol) {
      active[mat] = 0;
    } else if (iters[mat] >= max_iter || !(worst < 0.5 * prev[mat])) {
      active[mat] = 2;
      state = 2;
    } else {
      state = 1;
    }
    prev[mat] = worst;
  }
  // one atomic per wave and counter (thousands of systems would otherwise
  // serialise on the two addresses)
  wave_count(state == 1, n_active);
  wave_count(state == 2, n_stop);
}

// Systems that did not converge (active == 2): mask for the FP64 fallback.
__global__ void fallback_mask_kernel(int64_t count, const int32_t* active, int32_t* mask, int32_t* iters) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int fb = active[i] == 2 ? 1 : 0;
  mask[i] = fb;
  if (fb) iters[i] = -1;
}

// Per host thread: pinned flags for the convergence read-back and the event
// that marks their arrival (kept until the process ends).
struct FlagSlot {
  int32_t* host = nullptr;
  hipEvent_t ev = nullptr;
};
int flag_slot(FlagSlot** out) {
  thread_local FlagSlot slot;
  if (!slot.host) {
    void* p = nullptr;
    NPPH_HIP(hipHostMalloc(&p, 64, hipHostMallocDefault));
    hipEvent_t e = nullptr;
    const hipError_t err = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (err != hipSuccess) {
      (void)hipHostFree(p);
      NPPH_HIP(err);
    }
    slot.host = static_cast<int32_t*>(p);
    slot.ev = e;
  }
  *out = &slot;
  return NPPH_OK;
}

__global__ void copy_masked_kernel(int64_t count, int64_t per, const double* src, double* dst, const int32_t* mask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count * per) return;
  if (mask[i / per]) dst[i] = src[i];
}

inline dim3 g1(int64_t n, int b = 256) { return dim3((unsigned)ceil_div(n, b)); }

}  // namespace

// Residual + per-(system, column) maxima of |r| and |x| for the current X.
int residual_launch(int64_t count, int n, int nrhs, const double* A, const double* X, const double* B, double* R,
                    double* rmax, double* xmax, const int32_t* active, hipStream_t s) {
  if (n <= RSM_N && count <= 2147483647LL && ceil_div(nrhs, 16) <= 65535) {
    ProfScope ps(kProfResidual, 2.0 * count * n * n * nrhs, 8.0 * count * ((double)n * n + 3.0 * n * nrhs), s);
    dim3 grid((unsigned)count, (unsigned)ceil_div(nrhs, 16));
    const bool vec = (n % 2 == 0) && n >= 4 && (reinterpret_cast<uintptr_t>(A) % 16 == 0);
    if (vec)
      hipLaunchKernelGGL(residual_small_kernel<true>, grid, dim3(256), 0, s, n, nrhs, A, X, B, R, rmax, xmax,
                         active);
    else
      hipLaunchKernelGGL(residual_small_kernel<false>, grid, dim3(256), 0, s, n, nrhs, A, X, B, R, rmax, xmax,
                         active);
    NPPH_HIP(hipGetLastError());
    return NPPH_OK;
  }
  const int nchunk = (int)ceil_div(n, RES_JC), nz = (int)ceil_div(nrhs, RES_NC);
  double* part;
  NPPH_TRY(workspace(kWsPart, sizeof(double) * (size_t)count * nchunk * n * nrhs, (void**)&part));
  // rows per workgroup: enough that the grid is one wave of resident
  // workgroups (at least RES_PASS passes each): long-lived workgroups stream
  // A at a higher rate than many short ones, and a partial second wave of
  // workgroups leaves most CUs idle at the end (64 x 2048: 1.40 -> 1.27 ms
  // for three residuals)
  static const int64_t slots = [] {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, residual_partial_kernel, 256, 0) != hipSuccess ||
        per_cu <= 0)
      per_cu = 1;
    return (int64_t)(cu_count() > 0 ? cu_count() : 256) * per_cu;
  }();
  const int64_t pairs = count * nchunk * nz;
  const int64_t blocks_per_pair = std::max<int64_t>(1, slots / pairs);
  const int npass = (int)std::max<int64_t>(RES_PASS, ceil_div(ceil_div(n, RES_ROWS), blocks_per_pair));
  dim3 grid((unsigned)ceil_div(n, RES_ROWS * npass), (unsigned)count, (unsigned)(nchunk * nz));
  ProfScope ps(kProfResidual, 2.0 * count * n * n * nrhs, 8.0 * count * ((double)n * n + 3.0 * n * nrhs), s);
  hipLaunchKernelGGL(residual_partial_kernel, grid, dim3(256), 0, s, n, nrhs, nchunk, npass, A, X, part, xmax, active);
  hipLaunchKernelGGL(residual_finish_kernel, g1(count * n * nrhs), dim3(256), 0, s, count, n, nrhs, nchunk, part, B,
                     R, rmax, active);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// Bytes of factor storage for a uniform stack (see RefineFactors).
size_t refine_factor_bytes(int precision, int64_t count, int n) {
  // the pieces refine_factor_carve takes, each rounded up to 256 bytes
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t nn = (size_t)count * n * n;
  const size_t lu = precision == NPPH_PREC_FP64 ? nn * 8 : nn * 4;
  return al(lu) + al(8 * (size_t)count) + 2 * al(sizeof(int32_t) * (size_t)count * n) +
         al(sizeof(int32_t) * (size_t)count);
}

// Carves a RefineFactors out of one caller-owned device buffer of
// refine_factor_bytes() bytes.
void refine_factor_carve(int precision, int64_t count, int n, void* buf, RefineFactors& f) {
  char* p = static_cast<char*>(buf);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 255) & ~(size_t)255;
    return q;
  };
  f.precision = precision;
  f.count = count;
  f.n = n;
  const size_t nn = (size_t)count * n * n;
  f.lu64 = nullptr;
  f.lu32 = nullptr;
  if (precision == NPPH_PREC_FP64) f.lu64 = (double*)take(nn * 8);
  else f.lu32 = (float*)take(nn * 4);
  f.anorm = (double*)take(8 * (size_t)count);
  f.ipiv = (int32_t*)take(sizeof(int32_t) * (size_t)count * n);
  f.perm = (int32_t*)take(sizeof(int32_t) * (size_t)count * n);
  f.info = (int32_t*)take(sizeof(int32_t) * (size_t)count);
}

int refine_factor(const npph_refine_opts& o, const double* A, RefineFactors& f, hipStream_t s, bool prepped) {
  const int64_t count = f.count;
  const int n = f.n;
  if (count <= 0 || n <= 0) return NPPH_OK;
  if (count > 65535) { set_error("refined solve: count > 65535"); return NPPH_ERR_INVALID; }
  if (!prepped) hipLaunchKernelGGL(fill_kernel, g1(count), dim3(256), 0, s, f.anorm, count, 0.0);
  if (o.precision == NPPH_PREC_FP64) {
    const int64_t per_a = (int64_t)n * n;
    if (!prepped) {
      hipLaunchKernelGGL(demote_kernel<double>, g1(ceil_div(count * n, kDemoteRows) * 64), dim3(256), 0, s, count,
                         n, A, f.lu64, f.anorm);
    }
    if (n <= 64) {
      UniformBatch u{count, n, f.lu64, per_a, f.lu64, f.ipiv, f.info, 0, nullptr, 0, nullptr};
      NPPH_TRY(small_lu_launch(u, s));
      NPPH_TRY(ipiv_to_perm(count, n, f.ipiv, f.perm, s));
    } else {
      NPPH_TRY(blocked_getrf_perm<double>(count, n, f.lu64, per_a, f.ipiv, f.perm, f.info, 0, nullptr, s));
    }
    NPPH_HIP(hipGetLastError());
    return NPPH_OK;
  }
  if (o.precision != NPPH_PREC_FP32 && o.precision != NPPH_PREC_BF16) {
    set_error("refined solve: unknown precision %d", o.precision);
    return NPPH_ERR_INVALID;
  }
  const int terms = o.precision == NPPH_PREC_BF16 ? (o.bf16_terms > 0 ? o.bf16_terms : 6) : 0;
  if (!prepped) hipLaunchKernelGGL(demote_kernel<float>, g1(ceil_div(count * n, kDemoteRows) * 64), dim3(256), 0, s, count, n, A, f.lu32, f.anorm);
  NPPH_HIP(hipGetLastError());
  return blocked_getrf_perm<float>(count, n, f.lu32, (int64_t)n * n, f.ipiv, f.perm, f.info, terms, nullptr, s);
}

namespace {

// Plain FP64 solve of the systems selected by `mask` (all when masked is
// false), from a fresh FP64 factorisation of A.
int fp64_solve(RefineRun& r, bool masked, hipStream_t s) {
  const int64_t count = r.count, per_a = (int64_t)r.n * r.n, per_b = (int64_t)r.n * r.nrhs;
  const int64_t nb_all = count * per_b;
  double *Aw, *Bw;
  int32_t *ipiv, *info64;
  NPPH_TRY(workspace(kWsA, sizeof(double) * (size_t)(count * per_a), (void**)&Aw));
  NPPH_TRY(workspace(kWsB, sizeof(double) * (size_t)nb_all, (void**)&Bw));
  NPPH_TRY(workspace(kWsPiv, sizeof(int32_t) * (size_t)(count * r.n), (void**)&ipiv));
  NPPH_TRY(workspace(kWsInfo, sizeof(int32_t) * (size_t)count, (void**)&info64));
  if (r.var.A)
    hipLaunchKernelGGL(pad_var_kernel, g1(count * per_a), dim3(256), 0, s, count, r.n, r.var.ids, r.var.nn,
                       r.var.aoff, r.var.A, Aw);
  else
    NPPH_HIP(hipMemcpyAsync(Aw, r.A, sizeof(double) * count * per_a, hipMemcpyDeviceToDevice, s));
  NPPH_HIP(hipMemcpyAsync(Bw, r.B, sizeof(double) * nb_all, hipMemcpyDeviceToDevice, s));
  NPPH_TRY(gesv_dev_impl(count, r.n, r.nrhs, Aw, per_a, Bw, per_b, ipiv, info64, s));
  if (masked) {
    hipLaunchKernelGGL(copy_masked_kernel, g1(nb_all), dim3(256), 0, s, count, per_b, Bw, r.X, r.mask);
    hipLaunchKernelGGL(copy_masked_int_kernel, g1(count), dim3(256), 0, s, count, info64, r.info, r.mask);
  } else {
    NPPH_HIP(hipMemcpyAsync(r.X, Bw, sizeof(double) * nb_all, hipMemcpyDeviceToDevice, s));
    NPPH_HIP(hipMemcpyAsync(r.info, info64, sizeof(int32_t) * count, hipMemcpyDeviceToDevice, s));
  }
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// R = B - A X of the run's systems (A uniform, or packed per system).
int residual_run(RefineRun& r, hipStream_t s) {
  if (!r.var.A) return residual_launch(r.count, r.n, r.nrhs, r.A, r.X, r.B, r.R, r.rmax, r.xmax, r.active, s);
  if (r.n > RSM_N || ceil_div(r.nrhs, 16) > 65535) {
    set_error("refined solve: packed matrices need n <= %d", RSM_N);
    return NPPH_ERR_INVALID;
  }
  ProfScope ps(kProfResidual, 2.0 * r.count * r.n * r.n * r.nrhs,
               8.0 * r.count * ((double)r.n * r.n + 3.0 * r.n * r.nrhs), s);
  hipLaunchKernelGGL(residual_var_kernel, dim3((unsigned)r.count, (unsigned)ceil_div(r.nrhs, 16)), dim3(256), 0, s,
                     r.n, r.nrhs, r.var.ids, r.var.nn, r.var.aoff, r.var.A, r.X, r.B, r.R, r.rmax, r.xmax, r.active);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// eta of the systems flagged active (1), without further sweeps.
int report(RefineRun& r, hipStream_t s) {
  NPPH_TRY(residual_run(r, s));
  hipLaunchKernelGGL(eta_kernel, g1(r.count), dim3(256), 0, s, r.count, r.nrhs, INFINITY, 0, r.f.anorm, r.bmax,
                     r.rmax, r.xmax, r.scale, r.prev, r.berr_p, r.active, r.iters, r.ctr, nullptr);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

int max_iter_of(const npph_refine_opts& o) { return o.max_iter > 0 ? o.max_iter : 10; }

// Residual and convergence test of loop iteration `it`: systems that meet
// the tolerance (or have run max_iter sweeps, or stagnate) leave state 1;
// ctr[0] receives the number still iterating, ctr[1] accumulates those
// stopped without convergence.
int eta_step(RefineRun& r, hipStream_t s) {
  NPPH_TRY(residual_run(r, s));
  hipLaunchKernelGGL(eta_kernel, g1(r.count), dim3(256), 0, s, r.count, r.nrhs, r.o.tol, max_iter_of(r.o), r.f.anorm,
                     r.bmax, r.rmax, r.xmax, r.scale, r.prev, r.berr_p, r.active, r.iters, r.ctr, r.ctr + 1);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// The correction of loop iteration `it` for the systems still iterating
// (none past max_iter).
int corr_step(RefineRun& r, int it, hipStream_t s) {
  if (it > max_iter_of(r.o)) {
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, s, r.ctr, 1);  // the next test's counter
    return NPPH_OK;
  }
  // r.ctr[0]: systems still iterating after the last test (cleared below)
  NPPH_TRY(getrs_refine(r.count, r.n, r.nrhs, r.f.lu32, (int64_t)r.n * r.n, r.f.perm, r.R, r.scale, r.Dlo, r.X, 0,
                        r.active, s, r.f.perm, r.ctr));
  hipLaunchKernelGGL(count_sweep_kernel, g1(r.count), dim3(256), 0, s, r.count, r.active, r.iters, r.ctr);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace

int refine_begin(RefineRun& r, hipStream_t s) {
  const int64_t count = r.f.count;
  const int n = r.f.n;
  r.count = count;
  r.n = n;
  r.it = 0;
  r.eta_issued = false;
  r.done = count <= 0 || n <= 0 || r.nrhs <= 0;
  if (r.done) return NPPH_OK;
  const int nrhs = r.nrhs;
  const int64_t per_b = (int64_t)n * nrhs;
  const int64_t nb_all = count * per_b, nk = count * nrhs;

  // Bookkeeping arrays share one workspace slot.
  const size_t dbl_cnt = (size_t)(5 * nk + count);  // bmax, rmax, xmax, scale, berr, prev
  const size_t int_cnt = (size_t)(4 * count + 2);   // info, active, iters, mask, counters
  char* misc;
  NPPH_TRY(workspace(kWsNorm, dbl_cnt * 8 + int_cnt * 4 + 64, (void**)&misc));
  r.bmax = (double*)misc;
  r.rmax = r.bmax + nk;
  r.xmax = r.rmax + nk;
  r.scale = r.xmax + nk;
  double* berr_ws = r.scale + nk;
  r.prev = berr_ws + nk;
  r.info = (int32_t*)(r.prev + count);
  r.active = r.info + count;
  r.iters = r.active + count;
  r.mask = r.iters + count;
  r.ctr = r.mask + count;  // [0] = active count, [1] = fallback count
  r.berr_p = r.berr_out ? r.berr_out : berr_ws;
  NPPH_TRY(workspace(kWsR, sizeof(double) * (size_t)nb_all, (void**)&r.R));

  hipLaunchKernelGGL(bnorm_kernel, dim3((unsigned)count), dim3(256), 0, s, n, nrhs, r.B, r.bmax, r.ctr);
  const bool fp64 = r.o.precision == NPPH_PREC_FP64;
  hipLaunchKernelGGL(begin_state_kernel, g1(count), dim3(256), 0, s, count, nrhs, r.bmax, r.rmax, r.xmax, r.scale,
                     r.berr_p, r.prev, r.f.info, r.info, r.active, r.iters, fp64 ? nullptr : r.ctr + 1);
  NPPH_HIP(hipGetLastError());

  if (r.o.precision == NPPH_PREC_FP64) {
    NPPH_TRY(blocked_getrs_perm<double>(count, n, nrhs, r.f.lu64, (int64_t)n * n, r.f.perm, r.B, per_b, r.X, per_b,
                                        r.f.info, nullptr, s,
                                        factor_rows_mapped(NPPH_PREC_FP64, n) ? r.f.perm : nullptr));
    NPPH_TRY(report(r, s));
    r.done = true;
    return NPPH_OK;
  }
  NPPH_TRY(workspace(kWsD, sizeof(float) * (size_t)nb_all, (void**)&r.Dlo));
  // Sweep 0 (x = 0, r = b).
  NPPH_TRY(getrs_refine(count, n, nrhs, r.f.lu32, (int64_t)n * n, r.f.perm, r.B, r.scale, r.Dlo, r.X, 1, r.active,
                        s, r.f.perm));
  hipLaunchKernelGGL(count_sweep_kernel, g1(count), dim3(256), 0, s, count, r.active, r.iters, r.ctr);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

int refine_steps(RefineRun& r, int budget, hipStream_t s) {
  if (r.done) return NPPH_OK;
  for (int k = 0; k < budget; ++k) {
    if (!r.eta_issued) NPPH_TRY(eta_step(r, s));
    r.eta_issued = false;
    NPPH_TRY(corr_step(r, r.it, s));
    ++r.it;
  }
  NPPH_TRY(eta_step(r, s));
  r.eta_issued = true;
  return NPPH_OK;
}

int refine_loop(RefineRun& r, int32_t* n_fb, hipStream_t s) {
  *n_fb = 0;
  if (r.done) return NPPH_OK;
  FlagSlot* fs;
  NPPH_TRY(flag_slot(&fs));
  // Each test's counts (systems still active, systems stopped so far) come
  // back through pinned memory while the next sweep is already queued: that
  // sweep only touches systems still active, so when none is left it changes
  // nothing, and the device never idles for the host's round trip.
  for (;;) {
    if (!r.eta_issued) NPPH_TRY(eta_step(r, s));
    r.eta_issued = false;
    NPPH_HIP(hipMemcpyAsync(fs->host, r.ctr, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    NPPH_HIP(hipEventRecord(fs->ev, s));
    // eta stops every system after max_iter sweeps, so the counts reach zero
    // by then; the bound only guards the loop
    NPPH_TRY(corr_step(r, r.it, s));
    NPPH_HIP(hipEventSynchronize(fs->ev));
    if (fs->host[0] == 0 || r.it > max_iter_of(r.o)) break;
    ++r.it;
  }
  *n_fb = fs->host[1];
  return NPPH_OK;
}

int refine_end(RefineRun& r, int32_t n_fb, hipStream_t s) {
  if (r.count <= 0 || r.n <= 0 || r.nrhs <= 0) return NPPH_OK;
  if (r.o.precision != NPPH_PREC_FP64 && r.o.fallback_fp64 && n_fb > 0) {
    hipLaunchKernelGGL(fallback_mask_kernel, g1(r.count), dim3(256), 0, s, r.count, r.active, r.mask, r.iters);
    NPPH_TRY(fp64_solve(r, true, s));
    NPPH_HIP(hipMemcpyAsync(r.active, r.mask, sizeof(int32_t) * r.count, hipMemcpyDeviceToDevice, s));
    NPPH_TRY(report(r, s));
  }
  if (r.iters_out)
    NPPH_HIP(hipMemcpyAsync(r.iters_out, r.iters, sizeof(int32_t) * r.count, hipMemcpyDeviceToDevice, s));
  if (r.info_out) NPPH_HIP(hipMemcpyAsync(r.info_out, r.info, sizeof(int32_t) * r.count, hipMemcpyDeviceToDevice, s));
  return NPPH_OK;
}

int refine_solve(const npph_refine_opts& o, const double* A, const RefineFactors& f, int nrhs, const double* B,
                 double* X, double* berr, int32_t* iters_out, int32_t* info_out, hipStream_t s) {
  RefineRun r;
  r.o = o;
  r.A = A;
  r.f = f;
  r.nrhs = nrhs;
  r.B = B;
  r.X = X;
  r.berr_out = berr;
  r.iters_out = iters_out;
  r.info_out = info_out;
  NPPH_TRY(refine_begin(r, s));
  int32_t n_fb = 0;
  NPPH_TRY(refine_loop(r, &n_fb, s));
  return refine_end(r, n_fb, s);
}

int refined_impl(const npph_refine_opts& o, int64_t count, int n, int nrhs, const double* A, const double* B,
                 double* X, double* berr, int32_t* iters_out, int32_t* info_out, hipStream_t s) {
  if (count <= 0 || n <= 0 || nrhs <= 0) return NPPH_OK;
  void* buf;
  NPPH_TRY(workspace(kWsLo, refine_factor_bytes(o.precision, count, n), &buf));
  RefineFactors f;
  refine_factor_carve(o.precision, count, n, buf, f);
  NPPH_TRY(refine_factor(o, A, f, s));
  return refine_solve(o, A, f, nrhs, B, X, berr, iters_out, info_out, s);
}

}  // namespace npph

extern "C" int npph_gesv_refined_batched_dev(const npph_refine_opts* opts, int64_t count, int64_t n, int64_t nrhs,
                                             const double* A, const double* B, double* X, double* berr,
                                             int32_t* iters, int32_t* info, void* stream) {
  NPPH_TRY(npph::ensure_init());
  if (!opts || count < 0 || n < 0 || nrhs < 0 || n > npph::max_blocked_n()) {
    npph::set_error("refined solve: bad arguments");
    return NPPH_ERR_INVALID;
  }
  if (B == X) {
    npph::set_error("refined solve: B and X must not alias");
    return NPPH_ERR_INVALID;
  }
  const hipStream_t s = stream ? (hipStream_t)stream : npph::default_stream();
  npph::CallScope cs(s);
  return npph::refined_impl(*opts, count, (int)n, (int)nrhs, A, B, X, berr, iters, info, s);
}
