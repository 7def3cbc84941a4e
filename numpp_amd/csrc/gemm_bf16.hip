// Split-BF16 trailing update for the low-precision LU: C -= A * B with FP32
// operands and FP32 accumulation on the gfx950 BF16 MFMA
// (v_mfma_f32_16x16x32_bf16: A[i][k] with i = lane%16, k = 8*(lane/16)+0..7;
// B[k][j] with j = lane%16, same k; four FP32 outputs C[4*(lane/16)+v][lane%16]).
//
// Two phases per update:
//  1. split: each operand is scaled by a power of two per panel (a 64-row block
//     of A spanning all of K, a 64-column block of B spanning all of K) so the
//     panel's largest magnitude lands in [1, 2), then split into BF16 planes
//     x = h + m + l (h = bf16(x), m = bf16(x - h), l = bf16(x - h - m)) and
//     stored K-contiguous (A row-major [row][k], B transposed [col][k]), K padded
//     with zeros to a multiple of 32 and rows/columns to a multiple of 128;
//  2. product: a plain BF16 MFMA GEMM over the planes, 128 x 128 tiles, 32-deep
//     K slices through one XOR-swizzled LDS slice, the partial products of one slice
//     accumulated smallest first into one FP32 accumulator per output; the
//     epilogue unscales by 2^(eA + eB) of the output's panels and subtracts.
// terms = 1 keeps only h (plain BF16); 3 uses two pieces (h, m) and the
// products hh, hm, mh (about 16 significant bits); 6 uses three pieces and hh,
// hm, mh, hl, mm, lh (about FP32 accuracy).
//
// The operands are split once per update instead of once per output tile, so
// the product phase is a pure MFMA stream.
#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int SPAN = 64;       // panel: rows of A / columns of B sharing one scale
constexpr int TM = 128, TN = 128, BK = 32;
// LDS rows are BK bf16 = four 16-byte chunks, unpadded; chunk c of row r is
// stored at c ^ ((r >> 2) & 3), so the 16 rows one fragment read covers and
// the rows a staging write covers fall in distinct banks.
__device__ __forceinline__ int swz(int r, int c) { return (c ^ ((r >> 2) & 3)) * 8; }
constexpr int NTH = 256;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Power-of-two exponent putting the panel's largest magnitude in [1, 2); a
// zero, Inf or NaN panel is not scaled (Inf and NaN then reach C as they are),
// and exponents are clamped so that the scale and its inverse stay finite.
__device__ __forceinline__ int panel_exp(float amax) {
  return (amax > 0.f && isfinite(amax)) ? min(max(ilogbf(amax), -126), 126) : 0;
}

__device__ __forceinline__ float block_max(float v, float* red) {
#pragma unroll
  for (int msk = 1; msk < 64; msk <<= 1) v = fmaxf(v, __shfl_xor(v, msk, 64));
  const int t = threadIdx.x;
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  v = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  return v;
}

// Splits x into NP BF16 pieces.
template <int NP>
__device__ __forceinline__ void split_val(float x, __bf16 (&o)[3]) {
  const __bf16 h = (__bf16)x;
  o[0] = h;
  if constexpr (NP >= 2) {
    const float r1 = x - (float)h;
    const __bf16 m = (__bf16)r1;
    o[1] = m;
    if constexpr (NP >= 3) o[2] = (__bf16)(r1 - (float)m);
  }
}

// A (M x K, row stride lda, matrix stride sA): one workgroup per 64-row panel.
// Planes: P[((bz * NP + p) * Mp + row) * Kp + k]; exponents eA[bz * (Mp/64) + panel].
template <int NP>
__global__ __launch_bounds__(NTH) void split_rows_kernel(int64_t M, int K, int Kp, const float* __restrict__ Ab,
                                                         int64_t lda, int64_t sA, __bf16* __restrict__ P, int64_t Mp,
                                                         int* __restrict__ eA, RowMaps maps) {
  __shared__ float red[4];
  const int64_t bz = blockIdx.y;
  const int panel = blockIdx.x;
  const float* A = Ab + bz * sA;
  const int t = threadIdx.x;
  const int r = t >> 2, q = t & 3;  // 4 threads per row
  const int64_t row = (int64_t)panel * SPAN + r;
  const bool live = row < M;
  const int64_t lrow = live ? row : 0;
  const float* ar = A + (maps.a ? (int64_t)maps.a[bz * maps.stride + lrow] : lrow) * lda;
  // 16-byte loads when every row starts on a 16-byte boundary
  const bool vec = ((reinterpret_cast<uintptr_t>(Ab) | (uintptr_t)(lda * 4) | (uintptr_t)(sA * 4)) & 15) == 0;
  float amax = 0.f;
  if (live) {
    const int K4 = vec ? (K & ~3) : 0;
    for (int k = 4 * q; k < K4; k += 16) {
      const float4 x = *reinterpret_cast<const float4*>(ar + k);
      amax = fmaxf(amax, fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w))));
    }
    for (int k = K4 + q; k < K; k += 4) amax = fmaxf(amax, fabsf(ar[k]));
  }
  amax = block_max(amax, red);
  const int e = panel_exp(amax);
  if (t == 0) eA[bz * (Mp / SPAN) + panel] = e;
  const float sc = ldexpf(1.f, -e);
  // thread q writes k = 8 * (q + 4 j) .. + 7 as 16-byte vectors
  for (int k0 = 8 * q; k0 < Kp; k0 += 32) {
    bf16x8 v[3];
    float a8[8];
    if (vec && k0 + 8 <= K) {
      const float4 x0 = *reinterpret_cast<const float4*>(ar + k0);
      const float4 x1 = *reinterpret_cast<const float4*>(ar + k0 + 4);
      a8[0] = x0.x; a8[1] = x0.y; a8[2] = x0.z; a8[3] = x0.w;
      a8[4] = x1.x; a8[5] = x1.y; a8[6] = x1.z; a8[7] = x1.w;
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u) a8[u] = ar[min(k0 + u, K - 1)];  // clamped: padding reads a real entry
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u;
      const float a = a8[u];
      const float x = (live && k < K) ? a * sc : 0.f;
      __bf16 o[3];
      split_val<NP>(x, o);
#pragma unroll
      for (int p = 0; p < NP; ++p) v[p][u] = o[p];
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
      *reinterpret_cast<bf16x8*>(P + ((bz * NP + p) * Mp + row) * Kp + k0) = v[p];
  }
}

// B (K x N, row stride ldb, matrix stride sB): one workgroup per 64-column
// panel, transposed through LDS in 64 x 64 chunks.
// Planes: P[((bz * NP + p) * Np + col) * Kp + k]; exponents eB[bz * (Np/64) + panel].
template <int NP>
__global__ __launch_bounds__(NTH) void split_cols_kernel(int64_t N, int K, int Kp, const float* __restrict__ Bb,
                                                         int64_t ldb, int64_t sB, __bf16* __restrict__ P, int64_t Np,
                                                         int* __restrict__ eB, RowMaps maps) {
  __shared__ float red[4];
  __shared__ float tile[SPAN][SPAN + 1];  // [k][col]
  // the B rows' physical indices, read once into LDS (K <= MAXK), so the
  // row loads below wait for an LDS read rather than a global one
  constexpr int MAXK = 2048;
  __shared__ int rows_s[MAXK];
  const int64_t bz = blockIdx.y;
  const int panel = blockIdx.x;
  const float* B = Bb + bz * sB;
  const int32_t* mapb = maps.b ? maps.b + bz * maps.stride : nullptr;
  const bool staged = mapb && K <= MAXK;
  const int t = threadIdx.x;
  if (staged) {
    for (int k = t; k < K; k += NTH) rows_s[k] = mapb[k];
    __syncthreads();
  }
  auto brow = [&](int k) { return (int64_t)(staged ? rows_s[k] : (mapb ? mapb[k] : k)) * ldb; };
  const int c = t & 63, kr = t >> 6;  // read: column c, rows kr, kr + 4, ...
  const int64_t col = (int64_t)panel * SPAN + c;
  const bool live = col < N;
  float amax = 0.f;
  if (live)
    for (int k = kr; k < K; k += 4) amax = fmaxf(amax, fabsf(B[brow(k) + col]));
  amax = block_max(amax, red);
  const int e = panel_exp(amax);
  if (t == 0) eB[bz * (Np / SPAN) + panel] = e;
  const float sc = ldexpf(1.f, -e);
  // write: column wc, k = kc0 .. kc0 + 15 of the chunk (two 16-byte vectors per plane)
  const int wc = t >> 2, kq = (t & 3) * 16;
  for (int kb = 0; kb < Kp; kb += SPAN) {
    __syncthreads();  // previous chunk's tile reads are done
#pragma unroll
    for (int u = 0; u < SPAN / 4; ++u) {
      const int k = kb + kr + 4 * u;
      const float b = B[brow(min(k, K - 1)) + (live ? col : 0)];  // clamped
      tile[kr + 4 * u][c] = (live && k < K) ? b * sc : 0.f;
    }
    __syncthreads();
    const int64_t wcol = (int64_t)panel * SPAN + wc;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k0 = kb + kq + 8 * h;
      if (k0 >= Kp) continue;
      bf16x8 v[3];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        __bf16 o[3];
        split_val<NP>(tile[kq + 8 * h + u][wc], o);
#pragma unroll
        for (int p = 0; p < NP; ++p) v[p][u] = o[p];
      }
#pragma unroll
      for (int p = 0; p < NP; ++p)
        *reinterpret_cast<bf16x8*>(P + ((bz * NP + p) * Np + wcol) * Kp + k0) = v[p];
    }
  }
}

// C -= 2^(eA + eB) * sum over products of the planes. Tile 128 x 128, four waves
// in a 2 x 2 grid of 64 x 64 sub-tiles (4 x 4 MFMA tiles each).
template <int NP>
__global__ __launch_bounds__(NTH) void bf16p_gemm_kernel(int64_t M, int64_t N, int Kp, const __bf16* __restrict__ PA,
                                                         int64_t Mp, const int* __restrict__ eA,
                                                         const __bf16* __restrict__ PB, int64_t Np,
                                                         const int* __restrict__ eB, float* __restrict__ Cb,
                                                         int64_t ldc, int64_t sC, int tiles_m, int tiles_n,
                                                         RowMaps maps) {
  // one LDS slice (NP = 3: 49 KB, so two workgroups share a CU and one's
  // MFMAs cover the other's global-load latency)
  __shared__ __bf16 As[NP][TM][BK];
  __shared__ __bf16 Bs[NP][TN][BK];
  // Work order: workgroups are dispatched round-robin over the 8 XCDs (each
  // with its own L2); the permutation below gives XCD x a contiguous eighth of
  // the (matrix, tile) sequence, so whole matrices stay on one XCD, and each
  // matrix's tiles go in groups of four tile rows, so the workgroups running
  // side by side share their A and B panels' K slices in that L2.
  const int64_t total = gridDim.x;
  int64_t idx = blockIdx.x;
  if ((total & 7) == 0) idx = (idx >> 3) + (idx & 7) * (total >> 3);
  const int64_t tiles = (int64_t)tiles_m * tiles_n;
  const int64_t bz = idx / tiles;
  int tm, tn;
  {
    constexpr int G = 4;
    const int r = (int)(idx - bz * tiles);
    const int g = r / (G * tiles_n), first = g * G, gs = min(tiles_m - first, G);
    const int q = r - g * G * tiles_n;
    tm = first + q % gs;
    tn = q / gs;
  }
  const int64_t m0 = (int64_t)tm * TM, n0 = (int64_t)tn * TN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  // staging: thread t copies the 16-byte k chunk t % 4 of rows t / 4 and
  // t / 4 + 64 of each plane (four neighbouring lanes read a row's 64-byte
  // slice in one piece)
  const int sr = t >> 2, sk = (t & 3) * 8;
  const __bf16* pa = PA + (bz * NP * Mp + m0 + sr) * Kp + sk;
  const __bf16* pb = PB + (bz * NP * Np + n0 + sr) * Kp + sk;
  const int64_t plA = Mp * Kp, plB = Np * Kp;

  bf16x8 ra[NP][2], rb[NP][2];
  auto load = [&](int k0) {
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        ra[p][h] = *reinterpret_cast<const bf16x8*>(pa + p * plA + 64 * h * (int64_t)Kp + k0);
        rb[p][h] = *reinterpret_cast<const bf16x8*>(pb + p * plB + 64 * h * (int64_t)Kp + k0);
      }
  };
  auto store = [&]() {
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        *reinterpret_cast<bf16x8*>(&As[p][sr + 64 * h][swz(sr + 64 * h, sk >> 3)]) = ra[p][h];
        *reinterpret_cast<bf16x8*>(&Bs[p][sr + 64 * h][swz(sr + 64 * h, sk >> 3)]) = rb[p][h];
      }
  };

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // per slice: the next slice's global loads go out first, the fragments are
  // read from LDS and multiplied, then (after a barrier: every wave holds its
  // fragments) the next slice is stored over the current one
  const int nk = Kp / BK;
  load(0);
  store();
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load((kt + 1) * BK);
    bf16x8 a[NP][4], b[NP][4];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = wm * 64 + i * 16 + li;
        a[p][i] = *reinterpret_cast<const bf16x8*>(&As[p][r][swz(r, lk >> 3)]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = wn * 64 + j * 16 + li;
        b[p][j] = *reinterpret_cast<const bf16x8*>(&Bs[p][r][swz(r, lk >> 3)]);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v4f c = acc[i][j];
        if constexpr (NP == 3) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[1][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2][j], c, 0, 0, 0);
        }
        if constexpr (NP >= 2) {
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[0][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[1][j], c, 0, 0, 0);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0][j], c, 0, 0, 0);
        acc[i][j] = c;
      }
    }
    if (kt + 1 < nk) {
      __syncthreads();
      store();
    }
    __syncthreads();
  }

  float* C = Cb + bz * sC;
  const int rq = (lane >> 4) * 4;
  const int* ea = eA + bz * (Mp / SPAN);
  const int* eb = eB + bz * (Np / SPAN);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t rbase = m0 + wm * 64 + i * 16;
    const int erow = ea[rbase / SPAN];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = n0 + wn * 64 + j * 16 + li;
      const int es = erow + eb[col / SPAN];
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = rbase + rq + v;
        if (row < M) {
          float* cp = C + (maps.c ? (int64_t)maps.c[bz * maps.stride + row] : row) * ldc + col;
          *cp = *cp - ldexpf(acc[i][j][v], es);
        }
      }
    }
  }
}

template <int NP>
int bf16_update_np(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, int64_t sA, const float* B,
                   int64_t ldb, int64_t sB, float* C, int64_t ldc, int64_t sC, int64_t batch, hipStream_t s,
                   const RowMaps& maps) {
  const int64_t Mp = ceil_div(M, TM) * TM, Np = ceil_div(N, TN) * TN;
  const int Kp = (int)(ceil_div(K, BK) * BK);
  const size_t ea_bytes = ((size_t)batch * (Mp / SPAN) * sizeof(int) + 255) & ~(size_t)255;
  const size_t eb_bytes = ((size_t)batch * (Np / SPAN) * sizeof(int) + 255) & ~(size_t)255;
  char *wa = nullptr, *wb = nullptr;
  NPPH_TRY(workspace(kWsSplitA, ea_bytes + (size_t)batch * NP * Mp * Kp * sizeof(__bf16), (void**)&wa));
  NPPH_TRY(workspace(kWsSplitB, eb_bytes + (size_t)batch * NP * Np * Kp * sizeof(__bf16), (void**)&wb));
  int* eA = reinterpret_cast<int*>(wa);
  int* eB = reinterpret_cast<int*>(wb);
  __bf16* PA = reinterpret_cast<__bf16*>(wa + ea_bytes);
  __bf16* PB = reinterpret_cast<__bf16*>(wb + eb_bytes);
  hipLaunchKernelGGL(split_rows_kernel<NP>, dim3((unsigned)(Mp / SPAN), (unsigned)batch), dim3(NTH), 0, s, M,
                     (int)K, Kp, A, lda, sA, PA, Mp, eA, maps);
  hipLaunchKernelGGL(split_cols_kernel<NP>, dim3((unsigned)(Np / SPAN), (unsigned)batch), dim3(NTH), 0, s, N,
                     (int)K, Kp, B, ldb, sB, PB, Np, eB, maps);
  const int tiles_n = (int)(Np / TN);
  hipLaunchKernelGGL(bf16p_gemm_kernel<NP>, dim3((unsigned)((Mp / TM) * tiles_n * batch)), dim3(NTH), 0, s, M, N, Kp,
                     PA, Mp, eA, PB, Np, eB, C, ldc, sC, (int)(Mp / TM), tiles_n, maps);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace

int gemm_bf16split_update(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, int64_t sA,
                          const float* B, int64_t ldb, int64_t sB, float* C, int64_t ldc, int64_t sC,
                          int64_t batch, int terms, hipStream_t s, const RowMaps& maps) {
  if (M <= 0 || N <= 0 || batch <= 0) return NPPH_OK;
  if (K <= 0 || K > (1 << 20)) { set_error("bf16 split update: bad K=%ld", (long)K); return NPPH_ERR_INVALID; }
  if (batch > 65535) { set_error("bf16 split update: batch too large"); return NPPH_ERR_INVALID; }
  if (ceil_div(M, TM) * ceil_div(N, TN) * batch > (1LL << 31) - 1) {
    set_error("bf16 split update: too many tiles");
    return NPPH_ERR_INVALID;
  }
  ProfScope ps(kProfGemmBf16, 2.0 * M * N * K * batch, 4.0 * batch * ((double)M * K + (double)K * N + 2.0 * M * N), s);
  if (terms == 1) return bf16_update_np<1>(M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, maps);
  if (terms == 3) return bf16_update_np<2>(M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, maps);
  if (terms == 6) return bf16_update_np<3>(M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, s, maps);
  set_error("bf16 split update: terms must be 1, 3 or 6 (got %d)", terms);
  return NPPH_ERR_INVALID;
}

}  // namespace npph

// Device-resident entry for tests: C -= A * B through the split update.
extern "C" int npph_bf16split_update_dev(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                                         const float* B, int64_t ldb, float* C, int64_t ldc, int terms,
                                         void* stream) {
  NPPH_TRY(npph::ensure_init());
  const hipStream_t s = stream ? (hipStream_t)stream : npph::default_stream();
  npph::CallScope cs(s);
  return npph::gemm_bf16split_update(M, N, K, A, lda, 0, B, ldb, 0, C, ldc, 0, 1, terms, s);
}
