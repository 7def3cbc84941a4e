// Split-BF16 trailing update for the low-precision LU: C -= A * B with FP32
// operands and FP32 accumulation on the gfx950 BF16 MFMA
// (v_mfma_f32_16x16x32_bf16: A[i][k] with i = lane%16, k = 8*(lane/16)+0..7;
// B[k][j] with j = lane%16, same k; four FP32 outputs C[4*(lane/16)+v][lane%16]).
//
// Each FP32 operand x is first scaled by a power of two chosen per workgroup
// tile panel (the 128 x 64 block of L21 and the 64 x 128 block of U12 it reads),
// so the largest magnitude in the panel lands in [1, 2); then split into BF16
// pieces x = h + m + l (h = bf16(x), m = bf16(x - h), l = bf16(x - h - m)).
// terms = 1 keeps only h (plain BF16); 3 uses h,l-two-piece products hh, hl,
// lh (about 16 significant bits); 6 uses three pieces and the products hh, hm,
// mh, hl, mm, lh (about FP32 accuracy). Products are accumulated smallest
// first into one FP32 accumulator, then unscaled in the epilogue.
//
// Workgroup: 256 threads (2 x 2 waves), tile 128 x 128. K is processed in
// chunks of 64 (zero padded to a multiple of 32); each chunk gets its own
// power-of-two scales and its partial products are unscaled into the FP32
// accumulator.
#include "common.hpp"
#include "gemm.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int TM = 128, TN = 128, KMAX = 64, KP = KMAX + 8;  // KP: padded LDS row
constexpr int NTH = 256;

template <int NP>
__global__ __launch_bounds__(NTH) void bf16split_kernel(int64_t M, int64_t N, int K, const float* Ab,
                                                         int64_t lda, int64_t sA, const float* Bb,
                                                         int64_t ldb, int64_t sB, float* Cb, int64_t ldc,
                                                         int64_t sC, int tiles_n) {
  __shared__ __bf16 As[NP][TM][KP];
  __shared__ __bf16 Bs[NP][TN][KP];
  __shared__ float red[2][NTH / 64];
  const int64_t bz = blockIdx.y;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x - tm * tiles_n;
  const float* __restrict__ A = Ab + bz * sA;
  const float* __restrict__ B = Bb + bz * sB;
  float* __restrict__ C = Cb + bz * sC;
  const int64_t m0 = (int64_t)tm * TM, n0 = (int64_t)tn * TN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  // Staging: thread t owns A row (t % 128) and B column (t % 128) of the
  // chunk, k range [32 * (t / 128), +32).
  const int r = t & 127, kh = (t >> 7) * 32;
  const int64_t gr = m0 + r, gc = n0 + r;

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int kc = 0; kc < K; kc += KMAX) {
    const int kn = min(KMAX, K - kc);
    float av[32], bv[32];
    float amax = 0.f, bmax = 0.f;
#pragma unroll
    for (int q = 0; q < 32; ++q) {
      const int k = kh + q;
      av[q] = (gr < M && k < kn) ? A[gr * lda + kc + k] : 0.f;
      bv[q] = (gc < N && k < kn) ? B[(int64_t)(kc + k) * ldb + gc] : 0.f;
      amax = fmaxf(amax, fabsf(av[q]));
      bmax = fmaxf(bmax, fabsf(bv[q]));
    }
#pragma unroll
    for (int msk = 1; msk < 64; msk <<= 1) {
      amax = fmaxf(amax, __shfl_xor(amax, msk, 64));
      bmax = fmaxf(bmax, __shfl_xor(bmax, msk, 64));
    }
    if (kc > 0) __syncthreads();  // previous chunk's LDS reads are done
    if (lane == 0) { red[0][wave] = amax; red[1][wave] = bmax; }
    __syncthreads();
    amax = fmaxf(fmaxf(red[0][0], red[0][1]), fmaxf(red[0][2], red[0][3]));
    bmax = fmaxf(fmaxf(red[1][0], red[1][1]), fmaxf(red[1][2], red[1][3]));
    // Power-of-two scales putting each chunk panel's largest magnitude in [1, 2).
    // (a zero, Inf or NaN panel is not scaled, so Inf and NaN reach C as they
    // are; exponents are clamped to [-126, 127] so that both scales are finite)
    const int ea = (amax > 0.f && isfinite(amax)) ? max(ilogbf(amax), -126) : 0;
    const int eb = (bmax > 0.f && isfinite(bmax)) ? max(ilogbf(bmax), -126) : 0;
    const float sa = ldexpf(1.f, -ea), sb = ldexpf(1.f, -eb);
    const float unscale = ldexpf(1.f, ea + eb);
#pragma unroll
    for (int q = 0; q < 32; ++q) {
      float x = av[q] * sa;
      __bf16 h = (__bf16)x;
      As[0][r][kh + q] = h;
      if constexpr (NP >= 2) {
        float r1 = x - (float)h;
        __bf16 mid = (__bf16)r1;
        As[1][r][kh + q] = mid;
        if constexpr (NP >= 3) As[2][r][kh + q] = (__bf16)(r1 - (float)mid);
      }
      float y = bv[q] * sb;
      __bf16 hb = (__bf16)y;
      Bs[0][r][kh + q] = hb;
      if constexpr (NP >= 2) {
        float r1 = y - (float)hb;
        __bf16 mid = (__bf16)r1;
        Bs[1][r][kh + q] = mid;
        if constexpr (NP >= 3) Bs[2][r][kh + q] = (__bf16)(r1 - (float)mid);
      }
    }
    __syncthreads();
    v4f part[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) part[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    const int kp = (kn + 31) & ~31;
    for (int k0 = 0; k0 < kp; k0 += 32) {
      v8bf a[NP][4], b[NP][4];
#pragma unroll
      for (int pc = 0; pc < NP; ++pc) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          a[pc][i] = *reinterpret_cast<const v8bf*>(&As[pc][wm * 64 + i * 16 + li][k0 + lk]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          b[pc][j] = *reinterpret_cast<const v8bf*>(&Bs[pc][wn * 64 + j * 16 + li][k0 + lk]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v4f c = part[i][j];
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
          part[i][j] = c;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] += part[i][j] * unscale;
  }

  const int rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = m0 + wm * 64 + i * 16 + rq + v;
      if (row >= M) continue;
      float* crow = C + row * ldc;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t col = n0 + wn * 64 + j * 16 + li;
        if (col < N) crow[col] = crow[col] - acc[i][j][v];
      }
    }
  }
}

}  // namespace

int gemm_bf16split_update(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda, int64_t sA,
                          const float* B, int64_t ldb, int64_t sB, float* C, int64_t ldc, int64_t sC,
                          int64_t batch, int terms, hipStream_t s) {
  if (M <= 0 || N <= 0 || batch <= 0) return NPPH_OK;
  if (K <= 0 || K > (1 << 20)) { set_error("bf16 split update: bad K=%ld", (long)K); return NPPH_ERR_INVALID; }
  if (batch > 65535) { set_error("bf16 split update: batch too large"); return NPPH_ERR_INVALID; }
  const int tiles_m = (int)ceil_div(M, TM), tiles_n = (int)ceil_div(N, TN);
  dim3 grid((unsigned)(tiles_m * tiles_n), (unsigned)batch);
  ProfScope ps(kProfGemmBf16, 2.0 * M * N * K * batch, 4.0 * batch * ((double)M * K + (double)K * N + 2.0 * M * N), s);
  if (terms == 1)
    hipLaunchKernelGGL(bf16split_kernel<1>, grid, dim3(NTH), 0, s, M, N, (int)K, A, lda, sA, B, ldb, sB, C, ldc, sC, tiles_n);
  else if (terms == 3)
    hipLaunchKernelGGL(bf16split_kernel<2>, grid, dim3(NTH), 0, s, M, N, (int)K, A, lda, sA, B, ldb, sB, C, ldc, sC, tiles_n);
  else if (terms == 6)
    hipLaunchKernelGGL(bf16split_kernel<3>, grid, dim3(NTH), 0, s, M, N, (int)K, A, lda, sA, B, ldb, sB, C, ldc, sC, tiles_n);
  else { set_error("bf16 split update: terms must be 1, 3 or 6 (got %d)", terms); return NPPH_ERR_INVALID; }
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// Device-resident entry for tests: C -= A * B through the split kernel.
}  // namespace npph

extern "C" int npph_bf16split_update_dev(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                                         const float* B, int64_t ldb, float* C, int64_t ldc, int terms,
                                         void* stream) {
  NPPH_TRY(npph::ensure_init());
  const hipStream_t s = stream ? (hipStream_t)stream : npph::default_stream();
  npph::CallScope cs(s);
  return npph::gemm_bf16split_update(M, N, K, A, lda, 0, B, ldb, 0, C, ldc, 0, 1, terms, s);
}
