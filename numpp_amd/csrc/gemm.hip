// Row-major batched GEMM on gfx950 MFMA: C = alpha * A * B + beta * C.
//
// FP64 uses v_mfma_f64_16x16x4f64, FP32 uses v_mfma_f32_16x16x4f32 (Mfma<T>,
// with its operand and result layouts, in common.hpp).
//
// Workgroup: 256 threads = 4 waves in a 2 x 2 grid; macro tile BM x BN = 128 x 128,
// each wave owns a 64 x 64 sub-tile (4 x 4 MFMA tiles, 16 accumulators).
// K is staged in BK = 16 slices through double-buffered LDS: A is stored
// k-major (transposed on the LDS write) so that both operands are read as 16
// consecutive elements per k row. The rows are padded to BM + 16 elements, so
// the four k rows one MFMA operand read touches fall in disjoint banks, and the
// loads are laid out so that the LDS writes are bank-free too: consecutive
// lanes take consecutive A rows (one element each per transposed write) and
// consecutive 16-byte B chunks. Global loads for slice t+1 are issued before
// the MFMAs of slice t so their latency hides behind the matrix core.
//
// Tile order: blockIdx.x is remapped so that consecutive tiles land on the same
// XCD (workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2), then rastered in groups of 8 tile rows so an XCD's tiles share A row
// panels and B column panels in its L2.
#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int BK = 16, NT = 256;
constexpr int PAD = 16;

__device__ __forceinline__ void tile_coords(int bid, int tiles_m, int tiles_n, int& tm, int& tn) {
  const int total = tiles_m * tiles_n;
  int idx = bid;
  if ((total & 7) == 0) idx = (bid >> 3) + (bid & 7) * (total >> 3);
  constexpr int G = 8;
  const int per_group = G * tiles_n;
  const int g = idx / per_group;
  const int first_m = g * G;
  const int gsize = min(tiles_m - first_m, G);
  const int r = idx - g * per_group;
  tm = first_m + r % gsize;
  tn = r / gsize;
}

// Tile BM x BN with BM, BN in {128, 64}: the waves form a 2 x 2 grid of
// (BM/2) x (BN/2) sub-tiles (64-wide tiles serve narrow or short products, where
// a 128-wide tile would leave MFMA rows or columns empty).
// MAP: rows reached through p.maps (the recursive LU's unmoved rows); without
// it the kernel is the plain strided GEMM and keeps its registers for tiles.
template <class T, int BM, int BN, bool MAP>
__global__ __launch_bounds__(NT, 2) void gemm_kernel(GemmArgs<T> p) {
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int IT = BM / 32;      // 16-row MFMA tiles per wave
  constexpr int JT = BN / 32;      // 16-column MFMA tiles per wave
  constexpr int AE = BM / 16;      // A elements per thread per slice
  constexpr int BE = BN / 16;      // B elements per thread per slice
  __shared__ T As[2][BK][BM + PAD];
  __shared__ T Bs[2][BK][BN + PAD];

  int tm, tn;
  tile_coords(blockIdx.x, p.tiles_m, p.tiles_n, tm, tn);
  const int64_t bz = blockIdx.y;
  const T* __restrict__ A = p.A + bz * p.sA;
  const T* __restrict__ B = p.B + bz * p.sB;
  T* __restrict__ C = p.C + bz * p.sC;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  // Global-load mapping. A tile (BM rows x BK): AE consecutive k of one row per
  // thread, consecutive lanes on consecutive rows. B tile (BK rows x BN): 16
  // threads per row, each taking BE / VW chunks of VW elements 16 chunks apart.
  constexpr int VW = 16 / sizeof(T);  // elements per 16-byte vector
  constexpr int BV = BE / VW;         // B vectors per thread per slice
  static_assert((BK / AE) * BM == NT && BK * 16 == NT && BE % VW == 0, "tile mapping");
  const int a_row = tid % BM, a_k = (tid / BM) * AE;
  const int b_k = tid >> 4, b_c = tid & 15;
  auto bcol = [&](int i) { return (b_c + 16 * (i / VW)) * VW + i % VW; };
  const bool mn_full = (m0 + BM <= p.M) && (n0 + BN <= p.N);
  // physical rows (row maps): this thread's A row, and B rows per K slice
  const int32_t* mapb = MAP && p.maps.b ? p.maps.b + bz * p.maps.stride : nullptr;
  const int64_t a_phys =
      MAP && p.maps.a ? (int64_t)p.maps.a[bz * p.maps.stride + min(m0 + a_row, p.M - 1)] : m0 + a_row;
  const int lk = lane >> 4, lr = lane & 15;
  // the epilogue's C rows, fetched now with the other map entries so that the
  // kernel pays one map latency, not one at each end
  int32_t c_phys[MAP ? IT : 1][4];
  if constexpr (MAP) {
#pragma unroll
    for (int i = 0; i < IT; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t row = min(m0 + wm * (BM / 2) + i * 16 + Mfma<T>::row(lk, v), p.M - 1);
        c_phys[i][v] = p.maps.c ? p.maps.c[bz * p.maps.stride + row] : (int32_t)row;
      }
  }

  // B's physical row for the slice starting at k0 (a map entry is fetched two
  // slices ahead of its use, so its latency is not added to the tile loads)
  auto brow = [&](int64_t k0) -> int64_t {
    if constexpr (MAP) return mapb ? mapb[min(k0 + b_k, p.K - 1)] : k0 + b_k;
    return k0 + b_k;
  };
  T ra[AE], rb[BE];
  auto load = [&](int64_t k0, int64_t b_phys) {
    const bool full = mn_full && (k0 + BK <= p.K) && p.vec;
    const int64_t gr = m0 + a_row, gk = k0 + a_k;
    const int64_t bk = k0 + b_k;
    if (full) {
      const T* pa = A + a_phys * p.lda + gk;
      const T* pb = B + b_phys * p.ldb + n0 + b_c * VW;
      if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int i = 0; i < AE; i += 2) {
          double2 va = *reinterpret_cast<const double2*>(pa + i);
          ra[i] = va.x; ra[i + 1] = va.y;
        }
#pragma unroll
        for (int q = 0; q < BV; ++q) {
          double2 vb = *reinterpret_cast<const double2*>(pb + 16 * VW * q);
          rb[2 * q] = vb.x; rb[2 * q + 1] = vb.y;
        }
      } else {
#pragma unroll
        for (int i = 0; i < AE; i += 4) {
          float4 va = *reinterpret_cast<const float4*>(pa + i);
          ra[i] = va.x; ra[i + 1] = va.y; ra[i + 2] = va.z; ra[i + 3] = va.w;
        }
#pragma unroll
        for (int q = 0; q < BV; ++q) {
          float4 vb = *reinterpret_cast<const float4*>(pb + 16 * VW * q);
          rb[4 * q] = vb.x; rb[4 * q + 1] = vb.y; rb[4 * q + 2] = vb.z; rb[4 * q + 3] = vb.w;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < AE; ++i) ra[i] = (gr < p.M && gk + i < p.K) ? A[a_phys * p.lda + gk + i] : T(0);
#pragma unroll
      for (int i = 0; i < BE; ++i) {
        const int64_t bc = n0 + bcol(i);
        rb[i] = (bk < p.K && bc < p.N) ? B[b_phys * p.ldb + bc] : T(0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < AE; ++i) As[buf][a_k + i][a_row] = ra[i];
#pragma unroll
    for (int i = 0; i < BE; ++i) Bs[buf][b_k][bcol(i)] = rb[i];
  };

  acc_t acc[IT][JT];
#pragma unroll
  for (int i = 0; i < IT; ++i)
#pragma unroll
    for (int j = 0; j < JT; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // The 64 x 64 mapped tile (the recursive LU's short updates, a few K slices
  // each) fetches its C values with the first slice, so their latency overlaps
  // the product instead of following it.
  constexpr bool PC = MAP && BM == 64 && BN == 64;
  const bool beta0 = (p.beta == T(0));
  T cpre[PC ? IT : 1][4][PC ? JT : 1];
  const int kt_total = (int)((p.K + BK - 1) / BK);
  int64_t bnext = brow(BK);
  load(0, brow(0));
  if constexpr (PC) {
    if (!beta0) {
#pragma unroll
      for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t row = m0 + wm * (BM / 2) + i * 16 + Mfma<T>::row(lk, v);
          const T* crow = C + (int64_t)c_phys[i][v] * p.ldc;
#pragma unroll
          for (int j = 0; j < JT; ++j) {
            const int64_t col = n0 + wn * (BN / 2) + j * 16 + lr;
            cpre[i][v][j] = (row < p.M && col < p.N) ? crow[col] : T(0);
          }
        }
    }
  }
  store(0);
  __syncthreads();
  for (int kt = 0; kt < kt_total; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < kt_total) {
      // the map entry for slice kt + 2 is issued before slice kt + 1's tile
      // loads, so waiting for it later never waits for those (vmcnt is in order)
      const int64_t b2 = brow((int64_t)(kt + 2) * BK);
      load((int64_t)(kt + 1) * BK, bnext);
      bnext = b2;
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[IT], bf[JT];
#pragma unroll
      for (int i = 0; i < IT; ++i) af[i] = As[cur][kk + lk][wm * (BM / 2) + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < JT; ++j) bf[j] = Bs[cur][kk + lk][wn * (BN / 2) + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int j = 0; j < JT; ++j) acc[i][j] = Mfma<T>::run(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < kt_total) store(cur ^ 1);
    __syncthreads();
  }

  // Epilogue: C[row][col] = alpha * acc + beta * C.
#pragma unroll
  for (int i = 0; i < IT; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = m0 + wm * (BM / 2) + i * 16 + Mfma<T>::row(lk, v);
      if (row >= p.M) continue;
      T* crow = C + (MAP ? (int64_t)c_phys[MAP ? i : 0][v] : row) * p.ldc;
#pragma unroll
      for (int j = 0; j < JT; ++j) {
        const int64_t col = n0 + wn * (BN / 2) + j * 16 + lr;
        if (col < p.N) {
          T r = p.alpha * acc[i][j][v];
          if (!beta0) r += p.beta * (PC ? cpre[PC ? i : 0][v][PC ? j : 0] : crow[col]);
          crow[col] = r;
        }
      }
    }
  }
}

template <class T>
bool vec_ok(const GemmArgs<T>& p) {
  constexpr int64_t w = 16 / sizeof(T);  // elements per 16-byte load
  auto al = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  return al(p.A) && al(p.B) && p.lda % w == 0 && p.ldb % w == 0 && p.sA % w == 0 && p.sB % w == 0;
}

}  // namespace

template <class T>
int gemm_launch(GemmArgs<T> p, int64_t batch, hipStream_t s, int prof) {
  if (p.M <= 0 || p.N <= 0 || batch <= 0) return NPPH_OK;
  if (p.K <= 0) {
    // C = beta * C handled by a K=1 pass with a zero A column is wasteful; the
    // LU and matmul callers never pass K = 0, so treat it as invalid.
    set_error("gemm: K must be positive");
    return NPPH_ERR_INVALID;
  }
  if (p.M >= (1LL << 31)) {  // row indices are kept in 32 bits
    set_error("gemm: M too large");
    return NPPH_ERR_INVALID;
  }
  if (p.lda < p.K || p.ldb < p.N || p.ldc < p.N) {
    set_error("gemm: leading dimension too small");
    return NPPH_ERR_INVALID;
  }
  // 64-wide tiles for narrow (N <= 64) or short (M <= 64) products (the
  // lowest recursion levels), where a 128-wide tile would leave half of its
  // MFMA columns or rows empty; and for products of at most 2048 x 1024
  // outputs whose batch gives fewer than 16 128 x 128 tiles per CU (the
  // recursion's middle levels), where narrower tiles put more CUs to work
  // (measured)
  bool narrow = p.N <= 64, shortm = p.M <= 64;
  const int64_t cus = std::max(cu_count(), 1);
  auto wgs = [&](int bm, int bn) { return ceil_div(p.M, bm) * ceil_div(p.N, bn) * batch; };
  if (p.M * p.N <= 2048 * 1024 && wgs(shortm ? 64 : 128, narrow ? 64 : 128) < 16 * cus) {
    if (!narrow && p.N >= p.M) narrow = true;
    else if (!shortm) shortm = true;
    if (wgs(shortm ? 64 : 128, narrow ? 64 : 128) < 16 * cus) narrow = shortm = true;
  }
  const int BN = narrow ? 64 : 128, BM = shortm ? 64 : 128;
  p.tiles_m = (int)ceil_div(p.M, BM);
  p.tiles_n = (int)ceil_div(p.N, BN);
  p.vec = vec_ok(p) ? 1 : 0;
  if (batch > 65535) { set_error("gemm: batch too large"); return NPPH_ERR_INVALID; }
  dim3 grid((unsigned)(p.tiles_m * p.tiles_n), (unsigned)batch);
  const double es = sizeof(T), fl = 2.0 * p.M * p.N * p.K * batch;
  const double by = es * batch * ((double)p.M * p.K + (double)p.K * p.N + (p.beta == T(0) ? 1.0 : 2.0) * p.M * p.N);
  const int kind = prof == -1 ? (sizeof(T) == 8 ? kProfGemmF64 : kProfGemmF32) : prof;
  ProfScope ps(kind, fl, by, s);
  const bool mapped = p.maps.a || p.maps.b || p.maps.c;
  if (mapped) {
    if (shortm && narrow) hipLaunchKernelGGL((gemm_kernel<T, 64, 64, true>), grid, dim3(NT), 0, s, p);
    else if (shortm) hipLaunchKernelGGL((gemm_kernel<T, 64, 128, true>), grid, dim3(NT), 0, s, p);
    else if (narrow) hipLaunchKernelGGL((gemm_kernel<T, 128, 64, true>), grid, dim3(NT), 0, s, p);
    else hipLaunchKernelGGL((gemm_kernel<T, 128, 128, true>), grid, dim3(NT), 0, s, p);
  } else {
    if (shortm && narrow) hipLaunchKernelGGL((gemm_kernel<T, 64, 64, false>), grid, dim3(NT), 0, s, p);
    else if (shortm) hipLaunchKernelGGL((gemm_kernel<T, 64, 128, false>), grid, dim3(NT), 0, s, p);
    else if (narrow) hipLaunchKernelGGL((gemm_kernel<T, 128, 64, false>), grid, dim3(NT), 0, s, p);
    else hipLaunchKernelGGL((gemm_kernel<T, 128, 128, false>), grid, dim3(NT), 0, s, p);
  }
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

template int gemm_launch<double>(GemmArgs<double>, int64_t, hipStream_t, int);
template int gemm_launch<float>(GemmArgs<float>, int64_t, hipStream_t, int);

}  // namespace npph

using namespace npph;

extern "C" {

int npph_dgemm_dev(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                   int64_t sA, const double* B, int64_t ldb, int64_t sB, double beta, double* C,
                   int64_t ldc, int64_t sC, int64_t batch, void* stream) {
  NPPH_TRY(ensure_init());
  const hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  CallScope cs(s);
  GemmArgs<double> p{m, n, k, alpha, beta, A, lda, sA, B, ldb, sB, C, ldc, sC, 0, 0, 0};
  return gemm_launch<double>(p, batch, s);
}

int npph_sgemm_dev(int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                   int64_t sA, const float* B, int64_t ldb, int64_t sB, float beta, float* C,
                   int64_t ldc, int64_t sC, int64_t batch, void* stream) {
  NPPH_TRY(ensure_init());
  const hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  CallScope cs(s);
  GemmArgs<float> p{m, n, k, alpha, beta, A, lda, sA, B, ldb, sB, C, ldc, sC, 0, 0, 0};
  return gemm_launch<float>(p, batch, s);
}

int npph_gemm(int dtype, int64_t m, int64_t n, int64_t k, const void* A, const void* B, void* C) {
  NPPH_TRY(ensure_init());
  if (dtype != NPPH_DT_F64 && dtype != NPPH_DT_F32) {
    set_error("npph_gemm: unsupported dtype %d", dtype);
    return NPPH_ERR_UNSUPPORTED;
  }
  if (m < 0 || n < 0 || k < 0) { set_error("npph_gemm: negative size"); return NPPH_ERR_INVALID; }
  const size_t es = dtype == NPPH_DT_F64 ? 8 : 4;
  const size_t ab = (size_t)m * k * es, bb = (size_t)k * n * es, cb = (size_t)m * n * es;
  if (m == 0 || n == 0) return NPPH_OK;
  hipStream_t s = default_stream();
  if (k == 0) {
    std::memset(C, 0, cb);
    return NPPH_OK;
  }
  // the call scope is held through the final synchronisation, so concurrent
  // host-pointer calls take turns on the library stream and its buffers
  CallScope cs(s);
  void *dA, *dB, *dC;
  NPPH_TRY(workspace(kWsA, ab, &dA));
  NPPH_TRY(workspace(kWsB, bb, &dB));
  NPPH_TRY(workspace(kWsX, cb, &dC));
  NPPH_HIP(hipMemcpyAsync(dA, A, ab, hipMemcpyHostToDevice, s));
  NPPH_HIP(hipMemcpyAsync(dB, B, bb, hipMemcpyHostToDevice, s));
  int st;
  if (dtype == NPPH_DT_F64) {
    GemmArgs<double> p{m, n, k, 1.0, 0.0, (const double*)dA, k, 0, (const double*)dB, n, 0, (double*)dC, n, 0, 0, 0, 0};
    st = gemm_launch<double>(p, 1, s);
  } else {
    GemmArgs<float> p{m, n, k, 1.0f, 0.0f, (const float*)dA, k, 0, (const float*)dB, n, 0, (float*)dC, n, 0, 0, 0, 0};
    st = gemm_launch<float>(p, 1, s);
  }
  if (st != NPPH_OK) return st;
  NPPH_HIP(hipMemcpyAsync(C, dC, cb, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}

}  // extern "C"
