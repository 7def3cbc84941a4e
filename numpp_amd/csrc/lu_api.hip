// C-ABI entry points of the plain FP64 stacked solve (getrf / getrs / gesv).
//
// Size dispatch: n <= 64 runs the one-wave-per-matrix kernel (factor and
// solve in one launch, matrix in registers); larger n runs the blocked
// right-looking LU with MFMA trailing updates and the blocked getrs.
#include <mutex>
#include <vector>

#include "common.hpp"
#include "lu.hpp"

namespace npph {

int gesv_dev_impl(int64_t count, int n, int nrhs, double* A, int64_t strideA, double* B,
                  int64_t strideB, int32_t* ipiv, int32_t* info, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  int32_t* ip = ipiv;
  int32_t* inf = info;
  if (!ip) NPPH_TRY(workspace(kWsPiv, sizeof(int32_t) * (size_t)count * n, (void**)&ip));
  if (!inf) NPPH_TRY(workspace(kWsInfo, sizeof(int32_t) * (size_t)count, (void**)&inf));
  if (n <= 64) {
    UniformBatch u{count, n, A, strideA, A, ip, inf, nrhs, B, strideB, nrhs > 0 ? B : nullptr};
    return small_lu_launch(u, s);
  }
  int32_t* perm;
  NPPH_TRY(workspace(kWsPerm, sizeof(int32_t) * (size_t)count * n, (void**)&perm));
  NPPH_TRY(blocked_getrf_perm<double>(count, n, A, strideA, ip, perm, inf, 0, nullptr, s));
  // the factorisation leaves logical row p in physical row perm[p]: solve
  // through the map, then return the factors in row order
  if (nrhs <= 0) return unscramble_rows<double>(count, n, A, strideA, perm, s);
  double* X;
  const int64_t sx = (int64_t)n * nrhs;
  NPPH_TRY(workspace(kWsX, sizeof(double) * (size_t)(count * sx), (void**)&X));
  NPPH_TRY(blocked_getrs_perm<double>(count, n, nrhs, A, strideA, perm, B, strideB, X, sx, inf, nullptr, s, perm));
  NPPH_TRY(unscramble_rows<double>(count, n, A, strideA, perm, s));
  // Singular systems (info != 0) are skipped by getrs; their B content on
  // return is unspecified, as the reference raises instead of returning one.
  if (strideB == sx) {
    NPPH_HIP(hipMemcpyAsync(B, X, sizeof(double) * sx * count, hipMemcpyDeviceToDevice, s));
  } else {
    for (int64_t i = 0; i < count; ++i)
      NPPH_HIP(hipMemcpyAsync(B + i * strideB, X + i * sx, sizeof(double) * sx, hipMemcpyDeviceToDevice, s));
  }
  return NPPH_OK;
}

// Scratch bytes gesv_dev_impl asks for (each request rounded up to 256 bytes,
// as a ScratchArena carves them): pivots and info when the caller passes none,
// and for n > 64 the permutation, the factorisation's workspaces and getrs'
// output staging.
size_t gesv_scratch_bytes(int64_t count, int n, int nrhs, bool has_ipiv, bool has_info) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t t = 0;
  if (!has_ipiv) t += al(sizeof(int32_t) * (size_t)count * n);
  if (!has_info) t += al(sizeof(int32_t) * (size_t)count);
  if (n <= 64) return t;
  t += al(sizeof(int32_t) * (size_t)count * n);
  t += blocked_getrf_scratch_bytes(count, n, sizeof(double));
  if (nrhs > 0) t += al(sizeof(double) * (size_t)count * n * nrhs);
  return t;
}

namespace {

// Device-resident scratch of the allocation-free entry point: one copy per GPU,
// held until the process ends. Host code never uses the array's name as an
// address: each device's copy is looked up once with hipGetSymbolAddress and
// that pointer is what kernels receive.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_ptr(void** out) {
  int dev = 0;
  NPPH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) { set_error("no-alloc gesv: device %d out of range", dev); return NPPH_ERR_INVALID; }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    NPPH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NPPH_OK;
}

}  // namespace

}  // namespace npph

using namespace npph;

static hipStream_t pick(void* st) { return st ? (hipStream_t)st : default_stream(); }

extern "C" {

int npph_getrf_batched_dev(int64_t count, int64_t n, double* A, int64_t strideA, int32_t* ipiv,
                           int32_t* info, void* stream) {
  NPPH_TRY(ensure_init());
  if (count < 0 || n < 0 || n > max_blocked_n() || strideA < n * n) {
    set_error("getrf: bad sizes (count=%ld n=%ld stride=%ld)", (long)count, (long)n, (long)strideA);
    return NPPH_ERR_INVALID;
  }
  const hipStream_t s = pick(stream);
  CallScope cs(s);
  return gesv_dev_impl(count, (int)n, 0, A, strideA, nullptr, 0, ipiv, info, s);
}

int npph_getrs_batched_dev(int64_t count, int64_t n, int64_t nrhs, const double* LU, int64_t strideA,
                           const int32_t* ipiv, double* B, int64_t strideB, void* stream) {
  NPPH_TRY(ensure_init());
  if (count < 0 || n < 0 || nrhs < 0 || strideA < n * n || strideB < n * nrhs || !ipiv) {
    set_error("getrs: bad arguments");
    return NPPH_ERR_INVALID;
  }
  if (count == 0 || n == 0 || nrhs == 0) return NPPH_OK;
  hipStream_t s = pick(stream);
  CallScope cs(s);
  int32_t* perm;
  double* X;
  const int64_t sx = n * nrhs;
  NPPH_TRY(workspace(kWsPerm, sizeof(int32_t) * (size_t)(count * n), (void**)&perm));
  NPPH_TRY(workspace(kWsX, sizeof(double) * (size_t)(count * sx), (void**)&X));
  NPPH_TRY(ipiv_to_perm(count, (int)n, ipiv, perm, s));
  NPPH_TRY(blocked_getrs_perm<double>(count, (int)n, (int)nrhs, LU, strideA, perm, B, strideB, X, sx, nullptr,
                                      nullptr, s));
  if (strideB == sx) {
    NPPH_HIP(hipMemcpyAsync(B, X, sizeof(double) * (size_t)(count * sx), hipMemcpyDeviceToDevice, s));
  } else {
    for (int64_t i = 0; i < count; ++i)
      NPPH_HIP(hipMemcpyAsync(B + i * strideB, X + i * sx, sizeof(double) * sx, hipMemcpyDeviceToDevice, s));
  }
  return NPPH_OK;
}

int npph_gesv_batched_dev(int64_t count, int64_t n, int64_t nrhs, double* A, double* B, int32_t* ipiv,
                          int32_t* info, void* stream) {
  NPPH_TRY(ensure_init());
  if (count < 0 || n < 0 || nrhs < 0 || n > max_blocked_n()) {
    set_error("gesv: bad sizes (count=%ld n=%ld nrhs=%ld)", (long)count, (long)n, (long)nrhs);
    return NPPH_ERR_INVALID;
  }
  const hipStream_t s = pick(stream);
  CallScope cs(s);
  return gesv_dev_impl(count, (int)n, (int)nrhs, A, n * n, B, n * nrhs, ipiv, info, s);
}

int npph_gesv_batched_dev_noalloc(int64_t count, int64_t n, int64_t nrhs, double* A, double* B, int32_t* ipiv,
                                  int32_t* info, void* stream) {
  NPPH_TRY(ensure_init());
  if (count < 0 || n < 0 || nrhs < 0 || n > max_blocked_n()) {
    set_error("no-alloc gesv: bad sizes (count=%ld n=%ld nrhs=%ld)", (long)count, (long)n, (long)nrhs);
    return NPPH_ERR_INVALID;
  }
  if (count == 0 || n == 0) return NPPH_OK;
  const size_t need = gesv_scratch_bytes(count, (int)n, (int)nrhs, ipiv != nullptr, info != nullptr);
  if (need > kScratchBytes) {
    set_error("no-alloc gesv: needs %zu bytes of scratch for count=%ld n=%ld nrhs=%ld, above the %zu-byte maximum",
              need, (long)count, (long)n, (long)nrhs, kScratchBytes);
    return NPPH_ERR_INVALID;
  }
  void* base;
  NPPH_TRY(scratch_ptr(&base));
  ScratchArena arena;
  arena.base = static_cast<char*>(base);
  arena.bytes = kScratchBytes;
  arena_bind(&arena);
  // with the arena bound, workspace() carves from it and large
  // factorisations are issued as plain launches (no graph instantiation)
  const int st = gesv_dev_impl(count, (int)n, (int)nrhs, A, n * n, B, n * nrhs, ipiv, info, pick(stream));
  arena_bind(nullptr);
  return st;
}

int npph_gesv_batched(int dtype, int64_t count, int64_t n, int64_t nrhs, void* A, void* B, int32_t* ipiv,
                      int32_t* info) {
  NPPH_TRY(ensure_init());
  if (dtype != NPPH_DT_F64 && dtype != NPPH_DT_F32) {
    set_error("gesv_batched: unsupported dtype %d", dtype);
    return NPPH_ERR_UNSUPPORTED;
  }
  if (count < 0 || n < 0 || nrhs < 0 || n > max_blocked_n()) {
    set_error("gesv_batched: bad sizes (count=%ld n=%ld nrhs=%ld)", (long)count, (long)n, (long)nrhs);
    return NPPH_ERR_INVALID;
  }
  if (count == 0 || n == 0) return NPPH_OK;
  hipStream_t s = default_stream();
  const size_t na = (size_t)(count * n * n), nbv = (size_t)(count * n * nrhs);
  // NumPP computes in float64 whatever the input float dtype (linalg.cpp:20-26).
  std::vector<double> ha, hb;
  const double* srcA = static_cast<const double*>(A);
  const double* srcB = static_cast<const double*>(B);
  if (dtype == NPPH_DT_F32) {
    ha.resize(na);
    hb.resize(nbv);
    const float* fa = static_cast<const float*>(A);
    const float* fb = static_cast<const float*>(B);
    for (size_t i = 0; i < na; ++i) ha[i] = fa[i];
    for (size_t i = 0; i < nbv; ++i) hb[i] = fb[i];
    srcA = ha.data();
    srcB = hb.data();
  }
  // held through the final synchronisation (see npph_gemm)
  CallScope cs(s);
  double *dA, *dB;
  int32_t *dp, *di;
  NPPH_TRY(workspace(kWsA, sizeof(double) * na, (void**)&dA));
  NPPH_TRY(workspace(kWsB, sizeof(double) * (nbv ? nbv : 1), (void**)&dB));
  NPPH_TRY(workspace(kWsPiv, sizeof(int32_t) * (size_t)(count * n), (void**)&dp));
  NPPH_TRY(workspace(kWsInfo, sizeof(int32_t) * (size_t)count, (void**)&di));
  NPPH_HIP(hipMemcpyAsync(dA, srcA, sizeof(double) * na, hipMemcpyHostToDevice, s));
  if (nbv) NPPH_HIP(hipMemcpyAsync(dB, srcB, sizeof(double) * nbv, hipMemcpyHostToDevice, s));
  NPPH_TRY(gesv_dev_impl(count, (int)n, (int)nrhs, dA, n * n, dB, n * nrhs, dp, di, s));
  if (dtype == NPPH_DT_F64) {
    NPPH_HIP(hipMemcpyAsync(A, dA, sizeof(double) * na, hipMemcpyDeviceToHost, s));
    if (nbv) NPPH_HIP(hipMemcpyAsync(B, dB, sizeof(double) * nbv, hipMemcpyDeviceToHost, s));
  } else {
    NPPH_HIP(hipMemcpyAsync(ha.data(), dA, sizeof(double) * na, hipMemcpyDeviceToHost, s));
    if (nbv) NPPH_HIP(hipMemcpyAsync(hb.data(), dB, sizeof(double) * nbv, hipMemcpyDeviceToHost, s));
  }
  if (ipiv) NPPH_HIP(hipMemcpyAsync(ipiv, dp, sizeof(int32_t) * (size_t)(count * n), hipMemcpyDeviceToHost, s));
  if (info) NPPH_HIP(hipMemcpyAsync(info, di, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  if (dtype == NPPH_DT_F32) {
    float* fa = static_cast<float*>(A);
    float* fb = static_cast<float*>(B);
    for (size_t i = 0; i < na; ++i) fa[i] = (float)ha[i];
    for (size_t i = 0; i < nbv; ++i) fb[i] = (float)hb[i];
  }
  return NPPH_OK;
}

}  // extern "C"
