// NumPP GpuVTable provider for the MI355X HIP backend (C++20, host only).
//
// The drop-in for src/backend/cuda_backend.cu:288-320 (the CUDA provider's
// slot functions and accessor): every slot forwards to the C ABI in
// numpp_hip.h and returns false on any failure so NumPP's dispatcher falls back
// to its CPU kernel (gpu_vtable.hpp:14-16, backend.cpp:148-152). Slots never
// throw. The ScyPP slots (csr_spmv, pairwise_sqdist, separable_corr1d) are
// null, which the reference dispatcher already treats as "no device path".
#include <cstring>
#include <vector>

#include "numpp_hip.h"
#include "numpp_hip/gpu_vtable.hpp"
#include "numpp_hip/lapack_vtable.hpp"

namespace numpp {
namespace {

bool ew_binary_slot(int op, DTypeId dt, int64_t n, const void* a, const void* b, void* out) {
  return npph_elementwise_binary(op, static_cast<int>(dt), n, a, b, out) == NPPH_OK;
}

bool ew_unary_slot(int op, DTypeId dt, int64_t n, const void* a, void* out) {
  return npph_elementwise_unary(op, static_cast<int>(dt), n, a, out) == NPPH_OK;
}

bool reduce_slot(int op, DTypeId dt, int64_t n, const void* a, void* out) {
  return npph_reduce(op, static_cast<int>(dt), n, a, out) == NPPH_OK;
}

bool gemm_slot(DTypeId dt, int64_t m, int64_t n, int64_t k, const void* A, const void* B, void* C) {
  return npph_gemm(static_cast<int>(dt), m, n, k, A, B, C) == NPPH_OK;
}

bool gesv_batched_slot(DTypeId dt, int64_t count, int64_t n, int64_t nrhs, void* A, void* B, int32_t* ipiv,
                       int32_t* info) {
  return npph_gesv_batched(static_cast<int>(dt), count, n, nrhs, A, B, ipiv, info) == NPPH_OK;
}

// RAII device buffer for the host-pointer refined slot.
struct DevBuf {
  void* p = nullptr;
  bool ok = true;
  explicit DevBuf(uint64_t bytes) { ok = npph_malloc(&p, bytes ? bytes : 1) == NPPH_OK; }
  ~DevBuf() { if (p) npph_free(p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

bool gesv_refined_slot(int precision, DTypeId dt, int64_t count, int64_t n, int64_t nrhs, const void* A,
                       const void* B, void* X, double* berr, int32_t* iters, int32_t* info) {
  if (dt != DTypeId::Float64 || count < 0 || n < 0 || nrhs < 0) return false;
  if (count == 0 || n == 0 || nrhs == 0) return true;
  const uint64_t ab = 8ull * count * n * n, bb = 8ull * count * n * nrhs;
  DevBuf dA(ab), dB(bb), dX(bb), dE(8ull * count * nrhs), dI(4ull * count), dF(4ull * count);
  if (!dA.ok || !dB.ok || !dX.ok || !dE.ok || !dI.ok || !dF.ok) return false;
  if (npph_memcpy_htod(dA.p, A, ab, nullptr) != NPPH_OK) return false;
  if (npph_memcpy_htod(dB.p, B, bb, nullptr) != NPPH_OK) return false;
  npph_refine_opts o{precision, 10, 1e-12, 1, 6};
  if (npph_gesv_refined_batched_dev(&o, count, n, nrhs, static_cast<const double*>(dA.p),
                                    static_cast<const double*>(dB.p), static_cast<double*>(dX.p),
                                    static_cast<double*>(dE.p), static_cast<int32_t*>(dI.p),
                                    static_cast<int32_t*>(dF.p), nullptr) != NPPH_OK)
    return false;
  if (npph_memcpy_dtoh(X, dX.p, bb, nullptr) != NPPH_OK) return false;
  if (berr && npph_memcpy_dtoh(berr, dE.p, 8ull * count * nrhs, nullptr) != NPPH_OK) return false;
  if (iters && npph_memcpy_dtoh(iters, dI.p, 4ull * count, nullptr) != NPPH_OK) return false;
  if (info && npph_memcpy_dtoh(info, dF.p, 4ull * count, nullptr) != NPPH_OK) return false;
  return true;
}

bool det_slot(DTypeId dt, int64_t count, int64_t n, const void* A, void* det, int32_t* info) {
  return npph_det_batched(static_cast<int>(dt), count, n, A, det, nullptr, nullptr, info) == NPPH_OK;
}

bool slogdet_slot(DTypeId dt, int64_t count, int64_t n, const void* A, void* sign, void* logabsdet,
                  int32_t* info) {
  return npph_det_batched(static_cast<int>(dt), count, n, A, nullptr, sign, logabsdet, info) == NPPH_OK;
}

bool inv_slot(DTypeId dt, int64_t count, int64_t n, const void* A, void* Ainv, int32_t* info) {
  return npph_inv_batched(static_cast<int>(dt), count, n, A, Ainv, info) == NPPH_OK;
}

const GpuVTable g_vtable{
    "hip(gfx950)", &ew_binary_slot, &ew_unary_slot, &reduce_slot, &gemm_slot, nullptr, nullptr, nullptr,
    &gesv_batched_slot, &gesv_refined_slot, &det_slot, &slogdet_slot, &inv_slot,
};

bool lapack_gesv_slot(int n, int nrhs, DTypeId dt, void* A, void* B) {
  return npph_lapack_gesv(n, nrhs, static_cast<int>(dt), A, B) == NPPH_OK;
}

bool lapack_inv_slot(int n, DTypeId dt, const void* A, void* Ainv) {
  return npph_lapack_inv(n, static_cast<int>(dt), A, Ainv) == NPPH_OK;
}

const LapackVTable g_lapack{&lapack_gesv_slot, &lapack_inv_slot, nullptr, nullptr, nullptr, nullptr};

}  // namespace

// The LAPACK provider of a NumPP build that selects this backend for
// lapack_vtable() (lapack_vtable.hpp:30); null without a usable device.
const LapackVTable* lapack_vtable() { return npph_available() ? &g_lapack : nullptr; }

// Probed once per process (npph_init uses std::call_once); null without a
// usable gfx950 device, the reference's "no device" contract (gpu_vtable.hpp:56).
const GpuVTable* gpu_vtable() { return npph_available() ? &g_vtable : nullptr; }

}  // namespace numpp

extern "C" const void* npph_gpu_vtable(void) { return numpp::gpu_vtable(); }
extern "C" const void* npph_lapack_vtable(void) { return numpp::lapack_vtable(); }
