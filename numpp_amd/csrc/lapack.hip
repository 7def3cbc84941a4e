// LapackVTable-compatible gesv and inv (include/numpp/backend/lapack_vtable.hpp:
// 14-18 in NumPP) on the device.
//
// NumPP's linalg::solve tries LapackVTable::gesv before its portable kernel
// (linalg.cpp:372-382) and, unlike the portable kernel, computes in the output
// dtype: float32 systems are solved in float32 (LAPACKE_sgesv in the reference
// provider, lapack_backend.cpp:15-25). These entry points keep that contract:
// F64 runs the plain FP64 LU + getrs; F32 factors and solves in FP32 (the same
// recursive LU and getrs kernels in working precision float, NumPP's pivot
// rule). A is overwritten by the packed LU factors and B by X, as LAPACK's gesv
// does; an exact-zero pivot returns NPPH_ERR_SINGULAR (the reference provider
// returns false when LAPACK reports info > 0, and NumPP's portable kernel then
// raises "Singular matrix").
#include <vector>

#include "common.hpp"
#include "lu.hpp"

namespace npph {

int inv_dev_impl(int64_t count, int n, double* A, double* Ainv, int32_t* info, hipStream_t s);

namespace {

__global__ void identity_f32_kernel(int n, float* X) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (int64_t)n * n) X[e] = (e / n == e % n) ? 1.f : 0.f;
}

// FP32 LU + solve of one system in place (A -> LU, B -> X), info on the device.
int sgesv_dev(int n, int nrhs, float* A, float* B, int32_t* info, hipStream_t s) {
  int32_t *ipiv, *perm;
  float* X;
  NPPH_TRY(workspace(kWsPiv, sizeof(int32_t) * (size_t)n, (void**)&ipiv));
  NPPH_TRY(workspace(kWsPerm, sizeof(int32_t) * (size_t)n, (void**)&perm));
  NPPH_TRY(workspace(kWsD, sizeof(float) * (size_t)n * nrhs, (void**)&X));
  NPPH_TRY(blocked_getrf_perm<float>(1, n, A, (int64_t)n * n, ipiv, perm, info, 0, nullptr, s));
  NPPH_TRY(unscramble_rows<float>(1, n, A, (int64_t)n * n, perm, s));
  NPPH_TRY(blocked_getrs_perm<float>(1, n, nrhs, A, (int64_t)n * n, perm, B, (int64_t)n * nrhs, X,
                                     (int64_t)n * nrhs, info, nullptr, s));
  NPPH_HIP(hipMemcpyAsync(B, X, sizeof(float) * (size_t)n * nrhs, hipMemcpyDeviceToDevice, s));
  return NPPH_OK;
}

int check(const char* what, int dtype, int64_t n, int64_t nrhs) {
  if (dtype != NPPH_DT_F64 && dtype != NPPH_DT_F32) {
    set_error("%s: unsupported dtype %d", what, dtype);
    return NPPH_ERR_UNSUPPORTED;
  }
  if (n < 0 || nrhs < 0 || n > max_blocked_n()) {
    set_error("%s: bad sizes (n=%ld nrhs=%ld)", what, (long)n, (long)nrhs);
    return NPPH_ERR_INVALID;
  }
  return NPPH_OK;
}

int singular(const char* what, int32_t info) {
  set_error("%s: singular matrix (exact zero pivot at step %d)", what, info - 1);
  return NPPH_ERR_SINGULAR;
}

}  // namespace
}  // namespace npph

using namespace npph;

extern "C" {

int npph_lapack_gesv(int n, int nrhs, int dtype, void* A, void* B) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check("lapack gesv", dtype, n, nrhs));
  if (n == 0) return NPPH_OK;
  const hipStream_t s = default_stream();
  CallScope cs(s);  // held through the final synchronisation
  const size_t es = dtype == NPPH_DT_F64 ? 8 : 4;
  const size_t na = (size_t)n * n, nb = (size_t)n * nrhs;
  void *dA, *dB;
  int32_t* dinfo;
  NPPH_TRY(workspace(kWsA, es * na, &dA));
  NPPH_TRY(workspace(kWsB, es * (nb ? nb : 1), &dB));
  NPPH_TRY(workspace(kWsFlags, sizeof(int32_t), (void**)&dinfo));
  NPPH_HIP(hipMemcpyAsync(dA, A, es * na, hipMemcpyHostToDevice, s));
  if (nb) NPPH_HIP(hipMemcpyAsync(dB, B, es * nb, hipMemcpyHostToDevice, s));
  if (dtype == NPPH_DT_F64) {
    NPPH_TRY(gesv_dev_impl(1, n, nrhs, (double*)dA, na, (double*)dB, nb, nullptr, dinfo, s));
  } else if (nrhs > 0) {
    NPPH_TRY(sgesv_dev(n, nrhs, (float*)dA, (float*)dB, dinfo, s));
  } else {
    int32_t *ipiv, *perm;
    NPPH_TRY(workspace(kWsPiv, sizeof(int32_t) * (size_t)n, (void**)&ipiv));
    NPPH_TRY(workspace(kWsPerm, sizeof(int32_t) * (size_t)n, (void**)&perm));
    NPPH_TRY(blocked_getrf_perm<float>(1, n, (float*)dA, (int64_t)na, ipiv, perm, dinfo, 0, nullptr, s));
    NPPH_TRY(unscramble_rows<float>(1, n, (float*)dA, (int64_t)na, perm, s));
  }
  int32_t info = 0;
  NPPH_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(A, dA, es * na, hipMemcpyDeviceToHost, s));
  if (nb) NPPH_HIP(hipMemcpyAsync(B, dB, es * nb, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return info != 0 ? singular("lapack gesv", info) : NPPH_OK;
}

int npph_lapack_inv(int n, int dtype, const void* A, void* Ainv) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check("lapack inv", dtype, n, n));
  if (n == 0) return NPPH_OK;
  const hipStream_t s = default_stream();
  CallScope cs(s);
  const size_t es = dtype == NPPH_DT_F64 ? 8 : 4;
  const size_t na = (size_t)n * n;
  void *dA, *dX;
  int32_t* dinfo;
  NPPH_TRY(workspace(kWsA, es * na, &dA));
  NPPH_TRY(workspace(kWsB, es * na, &dX));
  NPPH_TRY(workspace(kWsFlags, sizeof(int32_t), (void**)&dinfo));
  NPPH_HIP(hipMemcpyAsync(dA, A, es * na, hipMemcpyHostToDevice, s));
  if (dtype == NPPH_DT_F64) {
    NPPH_TRY(inv_dev_impl(1, n, (double*)dA, (double*)dX, dinfo, s));
  } else {
    hipLaunchKernelGGL(identity_f32_kernel, dim3((unsigned)ceil_div((int64_t)na, 256)), dim3(256), 0, s, n,
                       (float*)dX);
    NPPH_HIP(hipGetLastError());
    NPPH_TRY(sgesv_dev(n, n, (float*)dA, (float*)dX, dinfo, s));
  }
  int32_t info = 0;
  NPPH_HIP(hipMemcpyAsync(&info, dinfo, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(Ainv, dX, es * na, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return info != 0 ? singular("lapack inv", info) : NPPH_OK;
}

}  // extern "C"
