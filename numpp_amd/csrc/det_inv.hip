// Determinants, signed log-determinants and inverses of a uniform stack from
// one device LU factorisation (the plain FP64 getrf of lu_api.hip).
//
// det      <- linalg.cpp:393-407: 0 when the factorisation meets an exact-zero
//             pivot, else permsign * u_00 * u_11 * ... in row order;
// slogdet  <- linalg.cpp:408-439 (real): (0, -inf) when singular, else
//             logabsdet = sum of log|u_ii| and sign = permsign * prod u_ii/|u_ii|,
//             both in row order;
// inv      <- linalg.cpp:105-115: getrs against the identity (n right-hand-side
//             columns); singular systems report info = k + 1.
// batched.cpp:75-76 and 102-108 map these over a stack one matrix at a time;
// here the whole stack is factored by one device call and the per-matrix
// reductions run one thread per matrix (a sequential loop keeps the
// reference's multiplication and summation order).
#include <cmath>
#include <vector>

#include "common.hpp"
#include "lu.hpp"

namespace npph {

namespace {

// One thread per matrix: permsign from the per-step pivot rows (a swap at step
// k when ipiv[k] != k, linalg.cpp:58-62), then the ordered products / sums.
// An empty matrix (n = 0) is never singular: its info is set to 0 here.
__global__ __launch_bounds__(256) void det_kernel(int64_t count, int n, const double* LU, const int32_t* ipiv,
                                                  int32_t* info, double* det, double* sign, double* logabs) {
  const int64_t mat = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (mat >= count) return;
  const double* a = LU + mat * (int64_t)n * n;
  const int32_t* p = ipiv + mat * (int64_t)n;
  if (n == 0) info[mat] = 0;
  else if (info[mat] != 0) {
    if (det) det[mat] = 0.0;
    if (sign) sign[mat] = 0.0;
    if (logabs) logabs[mat] = -INFINITY;
    return;
  }
  double sgn = 1.0;
  for (int k = 0; k < n; ++k) sgn = p[k] != k ? -sgn : sgn;
  double d = sgn, la = 0.0, sg = sgn;
  for (int i = 0; i < n; ++i) {
    const double u = a[(int64_t)i * n + i];
    const double m = fabs(u);
    d *= u;
    la += log(m);
    sg *= u / m;
  }
  if (det) det[mat] = d;
  if (sign) sign[mat] = sg;
  if (logabs) logabs[mat] = la;
}

__global__ __launch_bounds__(256) void identity_kernel(int64_t count, int n, double* X) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)n * n;
  if (e >= count * per) return;
  const int64_t w = e % per;
  X[e] = (w / n == w % n) ? 1.0 : 0.0;
}

inline dim3 g1(int64_t n) { return dim3((unsigned)ceil_div(n, 256)); }

}  // namespace

// A (count x n x n, FP64, device) is overwritten by its packed LU factors.
int det_dev_impl(int64_t count, int n, double* A, double* det, double* sign, double* logabs, int32_t* info,
                 hipStream_t s) {
  if (count <= 0) return NPPH_OK;
  int32_t *ip, *inf = info;
  NPPH_TRY(workspace(kWsPiv, sizeof(int32_t) * (size_t)count * (n > 0 ? n : 1), (void**)&ip));
  if (!inf) NPPH_TRY(workspace(kWsInfo, sizeof(int32_t) * (size_t)count, (void**)&inf));
  if (n == 0) {
    // the empty product: det 1, sign 1, logabsdet 0, never singular
    hipLaunchKernelGGL(det_kernel, g1(count), dim3(256), 0, s, count, 0, A, ip, inf, det, sign, logabs);
    NPPH_HIP(hipGetLastError());
    return NPPH_OK;
  }
  NPPH_TRY(gesv_dev_impl(count, n, 0, A, (int64_t)n * n, nullptr, 0, ip, inf, s));
  hipLaunchKernelGGL(det_kernel, g1(count), dim3(256), 0, s, count, n, A, ip, inf, det, sign, logabs);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// A is overwritten by its LU factors, Ainv (count x n x n) receives the inverse.
int inv_dev_impl(int64_t count, int n, double* A, double* Ainv, int32_t* info, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  hipLaunchKernelGGL(identity_kernel, g1(count * n * n), dim3(256), 0, s, count, n, Ainv);
  NPPH_HIP(hipGetLastError());
  return gesv_dev_impl(count, n, n, A, (int64_t)n * n, Ainv, (int64_t)n * n, nullptr, info, s);
}

}  // namespace npph

using namespace npph;

namespace {

hipStream_t pick(void* st) { return st ? (hipStream_t)st : default_stream(); }

// Host-pointer calls: inputs to FP64 (the reference computes in float64 and
// returns float32 only for float32 input, linalg.cpp:20-26).
int upload_f64(int dtype, const void* src, size_t n, double* dst, hipStream_t s, std::vector<double>& tmp) {
  if (dtype == NPPH_DT_F64) {
    NPPH_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, s));
    return NPPH_OK;
  }
  tmp.resize(n);
  const float* f = static_cast<const float*>(src);
  for (size_t i = 0; i < n; ++i) tmp[i] = f[i];
  NPPH_HIP(hipMemcpyAsync(dst, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  return NPPH_OK;
}

void store_out(int dtype, const double* src, size_t n, void* dst) {
  if (!dst) return;
  if (dtype == NPPH_DT_F64) {
    std::copy(src, src + n, static_cast<double*>(dst));
  } else {
    float* f = static_cast<float*>(dst);
    for (size_t i = 0; i < n; ++i) f[i] = (float)src[i];
  }
}

int check_args(const char* what, int dtype, int64_t count, int64_t n) {
  if (dtype != NPPH_DT_F64 && dtype != NPPH_DT_F32) {
    set_error("%s: unsupported dtype %d", what, dtype);
    return NPPH_ERR_UNSUPPORTED;
  }
  if (count < 0 || n < 0 || n > max_blocked_n()) {
    set_error("%s: bad sizes (count=%ld n=%ld)", what, (long)count, (long)n);
    return NPPH_ERR_INVALID;
  }
  return NPPH_OK;
}

}  // namespace

extern "C" {

int npph_det_batched_dev(int64_t count, int64_t n, double* A, double* det, double* sign, double* logabsdet,
                         int32_t* info, void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_args("det", NPPH_DT_F64, count, n));
  const hipStream_t s = pick(stream);
  CallScope cs(s);
  return det_dev_impl(count, (int)n, A, det, sign, logabsdet, info, s);
}

int npph_inv_batched_dev(int64_t count, int64_t n, double* A, double* Ainv, int32_t* info, void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_args("inv", NPPH_DT_F64, count, n));
  if (A == Ainv && count > 0 && n > 0) { set_error("inv: A and Ainv must not alias"); return NPPH_ERR_INVALID; }
  const hipStream_t s = pick(stream);
  CallScope cs(s);
  return inv_dev_impl(count, (int)n, A, Ainv, info, s);
}

int npph_det_batched(int dtype, int64_t count, int64_t n, const void* A, void* det, void* sign, void* logabsdet,
                     int32_t* info) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_args("det", dtype, count, n));
  if (count == 0) return NPPH_OK;
  const hipStream_t s = default_stream();
  CallScope cs(s);  // held through the final synchronisation
  const size_t na = (size_t)count * n * n;
  double *dA, *out;
  int32_t* di;
  NPPH_TRY(workspace(kWsA, sizeof(double) * (na ? na : 1), (void**)&dA));
  NPPH_TRY(workspace(kWsNorm, sizeof(double) * 3 * (size_t)count, (void**)&out));
  NPPH_TRY(workspace(kWsFlags, sizeof(int32_t) * (size_t)count, (void**)&di));
  std::vector<double> tmp;
  if (na) NPPH_TRY(upload_f64(dtype, A, na, dA, s, tmp));
  NPPH_TRY(det_dev_impl(count, (int)n, dA, out, out + count, out + 2 * count, di, s));
  std::vector<double> h(3 * (size_t)count);
  NPPH_HIP(hipMemcpyAsync(h.data(), out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  if (info) NPPH_HIP(hipMemcpyAsync(info, di, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  store_out(dtype, h.data(), (size_t)count, det);
  store_out(dtype, h.data() + count, (size_t)count, sign);
  store_out(dtype, h.data() + 2 * count, (size_t)count, logabsdet);
  return NPPH_OK;
}

int npph_inv_batched(int dtype, int64_t count, int64_t n, const void* A, void* Ainv, int32_t* info) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_args("inv", dtype, count, n));
  if (count == 0 || n == 0) return NPPH_OK;
  const hipStream_t s = default_stream();
  CallScope cs(s);  // held through the final synchronisation
  const size_t na = (size_t)count * n * n;
  double *dA, *dX;
  int32_t* di;
  NPPH_TRY(workspace(kWsA, sizeof(double) * na, (void**)&dA));
  NPPH_TRY(workspace(kWsB, sizeof(double) * na, (void**)&dX));
  NPPH_TRY(workspace(kWsFlags, sizeof(int32_t) * (size_t)count, (void**)&di));
  std::vector<double> tmp;
  NPPH_TRY(upload_f64(dtype, A, na, dA, s, tmp));
  NPPH_TRY(inv_dev_impl(count, (int)n, dA, dX, di, s));
  if (info) NPPH_HIP(hipMemcpyAsync(info, di, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToHost, s));
  if (dtype == NPPH_DT_F64) {
    NPPH_HIP(hipMemcpyAsync(Ainv, dX, sizeof(double) * na, hipMemcpyDeviceToHost, s));
    NPPH_HIP(hipStreamSynchronize(s));
  } else {
    tmp.resize(na);
    NPPH_HIP(hipMemcpyAsync(tmp.data(), dX, sizeof(double) * na, hipMemcpyDeviceToHost, s));
    NPPH_HIP(hipStreamSynchronize(s));
    store_out(dtype, tmp.data(), na, Ainv);
  }
  return NPPH_OK;
}

}  // extern "C"
