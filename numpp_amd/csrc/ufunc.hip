// Element-wise ufuncs and full reductions for NumPP's GpuVTable slots
// elementwise_binary / elementwise_unary / reduce (gpu_vtable.hpp:19-21, op
// codes gpu_vtable.hpp:10-12; callers gpu_or_binary / gpu_or_unary /
// gpu_or_reduce, ufunc.cpp:525-545, 607-617).
//
// Contract (openspec gpu-kernels): add, subtract, multiply, divide, negative and
// sqrt are IEEE-exact, so results equal the CPU kernel's bit for bit (HIP's
// float and double division and sqrt are correctly rounded); exp is declined
// (returns NPPH_ERR_UNSUPPORTED, the slot returns false and NumPP's CPU kernel
// runs), as in the reference's CUDA and OpenCL providers. Sum and product
// accumulate in the array's dtype, within the reduction tolerance of the CPU's
// sequential loop (ufunc.cpp:450-472): each workgroup reduces one contiguous
// chunk, a pairwise tree combines the lanes, and a second pass combines the
// chunk results in order, so the result is deterministic.
//
// Layout: flat contiguous arrays. The element-wise kernels move 16 bytes per
// thread per access (HBM-bound); the reduction reads the array once.
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace npph {

namespace {

constexpr int EW_T = 256;

template <class T>
__device__ __forceinline__ T ew_op(int op, T x, T y) {
  switch (op) {
    case NPPH_EW_ADD: return x + y;
    case NPPH_EW_SUB: return x - y;
    case NPPH_EW_MUL: return x * y;
    default: return x / y;
  }
}

template <class T>
__device__ __forceinline__ T un_op(int op, T x) {
  return op == NPPH_UN_SQRT ? sqrt(x) : -x;
}

// V elements per access: 16 / sizeof(T) when every pointer is 16-byte aligned
// (the tail n % V by the first threads), else 1.
template <class T, int V>
__global__ __launch_bounds__(EW_T) void binary_kernel(int op, int64_t n, const T* __restrict__ a,
                                                      const T* __restrict__ b, T* __restrict__ o) {
  typedef T vec __attribute__((ext_vector_type(V)));
  const int64_t nv = n / V;
  const int64_t stride = (int64_t)gridDim.x * EW_T;
  for (int64_t i = (int64_t)blockIdx.x * EW_T + threadIdx.x; i < nv; i += stride) {
    const vec x = reinterpret_cast<const vec*>(a)[i];
    const vec y = reinterpret_cast<const vec*>(b)[i];
    vec r;
#pragma unroll
    for (int q = 0; q < V; ++q) r[q] = ew_op<T>(op, x[q], y[q]);
    reinterpret_cast<vec*>(o)[i] = r;
  }
  const int64_t t = (int64_t)blockIdx.x * EW_T + threadIdx.x;
  if (t < n - nv * V) o[nv * V + t] = ew_op<T>(op, a[nv * V + t], b[nv * V + t]);
}

template <class T, int V>
__global__ __launch_bounds__(EW_T) void unary_kernel(int op, int64_t n, const T* __restrict__ a, T* __restrict__ o) {
  typedef T vec __attribute__((ext_vector_type(V)));
  const int64_t nv = n / V;
  const int64_t stride = (int64_t)gridDim.x * EW_T;
  for (int64_t i = (int64_t)blockIdx.x * EW_T + threadIdx.x; i < nv; i += stride) {
    const vec x = reinterpret_cast<const vec*>(a)[i];
    vec r;
#pragma unroll
    for (int q = 0; q < V; ++q) r[q] = un_op<T>(op, x[q]);
    reinterpret_cast<vec*>(o)[i] = r;
  }
  const int64_t t = (int64_t)blockIdx.x * EW_T + threadIdx.x;
  if (t < n - nv * V) o[nv * V + t] = un_op<T>(op, a[nv * V + t]);
}

template <class T>
__device__ __forceinline__ T red_op(int op, T x, T y) {
  return op == NPPH_RED_PROD ? x * y : x + y;
}

// Pass 1: workgroup g reduces elements [g chunk, (g+1) chunk): each thread
// folds a strided sequence, then the 256 thread results are combined pairwise
// through LDS. Pass 2 (one workgroup) combines the partials the same way.
template <class T>
__global__ __launch_bounds__(EW_T) void reduce_kernel(int op, int64_t n, int64_t chunk, const T* __restrict__ a,
                                                      T* __restrict__ part) {
  __shared__ T s[EW_T];
  const T id = op == NPPH_RED_PROD ? T(1) : T(0);
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  T acc = id;
  for (int64_t i = lo + threadIdx.x; i < hi; i += EW_T) acc = red_op<T>(op, acc, a[i]);
  s[threadIdx.x] = acc;
  __syncthreads();
#pragma unroll
  for (int h = EW_T / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = red_op<T>(op, s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

inline hipStream_t pick(void* st) { return st ? (hipStream_t)st : default_stream(); }

int elem_size(int dtype) { return dtype == NPPH_DT_F64 ? 8 : dtype == NPPH_DT_F32 ? 4 : 0; }

unsigned ew_grid(int64_t n, int es) {
  const int64_t nv = n * es / 16 + 1;
  const int64_t want = ceil_div(nv, EW_T);
  const int64_t cap = (int64_t)std::max(cu_count(), 1) * 16;  // grid-stride beyond 16 workgroups per CU
  return (unsigned)std::max<int64_t>(1, std::min(want, cap));
}

int check_ew(const char* what, int op, int lo, int hi, int dtype, int64_t n) {
  if (op < lo || op > hi) {
    set_error("%s: op %d not supported on the device", what, op);
    return NPPH_ERR_UNSUPPORTED;
  }
  if (!elem_size(dtype)) {
    set_error("%s: dtype %d not supported on the device", what, dtype);
    return NPPH_ERR_UNSUPPORTED;
  }
  if (n < 0) { set_error("%s: negative size", what); return NPPH_ERR_INVALID; }
  return NPPH_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <class T>
void binary_t(int op, int64_t n, const void* a, const void* b, void* out, hipStream_t s) {
  constexpr int V = 16 / sizeof(T);
  const bool vec = aligned16(a) && aligned16(b) && aligned16(out);
  const dim3 g(ew_grid(vec ? n : n * V, sizeof(T)));
  if (vec)
    hipLaunchKernelGGL((binary_kernel<T, V>), g, dim3(EW_T), 0, s, op, n, (const T*)a, (const T*)b, (T*)out);
  else
    hipLaunchKernelGGL((binary_kernel<T, 1>), g, dim3(EW_T), 0, s, op, n, (const T*)a, (const T*)b, (T*)out);
}

template <class T>
void unary_t(int op, int64_t n, const void* a, void* out, hipStream_t s) {
  constexpr int V = 16 / sizeof(T);
  const bool vec = aligned16(a) && aligned16(out);
  const dim3 g(ew_grid(vec ? n : n * V, sizeof(T)));
  if (vec) hipLaunchKernelGGL((unary_kernel<T, V>), g, dim3(EW_T), 0, s, op, n, (const T*)a, (T*)out);
  else hipLaunchKernelGGL((unary_kernel<T, 1>), g, dim3(EW_T), 0, s, op, n, (const T*)a, (T*)out);
}

int binary_launch(int op, int dtype, int64_t n, const void* a, const void* b, void* out, hipStream_t s) {
  if (n == 0) return NPPH_OK;
  if (dtype == NPPH_DT_F64) binary_t<double>(op, n, a, b, out, s);
  else binary_t<float>(op, n, a, b, out, s);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

int unary_launch(int op, int dtype, int64_t n, const void* a, void* out, hipStream_t s) {
  if (n == 0) return NPPH_OK;
  if (dtype == NPPH_DT_F64) unary_t<double>(op, n, a, out, s);
  else unary_t<float>(op, n, a, out, s);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// Reduces a device array into out (device, one element of the dtype).
template <class T>
int reduce_launch_t(int op, int64_t n, const T* a, T* out, hipStream_t s) {
  // chunks of at least 16 Ki elements, at most 4 per CU (and at most 1024, so
  // the second pass is one workgroup)
  const int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>({ceil_div(n, 16384),
                                                                  4 * (int64_t)std::max(cu_count(), 1), 1024}));
  const int64_t chunk = std::max<int64_t>(1, ceil_div(n, nchunk));
  T* part;
  NPPH_TRY(workspace(kWsPart, sizeof(T) * (size_t)nchunk, (void**)&part));
  hipLaunchKernelGGL(reduce_kernel<T>, dim3((unsigned)nchunk), dim3(EW_T), 0, s, op, n, chunk, a, part);
  hipLaunchKernelGGL(reduce_kernel<T>, dim3(1), dim3(EW_T), 0, s, op, nchunk, nchunk, (const T*)part, out);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

int reduce_launch(int op, int dtype, int64_t n, const void* a, void* out, hipStream_t s) {
  if (dtype == NPPH_DT_F64) return reduce_launch_t<double>(op, n, (const double*)a, (double*)out, s);
  return reduce_launch_t<float>(op, n, (const float*)a, (float*)out, s);
}

}  // namespace

}  // namespace npph

using namespace npph;

extern "C" {

int npph_elementwise_binary_dev(int op, int dtype, int64_t n, const void* a, const void* b, void* out,
                                void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_ew("elementwise_binary", op, NPPH_EW_ADD, NPPH_EW_DIV, dtype, n));
  return binary_launch(op, dtype, n, a, b, out, pick(stream));
}

int npph_elementwise_unary_dev(int op, int dtype, int64_t n, const void* a, void* out, void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_ew("elementwise_unary", op, NPPH_UN_NEG, NPPH_UN_SQRT, dtype, n));
  return unary_launch(op, dtype, n, a, out, pick(stream));
}

int npph_reduce_dev(int op, int dtype, int64_t n, const void* a, void* out, void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_ew("reduce", op, NPPH_RED_SUM, NPPH_RED_PROD, dtype, n));
  const hipStream_t s = pick(stream);
  CallScope cs(s);
  return reduce_launch(op, dtype, n, a, out, s);
}

int npph_elementwise_binary(int op, int dtype, int64_t n, const void* a, const void* b, void* out) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_ew("elementwise_binary", op, NPPH_EW_ADD, NPPH_EW_DIV, dtype, n));
  if (n == 0) return NPPH_OK;
  const hipStream_t s = default_stream();
  CallScope cs(s);  // held through the final synchronisation
  const size_t bytes = (size_t)n * elem_size(dtype);
  void *da, *db, *dout;
  NPPH_TRY(workspace(kWsA, bytes, &da));
  NPPH_TRY(workspace(kWsB, bytes, &db));
  NPPH_TRY(workspace(kWsX, bytes, &dout));
  NPPH_HIP(hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, s));
  NPPH_HIP(hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, s));
  NPPH_TRY(binary_launch(op, dtype, n, da, db, dout, s));
  NPPH_HIP(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}

int npph_elementwise_unary(int op, int dtype, int64_t n, const void* a, void* out) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_ew("elementwise_unary", op, NPPH_UN_NEG, NPPH_UN_SQRT, dtype, n));
  if (n == 0) return NPPH_OK;
  const hipStream_t s = default_stream();
  CallScope cs(s);
  const size_t bytes = (size_t)n * elem_size(dtype);
  void *da, *dout;
  NPPH_TRY(workspace(kWsA, bytes, &da));
  NPPH_TRY(workspace(kWsX, bytes, &dout));
  NPPH_HIP(hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, s));
  NPPH_TRY(unary_launch(op, dtype, n, da, dout, s));
  NPPH_HIP(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}

int npph_reduce(int op, int dtype, int64_t n, const void* a, void* out) {
  NPPH_TRY(ensure_init());
  NPPH_TRY(check_ew("reduce", op, NPPH_RED_SUM, NPPH_RED_PROD, dtype, n));
  const hipStream_t s = default_stream();
  CallScope cs(s);
  const int es = elem_size(dtype);
  const size_t bytes = (size_t)n * es;
  void *da, *dout;
  NPPH_TRY(workspace(kWsA, bytes ? bytes : 1, &da));
  NPPH_TRY(workspace(kWsX, 16, &dout));
  if (bytes) NPPH_HIP(hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, s));
  NPPH_TRY(reduce_launch(op, dtype, n, da, dout, s));
  NPPH_HIP(hipMemcpyAsync(out, dout, es, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}

}  // extern "C"
