// numpp::GpuVTable as provided by the MI355X HIP backend.
//
// Layout: the reference struct (include/numpp/backend/gpu_vtable.hpp:17-53 in
// NumPP) member for member, followed by appended trailing solve slots, so the
// existing aggregate initialisers of the other providers (cuda_backend.cu:
// 315-316, opencl_backend.cpp:391-392, refgpu_backend.cpp:110-111) keep
// compiling with the new members null. When built inside NumPP, the real
// numpp/core/dtype.hpp is used; standalone, a layout-compatible DTypeId
// (uint8_t, Float32 = 10, Float64 = 11: dtype.hpp:16-24) is declared.
#pragma once

#include <cstdint>

#if __has_include("numpp/core/dtype.hpp")
#include "numpp/core/dtype.hpp"
#else
namespace numpp {
enum class DTypeId : uint8_t {
  Bool = 0, Int8, Int16, Int32, Int64, UInt8, UInt16, UInt32, UInt64,
  Float16, Float32, Float64, Complex64, Complex128,
  String, Bytes, Datetime64, Timedelta64, Struct,
};
}  // namespace numpp
#endif

namespace numpp {

// Precision selector of the refined solve slot (NPPH_PREC_* in numpp_hip.h).
enum class SolvePrecision : int { Fp64 = 0, Fp32Refined = 1, Bf16Refined = 2 };

struct GpuVTable {
  const char* name;
  bool (*elementwise_binary)(int op, DTypeId dt, int64_t n, const void* a, const void* b, void* out);
  bool (*elementwise_unary)(int op, DTypeId dt, int64_t n, const void* a, void* out);
  bool (*reduce)(int op, DTypeId dt, int64_t n, const void* a, void* out);
  bool (*gemm)(DTypeId dt, int64_t m, int64_t n, int64_t k, const void* A, const void* B, void* C);
  bool (*csr_spmv)(DTypeId dt, int64_t rows, int64_t cols, int64_t nnz, const int64_t* indptr,
                   const int64_t* indices, const void* data, const void* x, void* y);
  bool (*pairwise_sqdist)(DTypeId dt, int64_t m, int64_t n, int64_t dim, const void* A, const void* B,
                          void* D);
  bool (*separable_corr1d)(DTypeId dt, int64_t lines, int64_t len, const void* in, const void* weights,
                           int64_t klen, int64_t origin, int mode, double cval, void* out);

  // ---- appended: stacked solve slots (null in every other provider) ----
  // Whole-stack LU with partial pivoting + getrs on host buffers, in place:
  // A (count x n x n) -> packed LU, B (count x n x nrhs) -> X. ipiv (count x n,
  // 0-based per-step pivot rows) and info (count; k+1 = exact zero pivot at
  // step k, else 0) are nullable. Computes in float64 (linalg.cpp:20-26).
  // Returns false on an unsupported dtype or a device error -> the caller
  // runs its per-matrix CPU loop (batched.cpp:90-100).
  bool (*gesv_batched)(DTypeId dt, int64_t count, int64_t n, int64_t nrhs, void* A, void* B,
                       int32_t* ipiv, int32_t* info);
  // Mixed-precision factorisation + FP64 refinement on host buffers (A and B
  // unchanged, X written; berr count x nrhs, iters and info count; all three
  // nullable). float64 only.
  bool (*gesv_refined_batched)(int precision, DTypeId dt, int64_t count, int64_t n, int64_t nrhs,
                               const void* A, const void* B, void* X, double* berr, int32_t* iters,
                               int32_t* info);
  // Determinants of a stack (batched::det / linalg::det, batched.cpp:76,
  // linalg.cpp:393-407): A count x n x n unchanged, det (count, input dtype).
  bool (*det_batched)(DTypeId dt, int64_t count, int64_t n, const void* A, void* det, int32_t* info);
  // batched::slogdet / linalg::slogdet (batched.cpp:102-108, linalg.cpp:408-439).
  bool (*slogdet_batched)(DTypeId dt, int64_t count, int64_t n, const void* A, void* sign, void* logabsdet,
                          int32_t* info);
  // batched::inv / linalg::inv (batched.cpp:75, linalg.cpp:105-115, 387-392):
  // Ainv count x n x n in the input dtype; info > 0 marks a singular matrix,
  // for which the caller raises linalg_error("Singular matrix").
  bool (*inv_batched)(DTypeId dt, int64_t count, int64_t n, const void* A, void* Ainv, int32_t* info);
};

// Null unless a usable gfx950 device is present (gpu_vtable.hpp:56).
const GpuVTable* gpu_vtable();

}  // namespace numpp
