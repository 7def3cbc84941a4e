/*
 * numpp_hip.h — C ABI of the MI355X (gfx950) device backend for NumPP's
 * numpp::matmul and stacked numpp::linalg::solve.
 *
 * Plain C: POD arguments, int status codes, no C++ or torch types. Two call
 * families:
 *   - host-pointer calls (npph_gemm, npph_gesv_batched, ...) follow the
 *     reference GpuVTable slot contract (include/numpp/backend/gpu_vtable.hpp:
 *     14-16, 22-26): contiguous row-major host buffers in the native dtype,
 *     synchronous, results in the host buffers on return;
 *   - device-resident calls (suffix _dev) take device pointers and a hipStream_t
 *     passed as void* (NULL = the library's own stream) and are asynchronous.
 *
 * Status: NPPH_OK (0) or a negative NPPH_ERR_*; npph_last_error() returns a
 * thread-local message for the last failure on the calling thread.
 * Row-major everywhere (NumPP is C-order). Pivots: ipiv[k] is the 0-based row
 * position swapped with row k at step k (LAPACK convention, 0-based); the
 * final permutation of linalg.cpp:58-62 follows by applying the swaps in order.
 * Singularity uses NumPP's exact-zero rule (linalg.cpp:57): info = k+1 when the
 * largest candidate pivot at step k is exactly 0.0, else 0.
 */
#ifndef NUMPP_HIP_H
#define NUMPP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NPPH_OK 0
#define NPPH_ERR_NO_DEVICE (-1)
#define NPPH_ERR_INVALID (-2)
#define NPPH_ERR_HIP (-3)
#define NPPH_ERR_OOM (-4)
#define NPPH_ERR_UNSUPPORTED (-5)
#define NPPH_ERR_BUDGET (-6)
#define NPPH_ERR_NOT_FOUND (-7)
#define NPPH_ERR_SINGULAR (-8)

/* dtype ids: the values of numpp::DTypeId (include/numpp/core/dtype.hpp:16-24). */
#define NPPH_DT_F32 10
#define NPPH_DT_F64 11

/* Factorisation precision for the refined solver. */
#define NPPH_PREC_FP64 0    /* plain FP64 LU, no refinement needed */
#define NPPH_PREC_FP32 1    /* FP32 LU on MFMA (FP32 accumulate) + FP64 refinement */
#define NPPH_PREC_BF16 2    /* BF16-split MFMA updates with per-panel scaling + FP64 refinement */

/* ---------------- runtime ---------------- */
/* Probes the device once (gfx950 check), creates the library stream.
 * No reference counterpart: the reference probes lazily on every call
 * (src/backend/cuda_backend.cu:120-123). */
int npph_init(void);
/* 1 when a usable gfx950 device is present (mirrors the non-null accessor
 * contract of gpu_vtable(), cuda_backend.cu:320), else 0. */
int npph_available(void);
const char* npph_last_error(void);
const char* npph_device_name(void);
int npph_cu_count(void);
/* HIP device ordinal the library runs on (the DLPack device_id of its arrays). */
int npph_device_id(void);

/* Device memory / stream plumbing for callers that keep data resident. */
int npph_malloc(void** ptr, uint64_t bytes);
int npph_free(void* ptr);
int npph_memcpy_htod(void* dst, const void* src, uint64_t bytes, void* stream);
int npph_memcpy_dtoh(void* dst, const void* src, uint64_t bytes, void* stream);
int npph_memcpy_dtod(void* dst, const void* src, uint64_t bytes, void* stream);
int npph_memset(void* dst, int value, uint64_t bytes, void* stream);
int npph_stream_create(void** stream);
int npph_stream_destroy(void* stream);
int npph_stream_sync(void* stream);
int npph_device_sync(void);
int npph_event_create(void** ev);
int npph_event_destroy(void* ev);
int npph_event_record(void* ev, void* stream);
int npph_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---------------- element-wise ufuncs and reductions ---------------- */
/* Op codes: the values of GpuBinOp / GpuUnOp / GpuRedOp
 * (include/numpp/backend/gpu_vtable.hpp:10-12). */
#define NPPH_EW_ADD 0
#define NPPH_EW_SUB 1
#define NPPH_EW_MUL 2
#define NPPH_EW_DIV 3
#define NPPH_UN_NEG 0
#define NPPH_UN_SQRT 1
#define NPPH_UN_EXP 2
#define NPPH_RED_SUM 0
#define NPPH_RED_PROD 1
/* Replace GpuVTable::elementwise_binary / elementwise_unary / reduce
 * (gpu_vtable.hpp:19-21; CUDA provider ew_binary / ew_unary / reduce_impl,
 * src/backend/cuda_backend.cu:270-287; callers ufunc.cpp:525-545, 607-617):
 * contiguous host buffers of n elements in dtype F32 or F64, synchronous.
 * Binary and unary results are IEEE-exact (the CPU kernel's, bit for bit);
 * exp and other dtypes return NPPH_ERR_UNSUPPORTED (the slot returns false and
 * NumPP's CPU kernel runs, as with the reference's CUDA/OpenCL providers).
 * reduce writes one element: the sum or product in the array's dtype, within
 * reduction tolerance of the CPU's sequential loop (ufunc.cpp:450-472). */
int npph_elementwise_binary(int op, int dtype, int64_t n, const void* a, const void* b, void* out);
int npph_elementwise_unary(int op, int dtype, int64_t n, const void* a, void* out);
int npph_reduce(int op, int dtype, int64_t n, const void* a, void* out);
/* Device-resident forms (device pointers; out of reduce is one device element). */
int npph_elementwise_binary_dev(int op, int dtype, int64_t n, const void* a, const void* b, void* out,
                                void* stream);
int npph_elementwise_unary_dev(int op, int dtype, int64_t n, const void* a, void* out, void* stream);
int npph_reduce_dev(int op, int dtype, int64_t n, const void* a, void* out, void* stream);

/* ---------------- GEMM ---------------- */
/* Replaces GpuVTable::gemm (gpu_vtable.hpp:22-26; CUDA provider gemm_impl/
 * gemm_t, src/backend/cuda_backend.cu:189-206, 288-293): row-major
 * C[m,n] = A[m,k] B[k,n] on host buffers, dtype F32 or F64, synchronous.
 * Returns NPPH_ERR_UNSUPPORTED for other dtypes (the vtable slot then returns
 * false and matmul falls back to the CPU, backend.cpp:144-152). */
int npph_gemm(int dtype, int64_t m, int64_t n, int64_t k, const void* A, const void* B, void* C);

/* Device-resident row-major C = alpha A B + beta C (FP64 MFMA), batched over
 * `batch` matrices with element strides sA/sB/sC. No reference counterpart
 * (the reference keeps no device-resident arrays, docs/gpu-backends.md:30-31);
 * it is the roofline path of the matmul slot above. */
int npph_dgemm_dev(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                   int64_t sA, const double* B, int64_t ldb, int64_t sB, double beta, double* C,
                   int64_t ldc, int64_t sC, int64_t batch, void* stream);
/* The same in FP32 (FP32 MFMA, FP32 accumulation). */
int npph_sgemm_dev(int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                   int64_t sA, const float* B, int64_t ldb, int64_t sB, float beta, float* C,
                   int64_t ldc, int64_t sC, int64_t batch, void* stream);

/* Device-resident C -= A B for FP32 matrices through the split-BF16 MFMA
 * update of the low-precision LU (terms 1, 3 or 6). No reference
 * counterpart; exposed so the update can be checked on its own. */
int npph_bf16split_update_dev(int64_t M, int64_t N, int64_t K, const float* A, int64_t lda,
                              const float* B, int64_t ldb, float* C, int64_t ldc, int terms,
                              void* stream);

/* ---------------- stacked solve (plain) ---------------- */
/* New GpuVTable slot gesv_batched: replaces the per-matrix loop of
 * batched::solve (src/linalg/batched.cpp:90-100) and solve_impl<double>
 * (src/linalg/linalg.cpp:92-103; lu_decompose 48-71, lu_solve 73-90) with one
 * device call for the whole stack. A: count x n x n, overwritten with the packed
 * LU factors (L unit-lower below the diagonal, U on and above). B: count x n x
 * nrhs, overwritten with X. ipiv (nullable): count x n. info (nullable): count.
 * dtype F64 (F32 is computed in F64 and rounded back, as linalg.cpp:20-26).
 * Returns NPPH_OK even when systems are singular; check info. */
int npph_gesv_batched(int dtype, int64_t count, int64_t n, int64_t nrhs, void* A, void* B,
                      int32_t* ipiv, int32_t* info);

/* Device-resident FP64 LU with partial pivoting of a uniform stack (element
 * stride strideA between matrices), getrs-style solve with those factors, and
 * both together. Same factor / pivot / info conventions as above. */
int npph_getrf_batched_dev(int64_t count, int64_t n, double* A, int64_t strideA, int32_t* ipiv,
                           int32_t* info, void* stream);
int npph_getrs_batched_dev(int64_t count, int64_t n, int64_t nrhs, const double* LU,
                           int64_t strideA, const int32_t* ipiv, double* B, int64_t strideB,
                           void* stream);
int npph_gesv_batched_dev(int64_t count, int64_t n, int64_t nrhs, double* A, double* B,
                          int32_t* ipiv, int32_t* info, void* stream);

/* Allocation-free form of npph_gesv_batched_dev, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. Same
 * arguments and results as npph_gesv_batched_dev; its scratch (pivots and info
 * when passed NULL, the permutation, row-move lists and staging, solve staging)
 * comes from a fixed 64 MiB device array per GPU instead of library
 * allocations, and it calls no allocating routine. Calls whose scratch would
 * exceed 64 MiB fail with NPPH_ERR_INVALID before any work is issued (n <= 64
 * needs at most the pivots and info; n = 2048 with 8 right-hand sides needs
 * about 18 MB per matrix). Every call on one GPU shares that array: calls that
 * overlap on one GPU (different streams or threads) race and must be ordered
 * by the caller. */
int npph_gesv_batched_dev_noalloc(int64_t count, int64_t n, int64_t nrhs, double* A, double* B,
                                  int32_t* ipiv, int32_t* info, void* stream);

/* ---------------- determinants and inverses ---------------- */
/* New GpuVTable slots det_batched / slogdet_batched: replace the per-matrix
 * loops of batched::det and batched::slogdet (src/linalg/batched.cpp:76,
 * 102-108) over linalg::det / linalg::slogdet (src/linalg/linalg.cpp:393-407,
 * 408-439) with one device LU of the whole stack. A: count x n x n host buffer
 * (not modified), dtype F64 or F32 (computed in F64, outputs rounded to the
 * input dtype as linalg.cpp:20-26). Outputs (each nullable, count entries, in
 * the input dtype): det = permsign * prod u_ii in row order, or 0 when an
 * exact-zero pivot appears; sign and logabsdet as slogdet, (0, -inf) when
 * singular. info (nullable): 0, or the first zero-pivot step + 1. */
int npph_det_batched(int dtype, int64_t count, int64_t n, const void* A, void* det, void* sign,
                     void* logabsdet, int32_t* info);
/* New GpuVTable slot inv_batched: replaces batched::inv (batched.cpp:75) over
 * linalg::inv / inv_impl (linalg.cpp:105-115, 387-392): LU of each matrix,
 * then lu_solve against the identity. A not modified; Ainv count x n x n in the
 * input dtype. info as above; a singular system's Ainv is unspecified (the
 * reference raises linalg_error("Singular matrix")). */
int npph_inv_batched(int dtype, int64_t count, int64_t n, const void* A, void* Ainv, int32_t* info);
/* Device-resident FP64 forms: A (count x n x n, stride n*n) is overwritten by
 * its packed LU factors; outputs are device arrays. */
int npph_det_batched_dev(int64_t count, int64_t n, double* A, double* det, double* sign, double* logabsdet,
                         int32_t* info, void* stream);
int npph_inv_batched_dev(int64_t count, int64_t n, double* A, double* Ainv, int32_t* info, void* stream);

/* ---------------- LapackVTable provider (one system, output dtype) ---------------- */
/* Replace LapackVTable::gesv / inv (include/numpp/backend/lapack_vtable.hpp:
 * 14-18; reference provider lapack_backend.cpp:15-34), which linalg::solve
 * tries before its portable kernel (linalg.cpp:372-382). Unlike the portable
 * kernel they compute in the given dtype: F32 systems are factored and solved
 * in FP32 (LAPACKE_sgesv's contract), F64 in FP64. Host buffers, row-major;
 * gesv overwrites A with the packed LU and B (n x nrhs) with X; inv writes
 * Ainv and leaves A unchanged. An exact-zero pivot returns NPPH_ERR_SINGULAR
 * (the slot returns false and NumPP's portable kernel raises). */
int npph_lapack_gesv(int n, int nrhs, int dtype, void* A, void* B);
int npph_lapack_inv(int n, int dtype, const void* A, void* Ainv);
/* Address of the numpp::LapackVTable provider struct (layout of
 * include/numpp_hip/lapack_vtable.hpp: gesv and inv filled, the other slots
 * null), or NULL without a usable device. */
const void* npph_lapack_vtable(void);

/* ---------------- mixed precision + FP64 iterative refinement ---------------- */
/* Options for the refined solver. No reference counterpart (NumPP always
 * factors in FP64, linalg.cpp:20-26); this is the precision addition. */
typedef struct npph_refine_opts {
  int precision;      /* NPPH_PREC_* */
  int max_iter;       /* refinement sweeps before FP64 fallback (e.g. 10) */
  double tol;         /* target normwise backward error, e.g. 1e-12 */
  int fallback_fp64;  /* 1: systems not converged after max_iter are re-solved in plain FP64 */
  int bf16_terms;     /* BF16 split products per update: 1, 3 (2-term split) or 6 (3-term) */
} npph_refine_opts;

/* Solves a uniform stack A X = B (A count x n x n, B/X count x n x nrhs, FP64,
 * device pointers; A and B are not modified). Factorises in opts->precision,
 * refines in FP64 on the device (r = b - A x, correction with the
 * low-precision factors, x += d) until every RHS column meets
 * eta = ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= tol.
 * Outputs (device, nullable): berr[count*nrhs] final eta per column,
 * iters[count] sweeps used (-1 = fell back to FP64), info[count] singular step
 * + 1 or 0. */
int npph_gesv_refined_batched_dev(const npph_refine_opts* opts, int64_t count, int64_t n,
                                  int64_t nrhs, const double* A, const double* B, double* X,
                                  double* berr, int32_t* iters, int32_t* info, void* stream);

/* ---------------- uneven batches ---------------- */
/* Many independent systems of different order n_i and nrhs_i in one call. No
 * reference counterpart (a NumPP stack must be uniform, batched.cpp:22-30); the
 * per-system semantics are those of npph_gesv_batched / the refined solver.
 * Packed layout: system i's A starts at element sum_{j<i} n_j^2, its B and X at
 * sum_{j<i} n_j*nrhs_j, its berr at sum_{j<i} nrhs_j; all row-major.
 * plan_create sorts the systems into padded size buckets (identity padding;
 * granularity chosen by a measured autotuner when `autotune` is 1, cached in
 * $NUMPP_HIP_TUNE_CACHE or ~/.cache/numpp_hip_tune.json) and allocates the
 * bucket buffers. execute packs, solves every bucket and unpacks; with
 * use_graph = 1 and FP64 precision the whole sequence is captured once into a
 * hipGraph (first call runs eagerly to size workspaces) and replayed.
 * n_i in [1, 4096], nrhs_i in [1, 4096]; host arrays n, nrhs. */
int npph_vbatch_plan_create(int64_t count, const int32_t* n, const int32_t* nrhs,
                            const npph_refine_opts* opts, int autotune, void** plan);
int npph_vbatch_plan_destroy(void* plan);
int npph_vbatch_plan_info(void* plan, int* nbuckets, int32_t* order, int64_t* counts, int32_t* rpad,
                          int cap);
int npph_vbatch_execute(void* plan, const double* A, const double* B, double* X, double* berr,
                        int32_t* iters, int32_t* info, int use_graph, void* stream);

/* ---------------- serving loop ---------------- */
/* Request loop for solves of different n and nrhs arriving over time. No
 * reference counterpart (NumPP's calls are synchronous and keep nothing on the
 * device, docs/gpu-backends.md:30-31). W worker threads, each with its own
 * stream and workspace, take requests from a FIFO queue when their device
 * bytes fit in the budget (resident factors are evicted LRU to make room);
 * queued one-shot requests of the same order are solved together as one stack
 * (continuous batching; results are bitwise the same alone and in the mix).
 * keep = 1 with a nonzero key leaves the factorisation resident; a later
 * request with that key and A = NULL solves with the resident factors.
 * Host buffers A, B, X, berr must stay valid until npph_server_wait returns. */
typedef struct npph_serve_config {
  double budget_bytes;  /* device byte budget for admitted work + resident factors */
  int workers;          /* concurrent worker streams (1..16) */
  int max_batch;        /* max same-order requests solved as one stack */
  int precision;        /* NPPH_PREC_* */
  int max_iter;
  double tol;
  int fallback_fp64;
  int bf16_terms;
} npph_serve_config;
int npph_server_create(const npph_serve_config* cfg, void** srv);
int npph_server_destroy(void* srv);
int npph_server_submit(void* srv, int64_t n, int64_t nrhs, const double* A, const double* B, double* X,
                       double* berr, uint64_t key, int keep, int64_t* req_id);
/* Blocks until the request is done; returns its status (NPPH_ERR_BUDGET when
 * it can never fit). latency_ms = submit to results on the host; queue_ms =
 * submit to admission. */
int npph_server_wait(void* srv, int64_t req_id, int32_t* iters, int32_t* info, double* latency_ms,
                     double* queue_ms);
int npph_server_stats(void* srv, int64_t* completed, int64_t* resident, double* bytes_in_use,
                      double* peak_bytes, int64_t* batched, int64_t* fallbacks, int64_t* evictions);
/* Replays a trace: request i is submitted t_arrive[i] seconds after the call
 * starts; per-request latency, iters, info and status are returned and
 * wall_s is the first-submit-to-last-completion time. */
int npph_server_replay(void* srv, int64_t count, const double* t_arrive, const int32_t* n,
                       const int32_t* nrhs, const double* const* A, const double* const* B,
                       double* const* X, double* const* berr, const uint64_t* key,
                       const int32_t* keep, double* latency_ms, int32_t* iters, int32_t* info,
                       int32_t* status, double* wall_s);

/* ---------------- profiling ---------------- */
/* Per-kernel-class timing with HIP events around every launch (off by default;
 * no reference counterpart: NumPP has no timers or benchmarks).
 * Kinds: 0 FP64 GEMM, 1 FP32 GEMM, 2 split-BF16 update, 3 LU panel,
 * 4 row-move + trsm, 5 getrs, 6 FP64 residual, 7 one-wave LU, 8 uneven batch.
 * npph_prof_get synchronises and returns summed milliseconds, algorithmic
 * flops and bytes, and the launch count of one kind since the last reset. */
#define NPPH_PROF_GEMM_F64 0
#define NPPH_PROF_GEMM_F32 1
#define NPPH_PROF_GEMM_BF16 2
#define NPPH_PROF_PANEL 3
#define NPPH_PROF_MOVE_TRSM 4
#define NPPH_PROF_GETRS 5
#define NPPH_PROF_RESIDUAL 6
#define NPPH_PROF_SMALL_LU 7
#define NPPH_PROF_VBATCH 8
int npph_prof_enable(int on);
int npph_prof_reset(void);
int npph_prof_get(int kind, double* ms, double* flops, double* bytes, int64_t* launches);

/* ---------------- drop-in vtable ---------------- */
/* Address of the numpp::GpuVTable provider struct (layout of
 * include/numpp_hip/gpu_vtable.hpp, i.e. the reference gpu_vtable.hpp:17-53
 * plus appended solve slots), or NULL when no usable device is present —
 * the same contract as numpp::gpu_vtable() (gpu_vtable.hpp:56). */
const void* npph_gpu_vtable(void);

#ifdef __cplusplus
}
#endif

#endif /* NUMPP_HIP_H */
