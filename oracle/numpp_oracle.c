/*
 * CPU oracle for the numpp matmul / stacked-solve path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP backend is
 * compared against; it is never linked into, called by, or used as a fallback
 * for the product library (numpp_amd/lib/libnumpp_hip.so). Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function restates NumPP's portable CPU kernel loop for loop, in plain
 * C, so that it computes in the same operation order:
 *   orc_lu_decompose   <- src/linalg/linalg.cpp:48-71   (lu_decompose<T>)
 *   orc_lu_solve       <- src/linalg/linalg.cpp:73-90   (lu_solve<T>)
 *   orc_solve          <- src/linalg/linalg.cpp:92-103  (solve_impl<double>)
 *   orc_batched_solve  <- src/linalg/batched.cpp:90-100 (batched::solve loop)
 *   orc_matmul         <- src/backend/backend.cpp:55-84 (cpu_matmul, f64)
 *   orc_matmul_f32     <- src/backend/backend.cpp:55-84 (cpu_matmul, f32)
 *   orc_det            <- src/linalg/linalg.cpp:393-407 (det)
 *   orc_slogdet        <- src/linalg/linalg.cpp:408-439 (slogdet, real case)
 *   orc_inv            <- src/linalg/linalg.cpp:105-115 (inv_impl)
 *   orc_nd_copy        <- src/core/ndarray.cpp:216-231, 257-285 (copy / astype:
 *                         element-by-element multi-index loop), for the
 *                         "NumPP-equivalent" CPU timings below
 *   orc_batched_solve_numpp <- batched.cpp:90-100 with the ndarray copies each
 *                         matrix goes through (copy(), to_vec, from_vec, stack)
 *   orc_matmul_numpp   <- backend.cpp:55-84 with its two astype copies
 *   orc_reduce_f64/f32 <- src/umath/ufunc.cpp:450-472 (reduce_row_numeric, sum
 *                         and prod: one sequential accumulator in the dtype)
 *   orc_ew_f64/f32     <- the portable element-wise kernels' arithmetic
 *                         (add, subtract, multiply, divide, negative, sqrt)
 * Helpers that NumPP does not have (pinned by mathematics, not by NumPP):
 *   orc_dot2_entry     compensated dot product (Ogita-Rump-Oishi Dot2) of one
 *                      GEMM entry, used to bound GEMM error at sizes where the
 *                      naive oracle is too slow.
 *   orc_backward_error normwise backward error of a computed solution.
 *
 * Build: -O3 -ffp-contract=off (NumPP's default x86-64 build emits no FMA, so
 * contraction stays off to keep the same rounding as the reference loops).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   lu_decompose: pivot rows equal LAPACK dgetrf ipiv (scipy.linalg.lu_factor
 *     goldens) on well-separated inputs; P.A = L.U reconstruction; exact
 *     rational Gaussian elimination for n <= 4; singular exact-zero rule.
 *   lu_solve / solve / batched_solve: numpy.linalg.solve goldens; exact
 *     rational solutions for n <= 4; identity / permutation / triangular cases.
 *   matmul: the reference's literal [[22,28],[49,64]] (tests/test_backend.cpp:
 *     46-55), the reference's frozen golden tests/golden/g15653376641833044941
 *     .npy, numpy goldens, exact integer products.
 *   dot2_entry: exact rational products on integer-valued inputs.
 *   det / slogdet / inv: numpy.linalg goldens (tests/golden/det_cases.npz);
 *     exact rational determinants and inverses for n <= 4; identity,
 *     permutation (sign of the permutation exactly) and triangular matrices
 *     (product of the diagonal exactly); exact-zero singular inputs (det 0,
 *     slogdet (0, -inf), inv singular); A inv(A) = I.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* linalg.cpp:48-71. In place, row-major n x n. piv[n] receives the final
 * permutation (piv[k] = original row now at position k), steps[n] (nullable)
 * the pivot row chosen at each step k (LAPACK-style ipiv, 0-based).
 * Returns -1 on success, or the step k at which the best pivot was exactly 0. */
int orc_lu_decompose(double* A, int n, int* piv, int* steps, double* permsign) {
  for (int i = 0; i < n; ++i) piv[i] = i;
  double sgn = 1.0;
  for (int k = 0; k < n; ++k) {
    int pr = k;
    double best = fabs(A[(int64_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      double v = fabs(A[(int64_t)i * n + k]);
      if (v > best) { best = v; pr = i; }
    }
    if (steps) steps[k] = pr;
    if (best == 0.0) { if (permsign) *permsign = sgn; return k; }
    if (pr != k) {
      for (int j = 0; j < n; ++j) {
        double t = A[(int64_t)k * n + j];
        A[(int64_t)k * n + j] = A[(int64_t)pr * n + j];
        A[(int64_t)pr * n + j] = t;
      }
      int t = piv[k]; piv[k] = piv[pr]; piv[pr] = t;
      sgn = -sgn;
    }
    const double akk = A[(int64_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double f = A[(int64_t)i * n + k] / akk;
      A[(int64_t)i * n + k] = f;
      double* ri = A + (int64_t)i * n;
      const double* rk = A + (int64_t)k * n;
      for (int j = k + 1; j < n; ++j) ri[j] -= f * rk[j];
    }
  }
  if (permsign) *permsign = sgn;
  return -1;
}

/* linalg.cpp:73-90. B is n x m row-major, replaced by X. */
void orc_lu_solve(const double* LU, int n, const int* piv, double* B, int m) {
  double* X = (double*)malloc(sizeof(double) * (size_t)n * (size_t)(m > 0 ? m : 1));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) X[(int64_t)i * m + j] = B[(int64_t)piv[i] * m + j];
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < i; ++k)
      for (int j = 0; j < m; ++j) X[(int64_t)i * m + j] -= LU[(int64_t)i * n + k] * X[(int64_t)k * m + j];
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k)
      for (int j = 0; j < m; ++j) X[(int64_t)i * m + j] -= LU[(int64_t)i * n + k] * X[(int64_t)k * m + j];
    double d = LU[(int64_t)i * n + i];
    for (int j = 0; j < m; ++j) X[(int64_t)i * m + j] /= d;
  }
  memcpy(B, X, sizeof(double) * (size_t)n * (size_t)m);
  free(X);
}

/* linalg.cpp:92-103: copy A and B, factor, solve. X receives the solution,
 * LU (nullable) the factors, steps (nullable) the per-step pivot rows.
 * Returns -1 on success or the singular step k (reference: linalg_error). */
int orc_solve(const double* A, const double* B, double* X, double* LU, int* steps, int n, int m) {
  double* W = (double*)malloc(sizeof(double) * (size_t)n * (size_t)(n > 0 ? n : 1));
  int* piv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  memcpy(W, A, sizeof(double) * (size_t)n * (size_t)n);
  memcpy(X, B, sizeof(double) * (size_t)n * (size_t)m);
  double sgn;
  int sing = orc_lu_decompose(W, n, piv, steps, &sgn);
  if (sing < 0) orc_lu_solve(W, n, piv, X, m);
  if (LU) memcpy(LU, W, sizeof(double) * (size_t)n * (size_t)n);
  free(W); free(piv);
  return sing;
}

/* linalg.cpp:393-407: the determinant from a copy of A: 0 when lu_decompose
 * reports singular, else permsign times the diagonal of U, multiplied in row
 * order starting from the sign. */
double orc_det(const double* A, int n) {
  double* W = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n * n : 1));
  int* piv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  memcpy(W, A, sizeof(double) * (size_t)n * (size_t)n);
  double sgn = 1.0, d;
  if (orc_lu_decompose(W, n, piv, NULL, &sgn) >= 0) {
    d = 0.0;
  } else {
    d = sgn;
    for (int i = 0; i < n; ++i) d *= W[(int64_t)i * n + i];
  }
  free(W); free(piv);
  return d;
}

/* linalg.cpp:408-439 for real matrices: singular -> sign 0, logabsdet -inf;
 * else logabsdet = sum of log|u_ii| and sign = permsign times the product of
 * u_ii / |u_ii|, both in row order. */
void orc_slogdet(const double* A, int n, double* sign, double* logabsdet) {
  double* W = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n * n : 1));
  int* piv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  memcpy(W, A, sizeof(double) * (size_t)n * (size_t)n);
  double sgn = 1.0;
  if (orc_lu_decompose(W, n, piv, NULL, &sgn) >= 0) {
    *sign = 0.0;
    *logabsdet = -INFINITY;
  } else {
    double la = 0.0, sg = sgn;
    for (int i = 0; i < n; ++i) {
      const double d = W[(int64_t)i * n + i];
      const double m = fabs(d);
      la += log(m);
      sg *= d / m;
    }
    *sign = sg;
    *logabsdet = la;
  }
  free(W); free(piv);
}

/* linalg.cpp:105-115: factor a copy of A and solve against the identity (n
 * right-hand-side columns). Returns -1 on success or the singular step (the
 * reference raises linalg_error("Singular matrix")). */
int orc_inv(const double* A, int n, double* Ainv) {
  double* W = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n * n : 1));
  int* piv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  memcpy(W, A, sizeof(double) * (size_t)n * (size_t)n);
  double sgn;
  const int sing = orc_lu_decompose(W, n, piv, NULL, &sgn);
  if (sing < 0) {
    for (int64_t i = 0; i < (int64_t)n * n; ++i) Ainv[i] = 0.0;
    for (int i = 0; i < n; ++i) Ainv[(int64_t)i * n + i] = 1.0;
    orc_lu_solve(W, n, piv, Ainv, n);
  }
  free(W); free(piv);
  return sing;
}

/* batched.cpp:90-100: per-matrix loop over a (count, n, n) stack and a
 * (count, n, m) right-hand side. info[i] = -1 ok or the singular step. The
 * reference raises on the first singular matrix; the return value is the index
 * of that matrix (or -1). */
int64_t orc_batched_solve(int64_t count, int n, int m, const double* A, const double* B,
                          double* X, int* info) {
  int64_t first = -1;
  const int64_t as = (int64_t)n * n, bs = (int64_t)n * m;
  for (int64_t i = 0; i < count; ++i) {
    int s = orc_solve(A + i * as, B + i * bs, X + i * bs, NULL, NULL, n, m);
    if (info) info[i] = s;
    if (s >= 0 && first < 0) first = i;
  }
  return first;
}

/* ndarray.cpp:216-231 (copy) and 257-285 (astype) for a 2-D float64 array:
 * a fresh buffer, then one element at a time through the multi-index loop
 * (offsets recomputed from the index vector and the strides per element). */
double* orc_nd_copy(const double* src, int64_t rows, int64_t cols) {
  const int64_t shape[2] = {rows, cols};
  const int64_t strides[2] = {cols * 8, 8};
  double* out = (double*)malloc(sizeof(double) * (size_t)(rows * cols > 0 ? rows * cols : 1));
  if (rows * cols == 0) return out;
  int64_t idx[2] = {0, 0};
  const char* sb = (const char*)src;
  char* db = (char*)out;
  for (;;) {
    int64_t soff = 0, doff = 0;
    for (int i = 0; i < 2; ++i) { soff += idx[i] * strides[i]; doff += idx[i] * strides[i]; }
    memcpy(db + doff, sb + soff, 8);
    int ax = 1;
    for (; ax >= 0; --ax) { if (++idx[ax] < shape[ax]) break; idx[ax] = 0; }
    if (ax < 0) break;
  }
  return out;
}

/* batched.cpp:90-100 as NumPP executes it, copies included: per matrix,
 * s.a3[i].copy() and b3[i].copy() (ndarray copy), to_vec (astype + vector
 * construction, linalg.cpp:35-40), the solve, from_vec (buffer memcpy + astype,
 * linalg.cpp:41-46), and stack's astype + memcpy into the output
 * (batched.cpp:34-48). X receives the solutions; returns as orc_batched_solve. */
int64_t orc_batched_solve_numpp(int64_t count, int n, int m, const double* A, const double* B, double* X) {
  int64_t first = -1;
  const int64_t as = (int64_t)n * n, bs = (int64_t)n * m;
  int* piv = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < count; ++i) {
    double* ac = orc_nd_copy(A + i * as, n, n);       /* s.a3[i].copy() */
    double* bc = orc_nd_copy(B + i * bs, n, m);       /* b3[i].copy() */
    double* at = orc_nd_copy(ac, n, n);               /* to_vec: astype */
    double* bt = orc_nd_copy(bc, n, m);
    double* av = (double*)malloc(sizeof(double) * (size_t)(as > 0 ? as : 1));  /* std::vector(p, p + size) */
    double* bv = (double*)malloc(sizeof(double) * (size_t)(bs > 0 ? bs : 1));
    memcpy(av, at, sizeof(double) * (size_t)as);
    memcpy(bv, bt, sizeof(double) * (size_t)bs);
    double sgn;
    const int s = orc_lu_decompose(av, n, piv, NULL, &sgn);
    if (s < 0) orc_lu_solve(av, n, piv, bv, m);
    else if (first < 0) first = i;
    double* r = (double*)malloc(sizeof(double) * (size_t)(bs > 0 ? bs : 1));  /* from_vec */
    memcpy(r, bv, sizeof(double) * (size_t)bs);
    double* ro = orc_nd_copy(r, n, m);                /* from_vec: astype(out) */
    double* rs = orc_nd_copy(ro, n, m);               /* stack: astype(cd) */
    memcpy(X + i * bs, rs, sizeof(double) * (size_t)bs);
    free(ac); free(bc); free(at); free(bt); free(av); free(bv); free(r); free(ro); free(rs);
  }
  free(piv);
  return first;
}

/* backend.cpp:55-84 for float64: zero the row, then i-p-j accumulation. */
void orc_matmul(int64_t m, int64_t n, int64_t k, const double* A, const double* B, double* C) {
  for (int64_t i = 0; i < m; ++i) {
    double* crow = C + i * n;
    for (int64_t j = 0; j < n; ++j) crow[j] = 0.0;
    for (int64_t p = 0; p < k; ++p) {
      const double aip = A[i * k + p];
      const double* brow = B + p * n;
      for (int64_t j = 0; j < n; ++j) crow[j] += aip * brow[j];
    }
  }
}

/* backend.cpp:55-84 as NumPP executes it for float64: the astype copies of
 * both operands (lines 59-60), then the i-p-j loop. */
void orc_matmul_numpp(int64_t m, int64_t n, int64_t k, const double* A, const double* B, double* C) {
  double* ac = orc_nd_copy(A, m, k);
  double* bc = orc_nd_copy(B, k, n);
  orc_matmul(m, n, k, ac, bc, C);
  free(ac); free(bc);
}

/* ufunc.cpp:450-472 for Sum (op 0) and Prod (op 1): acc starts at 0 or 1 and
 * takes every element in order. */
double orc_reduce_f64(int op, int64_t n, const double* p) {
  double acc = op == 1 ? 1.0 : 0.0;
  for (int64_t i = 0; i < n; ++i) acc = op == 1 ? acc * p[i] : acc + p[i];
  return acc;
}
float orc_reduce_f32(int op, int64_t n, const float* p) {
  float acc = op == 1 ? 1.0f : 0.0f;
  for (int64_t i = 0; i < n; ++i) acc = op == 1 ? acc * p[i] : acc + p[i];
  return acc;
}

/* Element-wise arithmetic as NumPP's CPU kernels compute it (IEEE operations
 * in the dtype): binary ops 0..3 = add, subtract, multiply, divide; unary ops
 * 10 = negative, 11 = sqrt (b unused). */
void orc_ew_f64(int op, int64_t n, const double* a, const double* b, double* o) {
  for (int64_t i = 0; i < n; ++i) {
    switch (op) {
      case 0: o[i] = a[i] + b[i]; break;
      case 1: o[i] = a[i] - b[i]; break;
      case 2: o[i] = a[i] * b[i]; break;
      case 3: o[i] = a[i] / b[i]; break;
      case 10: o[i] = -a[i]; break;
      default: o[i] = sqrt(a[i]); break;
    }
  }
}
void orc_ew_f32(int op, int64_t n, const float* a, const float* b, float* o) {
  for (int64_t i = 0; i < n; ++i) {
    switch (op) {
      case 0: o[i] = a[i] + b[i]; break;
      case 1: o[i] = a[i] - b[i]; break;
      case 2: o[i] = a[i] * b[i]; break;
      case 3: o[i] = a[i] / b[i]; break;
      case 10: o[i] = -a[i]; break;
      default: o[i] = sqrtf(a[i]); break;
    }
  }
}

/* backend.cpp:55-84 for float32 (compute dtype stays float32). */
void orc_matmul_f32(int64_t m, int64_t n, int64_t k, const float* A, const float* B, float* C) {
  for (int64_t i = 0; i < m; ++i) {
    float* crow = C + i * n;
    for (int64_t j = 0; j < n; ++j) crow[j] = 0.0f;
    for (int64_t p = 0; p < k; ++p) {
      const float aip = A[i * k + p];
      const float* brow = B + p * n;
      for (int64_t j = 0; j < n; ++j) crow[j] += aip * brow[j];
    }
  }
}

/* Error-free transformations (Knuth TwoSum, Dekker TwoProduct via splitting;
 * no FMA so the result does not depend on contraction). */
static void two_sum(double a, double b, double* s, double* e) {
  double x = a + b, z = x - a;
  *s = x; *e = (a - (x - z)) + (b - z);
}
static void split(double a, double* hi, double* lo) {
  double c = 134217729.0 * a; /* 2^27 + 1 */
  double h = c - (c - a);
  *hi = h; *lo = a - h;
}
static void two_prod(double a, double b, double* p, double* e) {
  double x = a * b, ah, al, bh, bl;
  split(a, &ah, &al); split(b, &bh, &bl);
  *p = x; *e = al * bl - (((x - ah * bh) - al * bh) - ah * bl);
}

/* Dot2 (Ogita, Rump, Oishi 2005, Algorithm 5.3): sum_p A[i,p]*B[p,j] as if
 * computed in twice the working precision, then rounded. */
double orc_dot2_entry(int64_t n, int64_t k, const double* A, const double* B, int64_t i, int64_t j) {
  double s = 0.0, c = 0.0;
  for (int64_t p = 0; p < k; ++p) {
    double h, r, q;
    two_prod(A[i * k + p], B[p * n + j], &h, &r);
    two_sum(s, h, &s, &q);
    c += q + r;
  }
  return s + c;
}

/* |A||B| entry (for the componentwise GEMM error bound gamma_k (|A||B|)_ij). */
double orc_absdot_entry(int64_t n, int64_t k, const double* A, const double* B, int64_t i, int64_t j) {
  double s = 0.0;
  for (int64_t p = 0; p < k; ++p) s += fabs(A[i * k + p]) * fabs(B[p * n + j]);
  return s;
}

/* Normwise backward error per right-hand-side column:
 *   eta_j = ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf)
 * with the residual accumulated by Dot2 so the checker itself adds no error of
 * the order being measured. A n x n, X and B n x m, row-major. */
void orc_backward_error(int n, int m, const double* A, const double* B, const double* X, double* eta) {
  double anorm = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += fabs(A[(int64_t)i * n + j]);
    if (s > anorm) anorm = s;
  }
  for (int c = 0; c < m; ++c) {
    double rmax = 0.0, xmax = 0.0, bmax = 0.0;
    for (int i = 0; i < n; ++i) {
      double s = B[(int64_t)i * m + c], e = 0.0;
      for (int p = 0; p < n; ++p) {
        double h, r, q;
        two_prod(-A[(int64_t)i * n + p], X[(int64_t)p * m + c], &h, &r);
        two_sum(s, h, &s, &q);
        e += q + r;
      }
      double ri = fabs(s + e);
      if (ri > rmax) rmax = ri;
      if (fabs(X[(int64_t)i * m + c]) > xmax) xmax = fabs(X[(int64_t)i * m + c]);
      if (fabs(B[(int64_t)i * m + c]) > bmax) bmax = fabs(B[(int64_t)i * m + c]);
    }
    double den = anorm * xmax + bmax;
    eta[c] = den > 0.0 ? rmax / den : (rmax > 0.0 ? INFINITY : 0.0);
  }
}
