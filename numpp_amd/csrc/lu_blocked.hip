// Blocked right-looking LU with partial pivoting for a uniform stack of
// matrices (n > 64), in working precision T (double for the plain path, float
// for the mixed-precision path), plus the matching getrs.
//
// Per outer step k (panel width NB = 64):
//   1. panel kernel (one 512-thread workgroup per matrix): factors the
//      (n-k) x nb column panel in inner blocks of IB columns. The inner block's
//      rows live in registers (R rows x IB columns per thread), the pivot search
//      is a workgroup argmax over (|a|, position) with NumPP's tie rule
//      (linalg.cpp:55-56), rows are moved only by rewriting their logical
//      positions, and the block is written back at the permuted positions.
//      Row moves of the other panel columns and of the permutation vector are
//      staged through LDS (at most 2*IB rows move per inner block). The
//      remaining panel columns get the IB x rem triangular solve and rank-IB
//      update from the registers. The kernel also emits the panel's net row
//      move list (<= 2*nb (src, dst) pairs) for step 2.
//   2. move + trsm kernel: applies the row moves to every column outside the
//      panel (full-row swaps, linalg.cpp:58-60) and solves U12 = L11^-1 A12 for
//      the columns right of the panel, one column per thread.
//   3. trailing update A22 -= L21 U12 on MFMA (gemm.hip), or the split-BF16
//      MFMA update for the low-precision path.
// The rank-1 updates of linalg.cpp:64-68 are regrouped into these blocks; the
// arithmetic per entry is the same sum in a different association order.
#include "common.hpp"
#include "gemm.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int NB = 64;   // outer panel width
constexpr int IB = 8;    // inner register block width
constexpr int PT = 512;  // panel kernel threads
constexpr int MVMAX = 2 * NB;

template <class T>
struct PanelArgs {
  T* A;
  int64_t strideA;
  int n, k, nb;
  int32_t* ipiv;   // count x n
  int32_t* perm;   // count x n: original row at each position
  int32_t* info;   // count
  int32_t* mv;     // count x (1 + 2*MVMAX): [cnt, src0, dst0, src1, dst1, ...]
  const int32_t* mask;  // nullable: skip matrices with mask == 0
};

template <class T, int R>
__global__ __launch_bounds__(PT) void panel_kernel(PanelArgs<T> a) {
  const int64_t mat = blockIdx.x;
  if (a.mask && a.mask[mat] == 0) return;
  T* __restrict__ A = a.A + mat * a.strideA;
  const int n = a.n, k = a.k, nb = a.nb;
  int32_t* ipiv = a.ipiv + mat * n;
  int32_t* perm = a.perm + mat * n;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int NW = PT / 64;

  __shared__ double red_v[2][NW];
  __shared__ int red_p[2][NW];
  __shared__ T prow[2][IB];
  __shared__ T L11[IB][IB + 1];
  __shared__ T U12[IB][NB];
  __shared__ int mv_src[2 * IB], mv_dst[2 * IB];
  __shared__ int mv_cnt;
  __shared__ T stage[2 * IB][NB];
  __shared__ int stage_perm[2 * IB];
  __shared__ int stage_pid[2 * IB];
  __shared__ int sing_s;
  extern __shared__ int pid[];  // n - k entries: panel-start position of the row now at each position

  const int m_all = n - k;
  for (int i = t; i < m_all; i += PT) pid[i] = i;
  if (t == 0) sing_s = a.info[mat];
  __syncthreads();

  for (int jb = 0; jb < nb; jb += IB) {
    const int ib = min(IB, nb - jb);
    const int r0 = k + jb;
    const int m = n - r0;
    T v[R][IB];
    int pos[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int i = t + PT * s;
      pos[s] = i < m ? i : (1 << 28);
#pragma unroll
      for (int c = 0; c < IB; ++c)
        v[s][c] = (i < m && c < ib) ? A[(int64_t)(r0 + i) * n + r0 + c] : T(0);
    }

#pragma unroll
    for (int c = 0; c < IB; ++c) {
      if (c >= ib) break;
      const int buf = c & 1;
      double bv = -1.0;
      int bp = 1 << 28;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] >= c && pos[s] < m) {
          const double av = fabs((double)v[s][c]);
          if (piv_better(av, pos[s], bv, bp)) { bv = av; bp = pos[s]; }
        }
      }
      wave_argmax(bv, bp);
      if (lane == 0) { red_v[buf][wave] = bv; red_p[buf][wave] = bp; }
      __syncthreads();
      bv = red_v[buf][0];
      bp = red_p[buf][0];
#pragma unroll
      for (int w = 1; w < NW; ++w) {
        const double ov = red_v[buf][w];
        const int op = red_p[buf][w];
        if (piv_better(ov, op, bv, bp)) { bv = ov; bp = op; }
      }
      const int p = bp;
      if (bv == 0.0 && t == 0 && sing_s == 0) sing_s = r0 + c + 1;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] == p) {
#pragma unroll
          for (int j = 0; j < IB; ++j) prow[buf][j] = v[s][j];
          pos[s] = c;
        } else if (pos[s] == c) {
          pos[s] = p;
        }
      }
      if (t == 0) ipiv[r0 + c] = r0 + p;
      __syncthreads();
      const T piv = prow[buf][c];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] > c && pos[s] < m) {
          const T l = piv != T(0) ? v[s][c] / piv : T(0);
          v[s][c] = l;
#pragma unroll
          for (int j = 0; j < IB; ++j)
            if (j > c) v[s][j] = v[s][j] - l * prow[buf][j];
        }
      }
    }

    // Write the factored block back at the permuted row positions; collect the
    // rows that moved; stash L11.
    if (t == 0) mv_cnt = 0;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int i = t + PT * s;
      if (pos[s] < m) {
        T* row = A + (int64_t)(r0 + pos[s]) * n + r0;
#pragma unroll
        for (int c = 0; c < IB; ++c)
          if (c < ib) row[c] = v[s][c];
        if (pos[s] < ib) {
#pragma unroll
          for (int c = 0; c < IB; ++c) L11[pos[s]][c] = v[s][c];
        }
        if (pos[s] != i) {
          const int e = atomicAdd(&mv_cnt, 1);
          mv_src[e] = i;
          mv_dst[e] = pos[s];
        }
      }
    }
    __syncthreads();
    // Move the other panel columns, the permutation and the panel-row ids.
    const int cnt = mv_cnt;
    const int left = r0 - k;              // panel columns left of the block
    const int rem = k + nb - (r0 + ib);   // panel columns right of the block
    const int oc = left + rem;
    for (int e = t; e < cnt * (oc + 2); e += PT) {
      const int q = e / (oc + 2), j = e - q * (oc + 2);
      const int src = r0 + mv_src[q];
      if (j < oc) {
        const int gc = j < left ? k + j : r0 + ib + (j - left);
        stage[q][j] = A[(int64_t)src * n + gc];
      } else if (j == oc) {
        stage_perm[q] = perm[src];
      } else {
        stage_pid[q] = pid[src - k];
      }
    }
    __syncthreads();
    for (int e = t; e < cnt * (oc + 2); e += PT) {
      const int q = e / (oc + 2), j = e - q * (oc + 2);
      const int dst = r0 + mv_dst[q];
      if (j < oc) {
        const int gc = j < left ? k + j : r0 + ib + (j - left);
        A[(int64_t)dst * n + gc] = stage[q][j];
      } else if (j == oc) {
        perm[dst] = stage_perm[q];
      } else {
        pid[dst - k] = stage_pid[q];
      }
    }
    __syncthreads();

    if (rem > 0) {
      // U12 = L11^-1 * A12 (IB x rem), one column per thread.
      if (t < rem) {
        const int gc = r0 + ib + t;
        T u[IB];
#pragma unroll
        for (int q = 0; q < IB; ++q) u[q] = q < ib ? A[(int64_t)(r0 + q) * n + gc] : T(0);
#pragma unroll
        for (int q = 1; q < IB; ++q)
#pragma unroll
          for (int p2 = 0; p2 < IB; ++p2)
            if (p2 < q) u[q] = u[q] - L11[q][p2] * u[p2];
#pragma unroll
        for (int q = 0; q < IB; ++q) {
          if (q < ib) {
            A[(int64_t)(r0 + q) * n + gc] = u[q];
            U12[q][t] = u[q];
          }
        }
      }
      __syncthreads();
      // Rows below the block: A[row, right] -= L21[row, :] * U12.
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] >= ib && pos[s] < m) {
          T* row = A + (int64_t)(r0 + pos[s]) * n + r0 + ib;
          for (int j = 0; j < rem; ++j) {
            T acc = row[j];
#pragma unroll
            for (int q = 0; q < IB; ++q)
              if (q < ib) acc = acc - v[s][q] * U12[q][j];
            row[j] = acc;
          }
        }
      }
      __syncthreads();
    }
  }

  // Net row moves of this panel for the columns outside it.
  if (t == 0) { mv_cnt = 0; a.info[mat] = sing_s; }
  __syncthreads();
  int32_t* mv = a.mv + mat * (1 + 2 * MVMAX);
  for (int i = t; i < m_all; i += PT) {
    const int src = pid[i];
    if (src != i) {
      const int e = atomicAdd(&mv_cnt, 1);
      mv[1 + 2 * e] = k + src;
      mv[2 + 2 * e] = k + i;
    }
  }
  __syncthreads();
  if (t == 0) mv[0] = mv_cnt;
}

// Row moves for the columns outside the panel, then U12 = L11^-1 A12 for the
// columns right of it. One thread per column, 256 columns per workgroup. The
// moved rows of a column are read into a per-column scratch area first and
// written to their destinations afterwards (same thread, so no barrier).
constexpr int MT = 256;
template <class T>
__global__ __launch_bounds__(MT) void move_trsm_kernel(T* Abase, int64_t strideA, int n, int k, int nb,
                                                       const int32_t* mvbase, T* scratch,
                                                       const int32_t* mask) {
  const int64_t mat = blockIdx.y;
  if (mask && mask[mat] == 0) return;
  T* __restrict__ A = Abase + mat * strideA;
  const int32_t* mv = mvbase + mat * (1 + 2 * MVMAX);
  const int t = threadIdx.x;
  const int nleft = k, nright = n - k - nb;
  const int j = blockIdx.x * MT + t;
  __shared__ int src_s[MVMAX], dst_s[MVMAX];
  __shared__ T L11[NB][NB + 1];
  const int cnt = mv[0];
  for (int e = t; e < cnt; e += MT) { src_s[e] = mv[1 + 2 * e]; dst_s[e] = mv[2 + 2 * e]; }
  const bool any_right = (int)(blockIdx.x * MT + MT - 1) >= nleft && nright > 0;
  if (any_right) {
    for (int e = t; e < nb * nb; e += MT) {
      const int r = e / nb, c = e - r * nb;
      L11[r][c] = A[(int64_t)(k + r) * n + k + c];
    }
  }
  __syncthreads();
  if (j >= nleft + nright) return;
  const int gc = j < nleft ? j : k + nb + (j - nleft);
  T* sc = scratch + (mat * MVMAX) * (int64_t)n + gc;
  for (int e = 0; e < cnt; ++e) sc[(int64_t)e * n] = A[(int64_t)src_s[e] * n + gc];
  for (int e = 0; e < cnt; ++e) A[(int64_t)dst_s[e] * n + gc] = sc[(int64_t)e * n];
  if (j < nleft) return;
  T x[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) x[q] = q < nb ? A[(int64_t)(k + q) * n + gc] : T(0);
#pragma unroll
  for (int q = 1; q < NB; ++q) {
    T acc = x[q];
#pragma unroll
    for (int p = 0; p < NB; ++p)
      if (p < q) acc = acc - L11[q][p] * x[p];
    x[q] = acc;
  }
#pragma unroll
  for (int q = 0; q < NB; ++q)
    if (q < nb) A[(int64_t)(k + q) * n + gc] = x[q];
}

__global__ void init_factor_kernel(int64_t count, int n, int32_t* perm, int32_t* info) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count * n) perm[i] = (int32_t)(i % n);
  if (i < count) info[i] = 0;
}

template <class T>
int panel_launch(PanelArgs<T> a, int64_t count, hipStream_t s) {
  const int m = a.n - a.k;
  const size_t shm = (size_t)m * sizeof(int);
  dim3 grid((unsigned)count), block(PT);
  ProfScope ps(kProfPanel, (double)count * m * a.nb * a.nb, 2.0 * sizeof(T) * count * m * a.nb, s);
  if (m <= PT) hipLaunchKernelGGL((panel_kernel<T, 1>), grid, block, shm, s, a);
  else if (m <= 2 * PT) hipLaunchKernelGGL((panel_kernel<T, 2>), grid, block, shm, s, a);
  else if (m <= 4 * PT) hipLaunchKernelGGL((panel_kernel<T, 4>), grid, block, shm, s, a);
  else if (m <= 8 * PT) hipLaunchKernelGGL((panel_kernel<T, 8>), grid, block, shm, s, a);
  else { set_error("blocked LU: n=%d exceeds the 4096 panel limit", a.n); return NPPH_ERR_UNSUPPORTED; }
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace

int max_blocked_n() { return 8 * PT; }

template <class T>
int blocked_getrf_perm(int64_t count, int n, T* A, int64_t strideA, int32_t* ipiv, int32_t* perm,
                       int32_t* info, int bf16_terms, const int32_t* mask, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  if (n > max_blocked_n()) { set_error("blocked LU: n=%d > %d", n, max_blocked_n()); return NPPH_ERR_UNSUPPORTED; }
  if (count > 65535) { set_error("blocked LU: count > 65535"); return NPPH_ERR_INVALID; }
  int32_t* mv;
  T* scratch;
  NPPH_TRY(workspace(kWsMisc, sizeof(int32_t) * (size_t)count * (1 + 2 * MVMAX), (void**)&mv));
  NPPH_TRY(workspace(kWsMove, sizeof(T) * (size_t)count * MVMAX * (size_t)n, (void**)&scratch));
  {
    const int64_t tot = count * n;
    hipLaunchKernelGGL(init_factor_kernel, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, s, count, n, perm, info);
    NPPH_HIP(hipGetLastError());
  }
  for (int k = 0; k < n; k += NB) {
    const int nb = n - k < NB ? n - k : NB;
    PanelArgs<T> pa{A, strideA, n, k, nb, ipiv, perm, info, mv, mask};
    NPPH_TRY(panel_launch<T>(pa, count, s));
    const int nout = n - nb;  // columns outside the panel
    if (nout > 0) {
      dim3 grid((unsigned)ceil_div(nout, MT), (unsigned)count);
      ProfScope ps(kProfMoveTrsm, (double)count * nb * nb * (n - k - nb),
                   2.0 * sizeof(T) * count * ((double)nb * (n - k - nb) + 2.0 * nb * nout), s);
      hipLaunchKernelGGL(move_trsm_kernel<T>, grid, dim3(MT), 0, s, A, strideA, n, k, nb, mv, scratch, mask);
      NPPH_HIP(hipGetLastError());
    }
    const int64_t rest = n - k - nb;
    if (rest > 0) {
      T* a21 = A + (int64_t)(k + nb) * n + k;
      T* a12 = A + (int64_t)k * n + k + nb;
      T* a22 = A + (int64_t)(k + nb) * n + k + nb;
      if constexpr (sizeof(T) == 4) {
        if (bf16_terms > 0) {
          NPPH_TRY(gemm_bf16split_update(rest, rest, nb, (const float*)a21, n, strideA, (const float*)a12, n,
                                         strideA, (float*)a22, n, strideA, count, bf16_terms, s));
          continue;
        }
      }
      GemmArgs<T> g{rest, rest, nb, T(-1), T(1), a21, n, strideA, a12, n, strideA, a22, n, strideA, 0, 0, 0};
      NPPH_TRY(gemm_launch<T>(g, count, s));
    }
  }
  return NPPH_OK;
}

template int blocked_getrf_perm<double>(int64_t, int, double*, int64_t, int32_t*, int32_t*, int32_t*, int,
                                        const int32_t*, hipStream_t);
template int blocked_getrf_perm<float>(int64_t, int, float*, int64_t, int32_t*, int32_t*, int32_t*, int,
                                       const int32_t*, hipStream_t);

// ---------------------------------------------------------------------------
// getrs: X = U^-1 L^-1 P B with the packed factors and the row permutation.
// One 256-thread workgroup per (matrix, 8-column RHS chunk). Right-looking
// blocked substitution with 64-row diagonal blocks: the diagonal block is
// solved by one wave with its rows in registers (v_readlane broadcasts), then
// every thread applies that block's contribution to the rows still below
// (forward) or above (back). Order of operations follows lu_solve
// (linalg.cpp:73-90): gather by the permutation, unit-lower forward sweep,
// upper backward sweep with division by the diagonal.
namespace {

constexpr int GT = 256;
constexpr int GB = 64;
constexpr int GNR = 8;

template <class T>
__global__ __launch_bounds__(GT) void getrs_kernel(int n, int nrhs, const T* LUb, int64_t strideA,
                                                   const int32_t* permb, const T* Bb, int64_t strideB,
                                                   T* Xb, int64_t strideX, const int32_t* info,
                                                   const int32_t* mask) {
  const int64_t mat = blockIdx.x;
  if (info && info[mat] != 0) return;
  if (mask && mask[mat] == 0) return;
  const int c0 = blockIdx.y * GNR;
  const int nc = min(GNR, nrhs - c0);
  const T* __restrict__ LU = LUb + mat * strideA;
  const int32_t* perm = permb + mat * n;
  const T* __restrict__ B = Bb + mat * strideB;
  T* __restrict__ X = Xb + mat * strideX;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  __shared__ T xb[GB][GNR];

  // gather X = P B (this chunk's columns)
  for (int e = t; e < n * GNR; e += GT) {
    const int i = e / GNR, c = e - i * GNR;
    if (c < nc) X[(int64_t)i * nrhs + c0 + c] = B[(int64_t)perm[i] * nrhs + c0 + c];
  }
  __syncthreads();

  // forward: unit lower
  for (int r0 = 0; r0 < n; r0 += GB) {
    const int bs = min(GB, n - r0);
    if (wave == 0) {
      const int i = r0 + lane;
      T x[GNR];
#pragma unroll
      for (int c = 0; c < GNR; ++c) x[c] = (lane < bs && c < nc) ? X[(int64_t)i * nrhs + c0 + c] : T(0);
      T lrow[GB];
#pragma unroll
      for (int q = 0; q < GB; ++q) lrow[q] = (lane < bs && q < lane) ? LU[(int64_t)i * n + r0 + q] : T(0);
#pragma unroll
      for (int q = 0; q < GB; ++q) {
#pragma unroll
        for (int c = 0; c < GNR; ++c) x[c] = x[c] - lrow[q] * readlane<T>(x[c], q);
      }
      if (lane < bs) {
#pragma unroll
        for (int c = 0; c < GNR; ++c) {
          xb[lane][c] = x[c];
          if (c < nc) X[(int64_t)i * nrhs + c0 + c] = x[c];
        }
      } else {
#pragma unroll
        for (int c = 0; c < GNR; ++c) xb[lane][c] = T(0);
      }
    }
    __syncthreads();
    for (int i = r0 + bs + t; i < n; i += GT) {
      T acc[GNR];
#pragma unroll
      for (int c = 0; c < GNR; ++c) acc[c] = c < nc ? X[(int64_t)i * nrhs + c0 + c] : T(0);
      const T* lr = LU + (int64_t)i * n + r0;
      for (int q = 0; q < bs; ++q) {
        const T l = lr[q];
#pragma unroll
        for (int c = 0; c < GNR; ++c) acc[c] = acc[c] - l * xb[q][c];
      }
#pragma unroll
      for (int c = 0; c < GNR; ++c)
        if (c < nc) X[(int64_t)i * nrhs + c0 + c] = acc[c];
    }
    __syncthreads();
  }

  // back: upper with division by the diagonal
  const int nblk = (n + GB - 1) / GB;
  for (int b = nblk - 1; b >= 0; --b) {
    const int r0 = b * GB;
    const int bs = min(GB, n - r0);
    if (wave == 0) {
      const int i = r0 + lane;
      T x[GNR];
#pragma unroll
      for (int c = 0; c < GNR; ++c) x[c] = (lane < bs && c < nc) ? X[(int64_t)i * nrhs + c0 + c] : T(0);
      T urow[GB];
#pragma unroll
      for (int q = 0; q < GB; ++q) urow[q] = (lane < bs && q >= lane && q < bs) ? LU[(int64_t)i * n + r0 + q] : T(0);
      T d = T(1);
#pragma unroll
      for (int q = 0; q < GB; ++q) d = (q == lane && lane < bs) ? urow[q] : d;
#pragma unroll
      for (int q = GB - 1; q >= 0; --q) {
        if (lane == q) {
#pragma unroll
          for (int c = 0; c < GNR; ++c) x[c] = x[c] / d;
        }
        const T f = lane < q ? urow[q] : T(0);
#pragma unroll
        for (int c = 0; c < GNR; ++c) x[c] = x[c] - f * readlane<T>(x[c], q);
      }
      if (lane < bs) {
#pragma unroll
        for (int c = 0; c < GNR; ++c) {
          xb[lane][c] = x[c];
          if (c < nc) X[(int64_t)i * nrhs + c0 + c] = x[c];
        }
      } else {
#pragma unroll
        for (int c = 0; c < GNR; ++c) xb[lane][c] = T(0);
      }
    }
    __syncthreads();
    for (int i = t; i < r0; i += GT) {
      T acc[GNR];
#pragma unroll
      for (int c = 0; c < GNR; ++c) acc[c] = c < nc ? X[(int64_t)i * nrhs + c0 + c] : T(0);
      const T* ur = LU + (int64_t)i * n + r0;
      for (int q = 0; q < bs; ++q) {
        const T u = ur[q];
#pragma unroll
        for (int c = 0; c < GNR; ++c) acc[c] = acc[c] - u * xb[q][c];
      }
#pragma unroll
      for (int c = 0; c < GNR; ++c)
        if (c < nc) X[(int64_t)i * nrhs + c0 + c] = acc[c];
    }
    __syncthreads();
  }
}

// Permutation from LAPACK-style pivots: one wave per matrix replays the swaps
// in LDS (used only by the standalone getrs entry point).
__global__ __launch_bounds__(64) void ipiv_to_perm_kernel(int n, const int32_t* ipivb, int32_t* permb) {
  const int64_t mat = blockIdx.x;
  const int32_t* ipiv = ipivb + mat * n;
  int32_t* perm = permb + mat * n;
  extern __shared__ int p_s[];
  for (int i = threadIdx.x; i < n; i += 64) p_s[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < n; ++k) {
      const int p = ipiv[k];
      if (p != k && p >= 0 && p < n) { const int tmp = p_s[k]; p_s[k] = p_s[p]; p_s[p] = tmp; }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 64) perm[i] = p_s[i];
}

}  // namespace

template <class T>
int blocked_getrs_perm(int64_t count, int n, int nrhs, const T* LU, int64_t strideA, const int32_t* perm,
                       const T* B, int64_t strideB, T* X, int64_t strideX, const int32_t* info,
                       const int32_t* mask, hipStream_t s) {
  if (count <= 0 || n <= 0 || nrhs <= 0) return NPPH_OK;
  if (B == X) { set_error("getrs: B and X must not alias"); return NPPH_ERR_INVALID; }
  if (count > 0x7fffffff) { set_error("getrs: count too large"); return NPPH_ERR_INVALID; }
  dim3 grid((unsigned)count, (unsigned)ceil_div(nrhs, GNR));
  ProfScope ps(kProfGetrs, 2.0 * count * n * n * nrhs, (double)count * (sizeof(T) * (double)n * n + 3.0 * sizeof(T) * n * nrhs), s);
  hipLaunchKernelGGL(getrs_kernel<T>, grid, dim3(GT), 0, s, n, nrhs, LU, strideA, perm, B, strideB, X,
                     strideX, info, mask);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

template int blocked_getrs_perm<double>(int64_t, int, int, const double*, int64_t, const int32_t*, const double*,
                                        int64_t, double*, int64_t, const int32_t*, const int32_t*, hipStream_t);
template int blocked_getrs_perm<float>(int64_t, int, int, const float*, int64_t, const int32_t*, const float*,
                                       int64_t, float*, int64_t, const int32_t*, const int32_t*, hipStream_t);

int ipiv_to_perm(int64_t count, int n, const int32_t* ipiv, int32_t* perm, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  hipLaunchKernelGGL(ipiv_to_perm_kernel, dim3((unsigned)count), dim3(64), (size_t)n * sizeof(int), s, n, ipiv,
                     perm);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

}  // namespace npph
