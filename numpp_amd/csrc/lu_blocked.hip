// Recursive LU with partial pivoting for a uniform stack of matrices (n > 64),
// in working precision T (double for the plain path, float for the
// mixed-precision path).
//
// getrf_rec(k0, w) factors the column range [k0, k0+w) over rows [k0, n):
//   w <= 64: one leaf launch (leaf_kernel);
//   else, with w1 = w/2 (a multiple of 64):
//     getrf_rec(k0, w1)
//     apply the left half's row moves to columns [k0+w1, k0+w)
//     trsm_rec: A12 <- L11^-1 A12            (recursive: leaf trsm + GEMM)
//     GEMM:     A22 <- A22 - A21 A12          (MFMA, K = w1)
//     getrf_rec(k0+w1, w-w1)
//     apply the right half's row moves to columns [k0, k0+w1)
// so most flops run in GEMMs with a large K (the top level is n/2 x n/2 x n/2)
// and the trailing matrix is read O(log n) times instead of O(n/64).
//
// The leaf factors an (n-k0) x w panel with one 512-thread workgroup per
// matrix, left-looking in inner blocks of 8 to 32 columns: the block's columns are
// gathered through the leaf's row permutation (rows are not moved until the
// block is done), brought up to date with the block's U rows (one wave per
// column solves with L11 via v_readlane) and the L columns already computed,
// then factored in registers with one workgroup barrier per column: each wave
// offers its local pivot candidate together with that row's values, so every
// thread picks the pivot (NumPP's rule: largest |a|, then smallest current
// position, linalg.cpp:55-56) and reads the pivot row in the same step. The
// per-step pivot rows (ipiv), the permutation (perm), the first exact-zero
// pivot (info) and the leaf's net row moves for the other columns are written.
// The rank-1 updates of linalg.cpp:64-68 are regrouped into these blocks; each
// entry receives the same terms in a different association order.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

namespace {

constexpr int LEAF = 64;   // leaf panel width
constexpr int PT = 512;    // leaf threads
constexpr int NW = PT / 64;
constexpr int MVLEAF = 2 * LEAF;  // max net moves of one leaf
constexpr int LSTRIDE = 1 + 2 * MVLEAF;

template <class T>
struct LeafArgs {
  T* A;
  int64_t strideA;
  int n, k0, w;
  int32_t* ipiv;
  int32_t* perm;
  int32_t* info;
  int32_t* mv;        // this leaf's list for matrix 0; matrix i at mv + i * mv_stride
  int64_t mv_stride;
  const int32_t* mask;
};

// NTH threads (512, or 256 for panels of at most 256 rows).
template <class T, int R, int IB, bool FULL, int NTH>
__global__ __launch_bounds__(NTH) void leaf_kernel(LeafArgs<T> a) {
  constexpr int NWL = NTH / 64;
  const int64_t mat = blockIdx.x;
  if (a.mask && a.mask[mat] == 0) return;
  T* __restrict__ A = a.A + mat * a.strideA;
  const int n = a.n, k0 = a.k0, w = a.w;
  const int m = n - k0;
  int32_t* ipiv = a.ipiv + mat * n;
  int32_t* perm = a.perm + mat * n;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  constexpr int QC = (R >= 4 && sizeof(T) == 8) ? 4 : 8;  // L columns per update chunk
  static_assert(IB % QC == 0 && LEAF % IB == 0, "inner blocks must hold whole update chunks");

  // rows/columns left of a block: at most LEAF - IB (none when IB = LEAF)
  constexpr int LEFT = LEAF - IB > 0 ? LEAF - IB : 1;
  // L rows of the blocks before the last, for the later blocks' U solves
  __shared__ T L11[LEFT][LEFT + 1];
  __shared__ T Ublk[LEFT][IB];
  __shared__ T red_v[2][NWL];  // candidate magnitudes (exact in T)
  __shared__ int red_p[2][NWL];
  __shared__ T red_row[2][NWL][IB];
  __shared__ int mv_old[2 * IB], mv_new[2 * IB];
  __shared__ int mv_cnt;
  __shared__ T stage[2 * IB][LEFT];
  __shared__ int stage_perm[2 * IB], stage_pid[2 * IB];
  __shared__ int sing_s;
  extern __shared__ int pid[];  // m entries: leaf-start row of the row now at each position

  for (int i = t; i < m; i += NTH) pid[i] = i;
  if (t == 0) sing_s = a.info[mat];
  __syncthreads();

  for (int jb = 0; jb < w; jb += IB) {
    const int ib = FULL ? IB : min(IB, w - jb);  // FULL: w is a multiple of IB
    const int cb = k0 + jb;  // first global column of the block
    T v[R][IB];
    int pos[R];
    // ---- gather the block's columns (lazy: rows still at leaf-start places)
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int p0 = jb + t + NTH * s;
      pos[s] = p0 < m ? p0 : (1 << 28);
      const T* src = A + (int64_t)(k0 + (p0 < m ? pid[p0] : 0)) * n + cb;
#pragma unroll
      // the row index is clamped, so the loads need no row guard; padding
      // slots (pos >= m) are excluded from the pivot search and every write
      for (int c = 0; c < IB; ++c) v[s][c] = (FULL || c < ib) ? src[c] : T(0);
    }
    {
      // U rows of the block (loads of every pass issued before any use)
      constexpr int UIT = ((LEAF - IB) * IB + NTH - 1) / NTH > 0 ? ((LEAF - IB) * IB + NTH - 1) / NTH : 1;
      T u[UIT];
#pragma unroll
      for (int i = 0; i < UIT; ++i) {
        const int e = t + i * NTH, q = e / IB, c = e % IB;
        const bool in = e < jb * IB && c < ib;
        u[i] = in ? A[(int64_t)(k0 + pid[in ? q : 0]) * n + cb + c] : T(0);
      }
#pragma unroll
      for (int i = 0; i < UIT; ++i) {
        const int e = t + i * NTH;
        if (e < jb * IB) Ublk[e / IB][e % IB] = u[i];
      }
    }
    __syncthreads();
    // ---- U rows [0, jb) of the block: wave w solves columns w, w+8, ...
    // (lane = row)
    if (jb > 0) {
      for (int cc = wave; cc < ib; cc += NWL) {
        T x = lane < jb ? Ublk[lane][cc] : T(0);
        for (int p = 0; p < jb - 1; ++p) {
          const T xp = readlane<T>(x, p);
          if (lane > p && lane < jb) x = x - L11[lane][p] * xp;
        }
        if (lane < jb) {
          Ublk[lane][cc] = x;
          A[(int64_t)(k0 + lane) * n + cb + cc] = x;
        }
      }
    }
    __syncthreads();
    // ---- left-looking update with the L columns [k0, cb) of each row, in
    // chunks of QC columns; the next chunk's loads are in flight while the
    // current one is applied
    if (jb > 0) {
      auto loadc = [&](T (&dst)[R][QC], int q0) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const T* lr = A + (int64_t)(k0 + (pos[s] < m ? pos[s] : 0)) * n + k0 + q0;
#pragma unroll
          // jb is a multiple of IB and of QC, so a chunk is all inside; rows
          // past m (clamped to row 0) only feed padding slots, which are
          // never read back
          for (int q = 0; q < QC; ++q) dst[s][q] = lr[q];
        }
      };
      auto applyc = [&](const T (&src)[R][QC], int q0) {
#pragma unroll
        for (int q = 0; q < QC; ++q) {
          if (q0 + q >= jb) break;
#pragma unroll
          for (int c = 0; c < IB; ++c) {
            const T u = Ublk[q0 + q][c];
#pragma unroll
            for (int s = 0; s < R; ++s) v[s][c] = v[s][c] - src[s][q] * u;
          }
        }
      };
      T la[R][QC], lb[R][QC];
      loadc(la, 0);
      for (int q0 = 0; q0 < jb; q0 += 2 * QC) {
        if (q0 + QC < jb) loadc(lb, q0 + QC);
        applyc(la, q0);
        if (q0 + 2 * QC < jb) loadc(la, q0 + 2 * QC);
        if (q0 + QC < jb) applyc(lb, q0 + QC);
      }
    }
    // ---- factor the block's columns, one barrier per column
    int opos[R];
#pragma unroll
    for (int s = 0; s < R; ++s) opos[s] = pos[s];
    // (compile-time column index: one copy of the step per column, so the
    // register rows are addressed statically)
    auto fstep = [&]<int c>() __attribute__((always_inline)) {
      if (c >= ib) return;
      const int col = jb + c;  // current leaf-relative step
      const int buf = c & 1;
      T bv = T(-1);
      int bp = 1 << 28, bs = -1;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] >= col && pos[s] < m) {
          const T av = fabs(v[s][c]);
          if (piv_better(av, pos[s], bv, bp)) { bv = av; bp = pos[s]; bs = s; }
        }
      }
      T wv = bv;
      int wp = bp;
      wave_argmax(wv, wp);
      if (bs >= 0 && wp == bp) {
        // this lane owns the wave's candidate row: offer it with its values
        red_v[buf][wave] = wv;
        red_p[buf][wave] = wp;
#pragma unroll
        for (int s = 0; s < R; ++s)
          if (s == bs) {
#pragma unroll
            for (int j = c; j < IB; ++j) red_row[buf][wave][j] = v[s][j];
          }
      } else if (lane == 0 && wv < T(0)) {
        red_v[buf][wave] = T(-1);
        red_p[buf][wave] = 1 << 28;
      }
      __syncthreads();
      int wbest = 0;
      bv = red_v[buf][0];
      bp = red_p[buf][0];
#pragma unroll
      for (int q = 1; q < NWL; ++q) {
        const T ov = red_v[buf][q];
        const int op = red_p[buf][q];
        if (piv_better(ov, op, bv, bp)) { bv = ov; bp = op; wbest = q; }
      }
      const int p = bp;
      if (bv == T(0) && t == 0 && sing_s == 0) sing_s = k0 + col + 1;
      if (t == 0) ipiv[k0 + col] = k0 + p;
      const T piv = red_row[buf][wbest][c];
      // FP32 factors (refined afterwards) scale by one reciprocal per column;
      // FP64 factors divide, as linalg.cpp:65
      const T rpiv = sizeof(T) == 4 && piv != T(0) ? T(1) / piv : T(0);
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] == p) pos[s] = col;
        else if (pos[s] == col) pos[s] = p;
      }
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] > col && pos[s] < m) {
          const T l = sizeof(T) == 4 ? v[s][c] * rpiv : (piv != T(0) ? v[s][c] / piv : T(0));
          v[s][c] = l;
#pragma unroll
          for (int j = c + 1; j < IB; ++j) v[s][j] = v[s][j] - l * red_row[buf][wbest][j];
        }
      }
    };
    [&]<int... C>(std::integer_sequence<int, C...>) __attribute__((always_inline)) {
      (fstep.template operator()<C>(), ...);
    }(std::make_integer_sequence<int, IB>{});
    // ---- L11 rows of the block's new pivot rows: block columns from
    // registers, left columns (from their pre-block places) loaded by all
    // threads in parallel
    // (only blocks before the last: rows and columns below LEFT)
    const bool need_l11 = jb + ib < w;
    if (t == 0) mv_cnt = 0;
    if (need_l11) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] >= jb && pos[s] < jb + ib) {
          stage_pid[pos[s] - jb] = opos[s];
#pragma unroll
          for (int c = 0; c < IB; ++c)
            if (c < ib) L11[pos[s]][jb + c] = v[s][c];
        }
      }
    }
    __syncthreads();
    if (need_l11 && jb > 0) {
      constexpr int LIT = (IB * (LEAF - IB) + NTH - 1) / NTH > 0 ? (IB * (LEAF - IB) + NTH - 1) / NTH : 1;
      T l[LIT];
#pragma unroll
      for (int i = 0; i < LIT; ++i) {
        const int e = t + i * NTH, rr = e / jb, q = e - rr * jb;
        l[i] = e < ib * jb ? A[(int64_t)(k0 + stage_pid[rr]) * n + k0 + q] : T(0);
      }
#pragma unroll
      for (int i = 0; i < LIT; ++i) {
        const int e = t + i * NTH, rr = e / jb, q = e - rr * jb;
        if (e < ib * jb) L11[jb + rr][q] = l[i];
      }
    }
    __syncthreads();
    // ---- write the block at final places; list rows whose place changed
#pragma unroll
    for (int s = 0; s < R; ++s) {
      if (pos[s] < m) {
        T* row = A + (int64_t)(k0 + pos[s]) * n + cb;
#pragma unroll
        for (int c = 0; c < IB; ++c)
          if (c < ib) row[c] = v[s][c];
        if (pos[s] != opos[s]) {
          const int e = atomicAdd(&mv_cnt, 1);
          mv_old[e] = opos[s];
          mv_new[e] = pos[s];
        }
      }
    }
    __syncthreads();
    // ---- move the up-to-date left columns, perm and pid of those rows
    const int cnt = mv_cnt;
    const int oc = jb;  // left columns [k0, cb)
    constexpr int MIT = (2 * IB * (LEAF - IB + 2) + NTH - 1) / NTH;
    const int tot = cnt * (oc + 2);
    if (tot > 0) {
      T mvv[MIT];
      int mvi[MIT];
#pragma unroll
      for (int i = 0; i < MIT; ++i) {
        const int e = t + i * NTH, q = e / (oc + 2), j = e - q * (oc + 2);
        const int src = e < tot ? mv_old[q] : 0;
        mvv[i] = (e < tot && j < oc) ? A[(int64_t)(k0 + src) * n + k0 + j] : T(0);
        mvi[i] = e < tot ? (j == oc ? perm[k0 + src] : (j > oc ? pid[src] : 0)) : 0;
      }
#pragma unroll
      for (int i = 0; i < MIT; ++i) {
        const int e = t + i * NTH, q = e / (oc + 2), j = e - q * (oc + 2);
        if (e < tot) {
          if (j < oc) stage[q][j] = mvv[i];
          else if (j == oc) stage_perm[q] = mvi[i];
          else stage_pid[q] = mvi[i];
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < MIT; ++i) {
        const int e = t + i * NTH, q = e / (oc + 2), j = e - q * (oc + 2);
        if (e < tot) {
          const int dst = mv_new[q];
          if (j < oc) A[(int64_t)(k0 + dst) * n + k0 + j] = stage[q][j];
          else if (j == oc) perm[k0 + dst] = stage_perm[q];
          else pid[dst] = stage_pid[q];
        }
      }
    }
    __syncthreads();
  }

  // Net row moves of this leaf for the columns outside it.
  if (t == 0) { mv_cnt = 0; a.info[mat] = sing_s; }
  __syncthreads();
  int32_t* mv = a.mv + mat * a.mv_stride;
  for (int i = t; i < m; i += NTH) {
    const int src = pid[i];
    if (src != i) {
      const int e = atomicAdd(&mv_cnt, 1);
      mv[1 + 2 * e] = k0 + src;
      mv[2 + 2 * e] = k0 + i;
    }
  }
  __syncthreads();
  if (t == 0) mv[0] = mv_cnt;
}

// Composes the move lists of leaves [l0, l1) (applied in order) into one net
// list of (src, dst) row pairs over rows [klo, n). One workgroup per matrix.
__global__ __launch_bounds__(256) void compose_kernel(int n, int klo, const int32_t* leafmv, int nleaf, int l0, int l1,
                                                      int32_t* out, int64_t out_stride, const int32_t* mask) {
  const int64_t mat = blockIdx.x;
  if (mask && mask[mat] == 0) return;
  extern __shared__ int cp[];  // n - klo entries
  __shared__ int tmp[MVLEAF];
  __shared__ int cnt_s;
  const int m = n - klo, t = threadIdx.x;
  for (int i = t; i < m; i += 256) cp[i] = i;
  __syncthreads();
  for (int l = l0; l < l1; ++l) {
    const int32_t* mv = leafmv + (mat * nleaf + l) * (int64_t)LSTRIDE;
    const int cnt = mv[0];
    for (int e = t; e < cnt; e += 256) tmp[e] = cp[mv[1 + 2 * e] - klo];
    __syncthreads();
    for (int e = t; e < cnt; e += 256) cp[mv[2 + 2 * e] - klo] = tmp[e];
    __syncthreads();
  }
  if (t == 0) cnt_s = 0;
  __syncthreads();
  int32_t* o = out + mat * out_stride;
  for (int i = t; i < m; i += 256) {
    if (cp[i] != i) {
      const int e = atomicAdd(&cnt_s, 1);
      o[1 + 2 * e] = klo + cp[i];
      o[2 + 2 * e] = klo + i;
    }
  }
  __syncthreads();
  if (t == 0) o[0] = cnt_s;
}

// Row moves for columns [c0, c0+nc): phase 0 copies the source rows into
// scratch, phase 1 writes them to their destinations. 256 columns x 16 moves
// per workgroup.
template <class T>
__global__ __launch_bounds__(256) void move_rows_kernel(T* Abase, int64_t strideA, int n, int c0, int nc,
                                                        const int32_t* list, int64_t list_stride, T* scratch,
                                                        int64_t scr_stride, int phase, const int32_t* mask) {
  const int64_t mat = blockIdx.z;
  if (mask && mask[mat] == 0) return;
  const int32_t* mv = list + mat * list_stride;
  const int cnt = mv[0];
  const int e0 = blockIdx.y * 16;
  if (e0 >= cnt) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nc) return;
  T* A = Abase + mat * strideA;
  T* sc = scratch + mat * scr_stride;
  const int e1 = min(cnt, e0 + 16);
  T buf[16];
  if (phase == 0) {
#pragma unroll
    for (int q = 0; q < 16; ++q) buf[q] = A[(int64_t)mv[1 + 2 * min(e0 + q, e1 - 1)] * n + c0 + c];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (e0 + q < e1) sc[(int64_t)(e0 + q) * nc + c] = buf[q];
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) buf[q] = sc[(int64_t)min(e0 + q, e1 - 1) * nc + c];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (e0 + q < e1) A[(int64_t)mv[2 + 2 * (e0 + q)] * n + c0 + c] = buf[q];
  }
}

// Row moves for columns [c0, c0+nc) in one pass when at most FMV rows move:
// a workgroup owns 16 columns, stages every moved row's values for them in
// LDS (the move list first, so the value loads are independent), then writes
// them to their destinations. Columns are disjoint between workgroups, so no
// other workgroup reads what this one overwrites.
constexpr int FMC = 16;
template <class T>
constexpr int fused_max_moves() { return sizeof(T) == 4 ? 512 : 256; }
template <class T>
__global__ __launch_bounds__(256) void move_rows_fused_kernel(T* Abase, int64_t strideA, int n, int c0, int nc,
                                                              const int32_t* list, int64_t list_stride,
                                                              const int32_t* mask) {
  constexpr int FMV = fused_max_moves<T>();
  const int64_t mat = blockIdx.y;
  if (mask && mask[mat] == 0) return;
  const int32_t* mv = list + mat * list_stride;
  const int cnt = mv[0];
  if (cnt == 0) return;
  T* A = Abase + mat * strideA;
  __shared__ T buf[FMV][FMC];
  __shared__ int src[FMV], dst[FMV];
  const int t = threadIdx.x, cl = t % FMC, g = t / FMC;
  for (int e = t; e < cnt; e += 256) {
    src[e] = mv[1 + 2 * e];
    dst[e] = mv[2 + 2 * e];
  }
  __syncthreads();
  const int c = blockIdx.x * FMC + cl;
  const int cc = c0 + min(c, nc - 1);
  constexpr int G = 256 / FMC;  // move groups
  for (int e0 = g; e0 < cnt; e0 += 4 * G) {
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = A[(int64_t)src[min(e0 + u * G, cnt - 1)] * n + cc];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e0 + u * G < cnt) buf[e0 + u * G][cl] = v[u];
  }
  __syncthreads();
  if (c < nc)
    for (int e = g; e < cnt; e += G) A[(int64_t)dst[e] * n + c0 + c] = buf[e][cl];
}

// B <- L^-1 B for the unit-lower s x s block L at (r0, r0), s <= 64, and the
// columns [c0, c0+nc). One thread per column (TT threads share one copy of L,
// stored transposed in LDS), column-oriented substitution: step p subtracts
// L[q][p] x[p] from every later row q, so the dependent chain is one FMA per
// step rather than a q-term dot product per row. TT = 64, 128 or 256 threads,
// the narrowest that covers nc (idle threads would hold registers and LDS).
template <class T, int TT>
__global__ __launch_bounds__(TT) void trsm_leaf_kernel(T* Abase, int64_t strideA, int n, int r0, int s, int c0,
                                                       int nc, const int32_t* mask) {
  const int64_t mat = blockIdx.y;
  if (mask && mask[mat] == 0) return;
  T* A = Abase + mat * strideA;
  __shared__ T LT[LEAF][LEAF + 1];  // LT[p][q] = L[q][p]
  {
    // coalesced clamped loads of L (row i, column j), transposed into LDS
    constexpr int PER = LEAF * LEAF / TT;
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * TT, i = e / LEAF, j = e - i * LEAF;
      v[u] = A[(int64_t)(r0 + min(i, s - 1)) * n + r0 + min(j, s - 1)];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * TT, i = e / LEAF, j = e - i * LEAF;
      LT[j][i] = (i < s && j < s) ? v[u] : T(0);
    }
  }
  __syncthreads();
  const int c = blockIdx.x * TT + threadIdx.x;
  if (c >= nc) return;
  T* base = A + (int64_t)r0 * n + c0 + c;  // row q of this column at base + q * n
  T x[LEAF];
#pragma unroll
  for (int q = 0; q < LEAF; ++q) x[q] = base[(int64_t)min(q, s - 1) * n];
#pragma unroll
  for (int q = 0; q < LEAF; ++q) x[q] = q < s ? x[q] : T(0);
#pragma unroll
  for (int p = 0; p < LEAF - 1; ++p) {
    // An opaque zero that depends on a recent result: step p's L reads cannot
    // be hoisted far above it (hoisting all 2016 LDS reads spills the kernel).
    int dep = 0;
    if (p >= 1) asm volatile("v_mov_b32 %0, 0" : "=v"(dep) : "v"(x[sizeof(T) == 4 ? p - 1 : p]));
    const T* lp = &LT[p][0] + dep;
    const T xp = x[p];
#pragma unroll
    for (int q = p + 1; q < LEAF; ++q) x[q] = x[q] - lp[q] * xp;
  }
#pragma unroll
  for (int q = 0; q < LEAF; ++q)
    if (q < s) base[(int64_t)q * n] = x[q];
}

__global__ void init_factor_kernel(int64_t count, int n, int32_t* perm, int32_t* info) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count * n) perm[i] = (int32_t)(i % n);
  if (i < count) info[i] = 0;
}

// Recursion driver state for one getrf call.
template <class T>
struct Rec {
  int64_t count;
  int n;
  T* A;
  int64_t strideA;
  int32_t *ipiv, *perm, *info;
  const int32_t* mask;
  int bf16_terms;
  hipStream_t s;
  int32_t* leafmv;  // count x nleaf x LSTRIDE
  int nleaf;
  int32_t* net;     // count x (1 + 2n)
  T* scratch;       // count x scr_stride (row-move staging)
  int64_t scr_stride;

  static int half(int w) {
    int h = (w / 2 + LEAF - 1) / LEAF * LEAF;
    return h >= w ? w - LEAF : h;
  }

  // FULL instantiation when every inner block of the panel is IB wide (the
  // column guards then fold away and the block's stores are unconditional).
  template <class U, int R, int IB, int NTH = PT>
  void launch_leaf(dim3 grid, size_t shm, const LeafArgs<U>& la) {
    if (la.w % IB == 0) hipLaunchKernelGGL((leaf_kernel<U, R, IB, true, NTH>), grid, dim3(NTH), shm, s, la);
    else hipLaunchKernelGGL((leaf_kernel<U, R, IB, false, NTH>), grid, dim3(NTH), shm, s, la);
  }

  int leaf(int k0, int w) {
    LeafArgs<T> la{A, strideA, n, k0, w, ipiv, perm, info, leafmv + (int64_t)(k0 / LEAF) * LSTRIDE,
                   (int64_t)nleaf * LSTRIDE, mask};
    const int m = n - k0;
    const size_t shm = (size_t)m * sizeof(int);
    dim3 grid((unsigned)count);
    ProfScope ps(kProfPanel, (double)count * m * w * w, 2.0 * sizeof(T) * count * m * w, s);
    // inner block width: as wide as the registers allow (fewer block-level
    // global round trips per leaf)
    constexpr bool F = sizeof(T) == 4;
    // short panels: 256 threads (half of a 512-thread workgroup would hold
    // no rows, and the per-column barrier spans fewer waves); taller panels
    // measured faster with 512 threads and fewer rows per thread
    // Many matrices (at least four per CU): one or two waves per matrix with
    // one row per lane up to 128 rows and two up to 256; the per-column pivot
    // search, pivot-row exchange and barrier then cost one or two waves'
    // instructions instead of four, and the other matrices' waves hide the
    // column chain's latency.
    static const int thin_env = [] {
      const char* e = std::getenv("NUMPP_HIP_LEAF_THIN");
      return e ? std::atoi(e) : 1;
    }();
    const bool thin = thin_env != 0 && count >= 4 * (int64_t)std::max(cu_count(), 1);
    // (FP64 rows take twice the registers: between 128 and 256 rows it uses
    // 16-column inner blocks; measured.)
    if (thin && m <= 64) launch_leaf<T, 1, 64, 64>(grid, shm, la);
    else if (thin && m <= 128) launch_leaf<T, 1, 64, 128>(grid, shm, la);
    else if (thin && m <= 256) {
      if constexpr (F) launch_leaf<T, 2, 64, 128>(grid, shm, la);
      else launch_leaf<T, 2, 16, 128>(grid, shm, la);
    } else if (m <= PT / 2) launch_leaf<T, 1, 64, PT / 2>(grid, shm, la);
    else if (m <= PT) launch_leaf<T, 1, 64>(grid, shm, la);
    else if (m <= 2 * PT) launch_leaf<T, 2, F ? 64 : 16>(grid, shm, la);
    else if (m <= 4 * PT) launch_leaf<T, 4, F ? 32 : 8>(grid, shm, la);
    else launch_leaf<T, 8, F ? 8 : 4>(grid, shm, la);
    NPPH_HIP(hipGetLastError());
    return NPPH_OK;
  }

  // Apply the moves of the leaves covering steps [kl, kl+wl) to columns [c0, c0+nc).
  int apply_moves(int kl, int wl, int c0, int nc) {
    if (nc <= 0) return NPPH_OK;
    const int l0 = kl / LEAF, l1 = (kl + wl - 1) / LEAF + 1;
    const int m = n - kl;
    hipLaunchKernelGGL(compose_kernel, dim3((unsigned)count), dim3(256), (size_t)m * sizeof(int), s, n, kl, leafmv,
                       nleaf, l0, l1, net, (int64_t)(1 + 2 * n), mask);
    NPPH_HIP(hipGetLastError());
    const int maxmv = std::min(m, 2 * wl);
    ProfScope ps(kProfMoveTrsm, 0.0, 4.0 * sizeof(T) * count * (double)maxmv * nc, s);
    if (maxmv <= fused_max_moves<T>()) {
      hipLaunchKernelGGL(move_rows_fused_kernel<T>, dim3((unsigned)ceil_div(nc, FMC), (unsigned)count), dim3(256), 0,
                         s, A, strideA, n, c0, nc, net, (int64_t)(1 + 2 * n), mask);
      NPPH_HIP(hipGetLastError());
      return NPPH_OK;
    }
    if ((int64_t)maxmv * nc > scr_stride) { set_error("LU: row-move scratch too small"); return NPPH_ERR_INVALID; }
    dim3 grid((unsigned)ceil_div(nc, 256), (unsigned)ceil_div(maxmv, 16), (unsigned)count);
    for (int phase = 0; phase < 2; ++phase) {
      hipLaunchKernelGGL(move_rows_kernel<T>, grid, dim3(256), 0, s, A, strideA, n, c0, nc, net, (int64_t)(1 + 2 * n),
                         scratch, scr_stride, phase, mask);
      NPPH_HIP(hipGetLastError());
    }
    return NPPH_OK;
  }

  int gemm_update(int64_t M, int64_t N, int64_t K, T* a, T* b, T* c) {
    if (M <= 0 || N <= 0 || K <= 0) return NPPH_OK;
    if constexpr (sizeof(T) == 4) {
      if (bf16_terms > 0)
        return gemm_bf16split_update(M, N, K, (const float*)a, n, strideA, (const float*)b, n, strideA, (float*)c, n,
                                     strideA, count, bf16_terms, s);
    }
    GemmArgs<T> g{M, N, K, T(-1), T(1), a, n, strideA, b, n, strideA, c, n, strideA, 0, 0, 0};
    return gemm_launch<T>(g, count, s);
  }

  // X <- L^-1 X for the unit-lower block at (r0, r0) of size sz and columns [c0, c0+nc).
  int trsm(int r0, int sz, int c0, int nc) {
    if (nc <= 0) return NPPH_OK;
    if (sz <= LEAF) {
      ProfScope ps(kProfMoveTrsm, (double)count * sz * sz * nc, 2.0 * sizeof(T) * count * (double)sz * nc, s);
      if (nc <= 64)
        hipLaunchKernelGGL((trsm_leaf_kernel<T, 64>), dim3(1u, (unsigned)count), dim3(64), 0, s, A, strideA, n, r0,
                           sz, c0, nc, mask);
      else if (nc <= 128)
        hipLaunchKernelGGL((trsm_leaf_kernel<T, 128>), dim3(1u, (unsigned)count), dim3(128), 0, s, A, strideA, n,
                           r0, sz, c0, nc, mask);
      else
        hipLaunchKernelGGL((trsm_leaf_kernel<T, 256>), dim3((unsigned)ceil_div(nc, 256), (unsigned)count),
                           dim3(256), 0, s, A, strideA, n, r0, sz, c0, nc, mask);
      NPPH_HIP(hipGetLastError());
      return NPPH_OK;
    }
    const int s1 = half(sz);
    NPPH_TRY(trsm(r0, s1, c0, nc));
    NPPH_TRY(gemm_update(sz - s1, nc, s1, A + (int64_t)(r0 + s1) * n + r0, A + (int64_t)r0 * n + c0,
                         A + (int64_t)(r0 + s1) * n + c0));
    return trsm(r0 + s1, sz - s1, c0, nc);
  }

  int getrf(int k0, int w) {
    if (w <= LEAF) return leaf(k0, w);
    const int w1 = half(w), w2 = w - w1;
    NPPH_TRY(getrf(k0, w1));
    NPPH_TRY(apply_moves(k0, w1, k0 + w1, w2));
    NPPH_TRY(trsm(k0, w1, k0 + w1, w2));
    NPPH_TRY(gemm_update(n - k0 - w1, w2, w1, A + (int64_t)(k0 + w1) * n + k0, A + (int64_t)k0 * n + k0 + w1,
                         A + (int64_t)(k0 + w1) * n + k0 + w1));
    NPPH_TRY(getrf(k0 + w1, w2));
    return apply_moves(k0 + w1, w2, k0, w1);
  }
};


// Sub-batch streams: a stack is factored as up to kGroups sub-batches on
// separate streams (default priority), each with its own workspace set, so the
// latency-bound leaves of one group overlap the GEMMs of another.
constexpr int kGroups = 4;
struct GroupPool {
  std::once_flag once;
  std::mutex mu;
  int status = NPPH_ERR_HIP;
  hipStream_t st[kGroups] = {};
  WsSet* ws[kGroups] = {};
  hipEvent_t ev[kGroups + 1] = {};
};
GroupPool& group_pool() {
  static GroupPool p;
  std::call_once(p.once, [] {
    for (int g = 0; g < kGroups; ++g) {
      if (hipStreamCreateWithFlags(&p.st[g], hipStreamNonBlocking) != hipSuccess) return;
      p.ws[g] = ws_create(p.st[g]);
    }
    for (int g = 0; g <= kGroups; ++g)
      if (hipEventCreateWithFlags(&p.ev[g], hipEventDisableTiming) != hipSuccess) return;
    p.status = NPPH_OK;
  });
  return p;
}

template <class T>
int getrf_group(int64_t count, int n, T* A, int64_t strideA, int32_t* ipiv, int32_t* perm, int32_t* info,
                int bf16_terms, const int32_t* mask, hipStream_t s) {
  Rec<T> r;
  r.count = count;
  r.n = n;
  r.A = A;
  r.strideA = strideA;
  r.ipiv = ipiv;
  r.perm = perm;
  r.info = info;
  r.mask = mask;
  r.bf16_terms = bf16_terms;
  r.s = s;
  r.nleaf = (int)ceil_div(n, LEAF);
  // The largest row move is at the top level: min(n, 2 w1) rows of w2 columns
  // with w1 >= n/2, at most n * n/2 elements.
  r.scr_stride = (int64_t)n * (n / 2 + LEAF);
  NPPH_TRY(workspace(kWsMisc, sizeof(int32_t) * (size_t)count * r.nleaf * LSTRIDE, (void**)&r.leafmv));
  NPPH_TRY(workspace(kWsNet, sizeof(int32_t) * (size_t)count * (1 + 2 * (size_t)n), (void**)&r.net));
  NPPH_TRY(workspace(kWsMove, sizeof(T) * (size_t)count * r.scr_stride, (void**)&r.scratch));
  {
    const int64_t tot = count * n;
    hipLaunchKernelGGL(init_factor_kernel, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, s, count, n, perm, info);
    NPPH_HIP(hipGetLastError());
  }
  return r.getrf(0, n);
}

}  // namespace

int max_blocked_n() { return 8 * PT; }

namespace {

// Sub-batches per factorisation (NUMPP_HIP_LU_GROUPS overrides).
int groups_for(int64_t count, int n, bool f32) {
  static const int env = [] {
    const char* e = std::getenv("NUMPP_HIP_LU_GROUPS");
    return e ? std::atoi(e) : -1;
  }();
  // measured on 64 x 2048^2: FP32 factorisations gain nothing from a second
  // sub-batch (its leaves and updates already fill the device between them),
  // FP64 ones gain about 1%
  int G = env > 0 ? env : ((count >= 4 && !f32) ? 2 : 1);
  if (n < 256) G = 1;
  return std::max(1, std::min<int>(G, (int)std::min<int64_t>(count, kGroups)));
}

// Issues the whole factorisation on s, forked over G sub-batch streams.
template <class T>
int issue_factor(int G, int64_t count, int n, T* A, int64_t strideA, int32_t* ipiv, int32_t* perm, int32_t* info,
                 int bf16_terms, const int32_t* mask, hipStream_t s) {
  if (G == 1) return getrf_group<T>(count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
  GroupPool& P = group_pool();
  if (P.status != NPPH_OK) return getrf_group<T>(count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
  NPPH_HIP(hipEventRecord(P.ev[kGroups], s));
  int st = NPPH_OK;
  for (int g = 0; g < G; ++g) {
    const int64_t c0 = count * g / G, c1 = count * (g + 1) / G;
    NPPH_HIP(hipStreamWaitEvent(P.st[g], P.ev[kGroups], 0));
    ws_bind(P.ws[g]);
    if (st == NPPH_OK)
      st = getrf_group<T>(c1 - c0, n, A + c0 * strideA, strideA, ipiv + c0 * n, perm + c0 * n, info + c0, bf16_terms,
                          mask ? mask + c0 : nullptr, P.st[g]);
    ws_bind(nullptr);
    NPPH_HIP(hipEventRecord(P.ev[g], P.st[g]));
    NPPH_HIP(hipStreamWaitEvent(s, P.ev[g], 0));
  }
  return st;
}

// Cached hipGraphs of whole factorisations, keyed by shape, pointers and the
// workspace generation. A new key runs eagerly once (sizing every workspace);
// its second call is captured and every later call replays the graph.
struct GraphKey {
  int tsize;
  int64_t count;
  int n, bf16, G;
  const void *A, *ipiv, *perm, *info, *mask;
  int64_t strideA;
  uint64_t gen;
  hipStream_t s;
  bool operator==(const GraphKey& o) const {
    return tsize == o.tsize && count == o.count && n == o.n && bf16 == o.bf16 && G == o.G && A == o.A &&
           ipiv == o.ipiv && perm == o.perm && info == o.info && mask == o.mask && strideA == o.strideA &&
           gen == o.gen && s == o.s;
  }
};
struct GraphEntry {
  GraphKey key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};
std::mutex g_graph_mu;
std::vector<GraphEntry> g_graphs;
std::vector<GraphKey> g_seen;
constexpr size_t kMaxGraphs = 8;

bool graphs_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("NUMPP_HIP_LU_GRAPHS");
    return !(e && e[0] == '0');
  }();
  return on;
}

}  // namespace

template <class T>
int blocked_getrf_perm(int64_t count, int n, T* A, int64_t strideA, int32_t* ipiv, int32_t* perm, int32_t* info,
                       int bf16_terms, const int32_t* mask, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  if (n > max_blocked_n()) { set_error("LU: n=%d > %d", n, max_blocked_n()); return NPPH_ERR_UNSUPPORTED; }
  if (count > 65535) { set_error("LU: count > 65535"); return NPPH_ERR_INVALID; }
  // A caller with its own bound workspace (a serving worker), an outer capture
  // or a caller that is about to capture gets a plain issue on its stream.
  if (ws_current() != nullptr || plain_issue() || stream_capturing(s))
    return getrf_group<T>(count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
  const int G = groups_for(count, n, sizeof(T) == 4);
  // Per-launch profiling needs an eager issue (a replayed graph has no
  // per-kernel events).
  if (!graphs_enabled() || prof_enabled())
    return issue_factor<T>(G, count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
  GraphKey key{(int)sizeof(T), count, n, bf16_terms, G, A, ipiv, perm, info, mask, strideA, workspace_generation(), s};
  std::lock_guard<std::mutex> lk(g_graph_mu);
  for (auto& e : g_graphs)
    if (e.key == key) {
      NPPH_HIP(hipGraphLaunch(e.exec, s));
      return NPPH_OK;
    }
  bool seen = false;
  for (auto& k : g_seen) seen = seen || (k == key);
  if (!seen) {
    const int st = issue_factor<T>(G, count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
    if (st == NPPH_OK) {
      key.gen = workspace_generation();  // the eager run may have grown the workspaces
      if (g_seen.size() >= 4 * kMaxGraphs) g_seen.erase(g_seen.begin());
      g_seen.push_back(key);
    }
    return st;
  }
  NPPH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int st = issue_factor<T>(G, count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (st != NPPH_OK || ec != hipSuccess || workspace_generation() != key.gen) {
    if (graph) (void)hipGraphDestroy(graph);
    if (st != NPPH_OK) return st;
    // capture failed: fall back to an eager issue of the same work
    return issue_factor<T>(G, count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
  }
  GraphEntry e;
  e.key = key;
  e.graph = graph;
  NPPH_HIP(hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0));
  if (g_graphs.size() >= kMaxGraphs) {
    (void)hipStreamSynchronize(s);
    (void)hipGraphExecDestroy(g_graphs.front().exec);
    (void)hipGraphDestroy(g_graphs.front().graph);
    g_graphs.erase(g_graphs.begin());
  }
  g_graphs.push_back(e);
  NPPH_HIP(hipGraphLaunch(e.exec, s));
  return NPPH_OK;
}

template int blocked_getrf_perm<double>(int64_t, int, double*, int64_t, int32_t*, int32_t*, int32_t*, int,
                                        const int32_t*, hipStream_t);
template int blocked_getrf_perm<float>(int64_t, int, float*, int64_t, int32_t*, int32_t*, int32_t*, int,
                                       const int32_t*, hipStream_t);

}  // namespace npph
