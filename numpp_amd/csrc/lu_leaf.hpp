// Panel leaf of the recursive LU (see lu_blocked.hip): the kernel template and
// its launch helpers. The instantiations are spread over the lu_leaf_*.hip
// translation units (one per working precision and panel family) so that they
// compile in parallel.
//
// Rows never move: position p of the factorisation lives in physical row
// perm[p], and pivoting only exchanges perm entries. The leaf factors the
// (n-k0) x w panel at positions [k0, n) (w <= 64) with one workgroup per
// matrix, left-looking in inner blocks of IB columns: each thread gathers its
// rows' block columns from their physical rows, brings them up to date with the
// block's U rows (one wave per column solves with L11 via v_readlane) and the L
// columns already computed, then factors them in registers with one workgroup
// barrier per column: each wave offers its local pivot candidate together with
// that row's values, so every thread picks the pivot (NumPP's rule: largest
// |a|, then smallest current position, linalg.cpp:55-56, with its NaN
// behaviour) and reads the pivot row in the same step. The block is written
// back to the same physical rows. The per-step pivot rows (ipiv), the
// permutation (perm) and the first exact-zero pivot (info) are written. The
// rank-1 updates of linalg.cpp:64-68 are regrouped into these blocks; each
// entry receives the same terms in a different association order.
#pragma once

#include <utility>

#include "common.hpp"

namespace npph {

constexpr int LEAF = 64;   // leaf panel width
constexpr int PT = 512;    // leaf threads

template <class T>
struct LeafArgs {
  T* A;
  int64_t strideA;
  int n, k0, w;
  int32_t* ipiv;
  int32_t* perm;  // count x n: physical row of each position
  int32_t* info;
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
  const unsigned lt = threadIdx.x;  // the thread index, unsigned for cheap / and %
  constexpr int QC = (R >= 4 && sizeof(T) == 8) ? 4 : 8;  // L columns per update chunk
  static_assert(IB % QC == 0 && LEAF % IB == 0, "inner blocks must hold whole update chunks");

  // rows/columns left of a block: at most LEAF - IB (none when IB = LEAF)
  constexpr int LEFT = LEAF - IB > 0 ? LEAF - IB : 1;
  // L rows of the blocks before the last, for the later blocks' U solves
  __shared__ T L11[LEFT][LEFT + 1];
  __shared__ T Ublk[LEFT][IB];
  __shared__ T red_v[2][NWL];  // candidate magnitudes (exact in T)
  __shared__ int red_p[2][NWL];
  __shared__ __attribute__((aligned(16))) T red_row[2][NWL][IB];
  __shared__ int stage_row[IB];  // physical rows of the block's pivot rows
  // pivot rows of the leaf's steps, written to HBM once after the last block:
  // a global store inside the column loop would make every column barrier
  // wait for it (the barrier's release fence)
  __shared__ int piv_s[LEAF];
  // Staging of the block's rows for panels of 512 threads whose inner-block
  // rows are 128 bytes (FP32 IB = 32, FP64 IB = 16) or 256 bytes (FP32
  // IB = 64): the gather and the write-back move each row as 16-byte pieces
  // read by LPR neighbouring lanes (whole cache lines), passing through LDS in
  // 512-row chunks (72 or 136 KB) to and from the one-row-per-lane register
  // layout. With one row per lane each wave instruction would touch 64 cache
  // lines for 16 bytes apiece, which one CU's memory pipe serves at a third of
  // the rate. Rows are padded by 16 bytes so the lanes' 16-byte LDS reads
  // start 4 banks apart. A thread moves its LPR pieces of a chunk in batches
  // of eight, one batch in flight at a time.
  constexpr int RB = IB * (int)sizeof(T);   // bytes per inner-block row
  constexpr bool STG = NTH == 512 && (RB == 128 || (RB == 256 && sizeof(T) == 4));
  constexpr int EPV = 16 / (int)sizeof(T);  // elements per 16-byte piece
  constexpr int SW = IB + EPV;              // staging row stride (elements)
  constexpr int LPR = RB / 16;              // lanes per row (= pieces per thread)
  constexpr int NBT = STG ? LPR / 8 : 1;    // batches of eight pieces per chunk
  typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
  __shared__ __attribute__((aligned(16))) T stage[STG ? NTH : 1][STG ? SW : 1];
  __shared__ int stage_phys[STG ? NTH : 1];
  extern __shared__ int rowof[];  // m entries: physical row of each leaf position

  // 16-byte pieces need 16-byte aligned rows
  const bool staged = ((reinterpret_cast<uintptr_t>(a.A) | (uintptr_t)(a.strideA * sizeof(T)) |
                        (uintptr_t)(a.n * sizeof(T))) & 15) == 0;
  for (int i = t; i < m; i += NTH) rowof[i] = perm[k0 + i];
  // first exact-zero pivot step + 1 (0: none yet); every thread tracks it, as
  // every thread sees each step's pivot value
  int sing = a.info[mat];
  __syncthreads();

  for (int jb = 0; jb < w; jb += IB) {
    const int ib = FULL ? IB : min(IB, w - jb);  // FULL: w is a multiple of IB
    const int cb = k0 + jb;  // first global column of the block
    T v[R][IB];
    int pos[R], phys[R];  // current positions and (fixed) physical rows
    // ---- gather the block's columns from the rows' physical places
    // (the row index is clamped, so the loads need no row guard; padding
    // slots (pos >= m) are excluded from the pivot search and every write)
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int p0 = jb + t + NTH * s;
      pos[s] = p0 < m ? p0 : (1 << 28);
      phys[s] = rowof[p0 < m ? p0 : 0];
    }
    if (STG && FULL && staged) {
      // batch g = (chunk s, batch b): piece lt % LPR of the chunk's rows
      // lt / LPR + (NTH / LPR) q for q in 8 b .. 8 b + 7 (the next batch's
      // loads are in flight while the current one is stored to LDS and, after
      // a chunk's last batch, while the chunk is read back)
      vec_t piece[8];
      auto load_batch = [&](int g) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int pp = jb + NTH * (g / NBT) + lt / LPR + (NTH / LPR) * (q + 8 * (g % NBT));
          const int ph = rowof[pp < m ? pp : 0];
          piece[q] = *reinterpret_cast<const vec_t*>(A + (int64_t)ph * n + cb + (lt % LPR) * EPV);
        }
      };
      load_batch(0);
#pragma unroll
      for (int g = 0; g < R * NBT; ++g) {
        if (g > 0 && g % NBT == 0) __syncthreads();  // the previous chunk's rows are read
#pragma unroll
        for (int q = 0; q < 8; ++q)
          *reinterpret_cast<vec_t*>(&stage[lt / LPR + (NTH / LPR) * (q + 8 * (g % NBT))][(lt % LPR) * EPV]) =
              piece[q];
        if (g + 1 < R * NBT) load_batch(g + 1);
        if (g % NBT == NBT - 1) {
          __syncthreads();
          const int s = g / NBT;
#pragma unroll
          for (int c = 0; c < IB; c += EPV) {
            const vec_t x = *reinterpret_cast<const vec_t*>(&stage[t][c]);
#pragma unroll
            for (int e = 0; e < EPV; ++e) v[s][c + e] = x[e];
          }
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const T* src = A + (int64_t)phys[s] * n + cb;
#pragma unroll
        for (int c = 0; c < IB; ++c) v[s][c] = (FULL || c < ib) ? src[c] : T(0);
      }
    }
    {
      // U rows of the block (loads of every pass issued before any use)
      constexpr int UIT = ((LEAF - IB) * IB + NTH - 1) / NTH > 0 ? ((LEAF - IB) * IB + NTH - 1) / NTH : 1;
      T u[UIT];
#pragma unroll
      for (int i = 0; i < UIT; ++i) {
        const int e = t + i * NTH, q = e / IB, c = e % IB;
        const bool in = e < jb * IB && c < ib;
        u[i] = in ? A[(int64_t)rowof[in ? q : 0] * n + cb + c] : T(0);
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
          A[(int64_t)rowof[lane] * n + cb + cc] = x;
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
          const T* lr = A + (int64_t)phys[s] * n + k0 + q0;
#pragma unroll
          // jb is a multiple of IB and of QC, so a chunk is all inside; padding
          // slots read a real row and are never written back
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
    // (compile-time column index: one copy of the step per column, so the
    // register rows are addressed statically)
    auto fstep = [&]<int c>() __attribute__((always_inline)) {
      if (c >= ib) return;
      const int col = jb + c;  // current leaf-relative step
      const int buf = c & 1;
      T bv = T(-1);
      int bp = 1 << 28, bs = -1;
      // (selects, no branches: a row outside [col, m) has key -1, which
      // never wins, and neither does a NaN below position col)
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const bool el = pos[s] >= col && pos[s] < m;
        const T av = el ? piv_key(v[s][c], pos[s], col) : T(-1);
        const bool tk = av >= T(0) && piv_better(av, pos[s], bv, bp);
        bv = tk ? av : bv;
        bp = tk ? pos[s] : bp;
        bs = tk ? s : bs;
      }
      T wv = bv;
      int wp = bp;
      wave_argmax(wv, wp);
      if (bs >= 0 && wp == bp) {
        // this lane owns the wave's candidate row: offer it with its values
        // (from a 16-byte boundary, so the stores are 16-byte vectors; the
        // entries left of c are never read)
        constexpr int C0 = c & ~(16 / (int)sizeof(T) - 1);
        red_v[buf][wave] = wv;
        red_p[buf][wave] = wp;
#pragma unroll
        for (int s = 0; s < R; ++s)
          if (s == bs) {
#pragma unroll
            for (int j = C0; j < IB; ++j) red_row[buf][wave][j] = v[s][j];
          }
      } else if (lane == 0 && wv < T(0)) {
        red_v[buf][wave] = T(-1);
        red_p[buf][wave] = 1 << 28;
      }
      __syncthreads();
      // cross-wave pick: lane q < NWL reads wave q's candidate, then the same
      // wave argmax (the row at position col is a candidate of some wave, so
      // the winner is a real row and its wave is unique)
      const int cp = lane < NWL ? red_p[buf][lane < NWL ? lane : 0] : (1 << 29);
      bv = lane < NWL ? red_v[buf][lane < NWL ? lane : 0] : T(-2);
      bp = cp;
      static_assert(NWL <= 8, "the cross-wave pick reduces over lanes 0..7");
      lanes8_argmax(bv, bp);
      const int wbest = (int)__builtin_ctzll(__ballot(lane < NWL && cp == bp));
      const int p = bp;
      sing = (bv == T(0) && sing == 0) ? k0 + col + 1 : sing;
      if (t == 0) piv_s[col] = k0 + p;
      const T piv = red_row[buf][wbest][c];
      // FP32 factors (refined afterwards) scale by one reciprocal per column;
      // FP64 factors divide, as linalg.cpp:65
      // (FP32 factors are refined afterwards: the hardware reciprocal, within
      // one ulp, instead of the correctly rounded division)
      T rpiv = T(0);
      if constexpr (sizeof(T) == 4) rpiv = piv != T(0) ? __builtin_amdgcn_rcpf(piv) : T(0);
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] == p) pos[s] = col;
        else if (pos[s] == col) pos[s] = p;
      }
      // the pivot row's entries right of c, read once for all R rows (from a
      // 16-byte boundary: vector loads)
      constexpr int C1 = (c + 1) & ~(16 / (int)sizeof(T) - 1);
      T urow[IB - C1 > 0 ? IB - C1 : 1];
#pragma unroll
      for (int j = C1; j < IB; ++j) urow[j - C1] = red_row[buf][wbest][j];
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] > col && pos[s] < m) {
          const T l = sizeof(T) == 4 ? v[s][c] * rpiv : (piv != T(0) ? v[s][c] / piv : T(0));
          v[s][c] = l;
#pragma unroll
          for (int j = c + 1; j < IB; ++j) v[s][j] = v[s][j] - l * urow[j - C1];
        }
      }
    };
    [&]<int... C>(std::integer_sequence<int, C...>) __attribute__((always_inline)) {
      (fstep.template operator()<C>(), ...);
    }(std::make_integer_sequence<int, IB>{});
    // ---- L11 rows of the block's new pivot rows: block columns from
    // registers, left columns (from the rows' physical places) loaded by all
    // threads in parallel
    // (only blocks before the last: rows and columns below LEFT)
    const bool need_l11 = jb + ib < w;
    if (need_l11) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] >= jb && pos[s] < jb + ib) {
          stage_row[pos[s] - jb] = phys[s];
#pragma unroll
          for (int c = 0; c < IB; ++c)
            if (c < ib) L11[pos[s]][jb + c] = v[s][c];
        }
      }
    }
    // ---- the block back to its rows' physical places
    if (STG && FULL && staged) {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (s > 0) __syncthreads();  // the previous chunk's pieces are stored
        stage_phys[t] = pos[s] < m ? phys[s] : -1;
#pragma unroll
        for (int c = 0; c < IB; c += EPV) {
          vec_t x;
#pragma unroll
          for (int e = 0; e < EPV; ++e) x[e] = v[s][c + e];
          *reinterpret_cast<vec_t*>(&stage[t][c]) = x;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LPR; ++q) {
          const int lr = lt / LPR + (NTH / LPR) * q;
          const int ph = stage_phys[lr];
          if (ph >= 0)
            *reinterpret_cast<vec_t*>(A + (int64_t)ph * n + cb + (lt % LPR) * EPV) =
                *reinterpret_cast<const vec_t*>(&stage[lr][(lt % LPR) * EPV]);
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (pos[s] < m) {
          T* row = A + (int64_t)phys[s] * n + cb;
#pragma unroll
          for (int c = 0; c < IB; ++c)
            if (c < ib) row[c] = v[s][c];
        }
      }
    }
    __syncthreads();  // every read of rowof (and of the staging rows) for this block is done
#pragma unroll
    for (int s = 0; s < R; ++s)
      if (pos[s] < m) rowof[pos[s]] = phys[s];
    if (need_l11 && jb > 0) {
      constexpr int LIT = (IB * (LEAF - IB) + NTH - 1) / NTH > 0 ? (IB * (LEAF - IB) + NTH - 1) / NTH : 1;
      T l[LIT];
#pragma unroll
      for (int i = 0; i < LIT; ++i) {
        const int e = t + i * NTH, rr = e / jb, q = e - rr * jb;
        l[i] = e < ib * jb ? A[(int64_t)stage_row[rr] * n + k0 + q] : T(0);
      }
#pragma unroll
      for (int i = 0; i < LIT; ++i) {
        const int e = t + i * NTH, rr = e / jb, q = e - rr * jb;
        if (e < ib * jb) L11[jb + rr][q] = l[i];
      }
    }
    __syncthreads();
  }

  // pivots, first exact-zero pivot and the positions' physical rows
  if (t == 0) a.info[mat] = sing;
  for (int c = t; c < w; c += NTH) ipiv[k0 + c] = piv_s[c];
  for (int i = t; i < m; i += NTH) perm[k0 + i] = rowof[i];
}


// FULL instantiation when every inner block of the panel is IB wide (the
// column guards then fold away and the block's stores are unconditional).
template <class T, int R, int IB, int NTH = PT>
inline void launch_leaf(dim3 grid, size_t shm, hipStream_t s, const LeafArgs<T>& la) {
  if (la.w % IB == 0) hipLaunchKernelGGL((leaf_kernel<T, R, IB, true, NTH>), grid, dim3(NTH), shm, s, la);
  else hipLaunchKernelGGL((leaf_kernel<T, R, IB, false, NTH>), grid, dim3(NTH), shm, s, la);
}

// Leaf launches by panel height m = n - k0 (one workgroup per matrix):
// thin: stacks of at least four matrices per CU, one or two waves per matrix
// (m <= 256); wide: 256 or 512 threads (any m up to 8 * 512).
int leaf_launch_thin(const LeafArgs<float>& la, int m, dim3 grid, size_t shm, hipStream_t s);
int leaf_launch_thin(const LeafArgs<double>& la, int m, dim3 grid, size_t shm, hipStream_t s);
int leaf_launch_wide(const LeafArgs<float>& la, int m, dim3 grid, size_t shm, hipStream_t s);
int leaf_launch_wide(const LeafArgs<double>& la, int m, dim3 grid, size_t shm, hipStream_t s);

}  // namespace npph
