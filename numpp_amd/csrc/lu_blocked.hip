// Recursive LU with partial pivoting for a uniform stack of matrices (n > 64),
// in working precision T (double for the plain path, float for the
// mixed-precision path).
//
// Rows never move during the factorisation: position p lives in physical row
// perm[p] (perm starts as the identity and the leaves exchange its entries as
// they pivot), and every kernel reaches the rows of a position range through
// perm. getrf_rec(k0, w) factors the column range [k0, k0+w) over positions
// [k0, n):
//   w <= 64: one leaf launch (leaf_kernel);
//   else, with w1 = w/2 (a multiple of 64):
//     getrf_rec(k0, w1)
//     trsm_rec: A12 <- L11^-1 A12            (recursive: leaf trsm + GEMM)
//     GEMM:     A22 <- A22 - A21 A12          (MFMA, K = w1)
//     getrf_rec(k0+w1, w-w1)
// so most flops run in GEMMs with a large K (the top level is n/2 x n/2 x n/2)
// and the trailing matrix is read O(log n) times instead of O(n/64). No row
// is ever copied: a factorisation leaves logical row p of L\U in physical row
// perm[p], which the solves read through perm; unscramble_rows gathers the
// packed row order for callers that return the factors.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "lu.hpp"
#include "prof.hpp"
#include "lu_leaf.hpp"

namespace npph {

namespace {

// One 16 x 16 MFMA tile from LDS: X[0..16)[0..K) Y[0..K)[0..16) for X and Y
// with row stride LD (K a multiple of 4).
template <class T, int LD, int K>
__device__ __forceinline__ typename Mfma<T>::acc_t lds_tile(const T* x, const T* y, int lane) {
  typename Mfma<T>::acc_t acc = {0, 0, 0, 0};
  const int r = lane & 15, kk = lane >> 4;
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 4) acc = Mfma<T>::run(x[r * LD + k0 + kk], y[(k0 + kk) * LD + r], acc);
  return acc;
}

template <class T, int LD>
__device__ __forceinline__ void lds_store_tile(const typename Mfma<T>::acc_t& acc, T* dst, T scale, int lane) {
#pragma unroll
  for (int v = 0; v < 4; ++v) dst[Mfma<T>::row(lane >> 4, v) * LD + (lane & 15)] = scale * acc[v];
}

// B <- L^-1 B for the unit-lower s x s block L at positions (r0, r0), s <= 64,
// and the columns [c0, c0+nc) (rows through perm). Each workgroup (256
// threads, 64 columns) forms M = L^-1 in LDS and applies it on the matrix core,
// so no step of a 64-long substitution chain is left on the critical path:
//   - the four 16 x 16 diagonal blocks of M by substitution (one thread per
//     column of a block, 15 steps);
//   - block doubling, M21 = -M22 (L21 M11), for the 16-row pairs and then the
//     32-row pair, on 16 x 16 MFMA tiles;
//   - the tile's 64 x 64 columns of B become M B (MFMA, M lower triangular).
// L is padded to 64 with the identity, so M's padding rows and columns are the
// identity's and the padding rows of B (read as zero) are never written. Each
// level's T = L21 M11 overwrites the L21 block it was formed from, and the B
// tile takes L's place once M is complete (two 64 x 64 arrays of LDS).
// Applying the explicit inverse instead of substituting changes only the
// rounding of B, within the usual bound for a unit-lower L whose entries are at
// most 1 in magnitude (partial pivoting).
template <class T>
__global__ __launch_bounds__(256) void trsm_inv_kernel(T* Abase, int64_t strideA, int n, int r0, int s, int c0,
                                                       int nc, const int32_t* permb, const int32_t* mask) {
  const int64_t mat = blockIdx.y;
  if (mask && mask[mat] == 0) return;
  T* A = Abase + mat * strideA;
  constexpr int LD = LEAF + 1;
  __shared__ T Ls[LEAF * LD];  // L padded with the identity; level products; the B tile
  __shared__ T Ms[LEAF * LD];  // M = L^-1
  __shared__ int rowp[LEAF];   // physical rows of positions r0 .. r0 + s - 1
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int32_t* perm = permb + mat * n;
  if (t < LEAF) rowp[t] = perm[r0 + min(t, s - 1)];
  __syncthreads();
  // this workgroup's 64 columns of B (registers until M is ready) and L
  const int cb = blockIdx.x * 64;
  constexpr int PER = LEAF * 64 / 256;
  T bv[PER], lv[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = t + 256 * u, i = e >> 6, j = e & 63;
    const T* row = A + (int64_t)rowp[i] * n;
    bv[u] = (i < s && cb + j < nc) ? row[c0 + cb + j] : T(0);
    lv[u] = row[r0 + min(j, s - 1)];  // clamped: a real entry of a real row
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = t + 256 * u, i = e >> 6, j = e & 63;
    Ls[i * LD + j] = (j < i && i < s) ? lv[u] : (i == j ? T(1) : T(0));
    Ms[i * LD + j] = T(0);  // M's blocks above the diagonal stay zero
  }
  __syncthreads();
  // diagonal 16 x 16 blocks of M: thread (b, j) substitutes unit vector e_j
  if (t < 64) {
    const int b = t >> 4, j = t & 15;
    const T* lb = Ls + 16 * b * LD + 16 * b;
    T x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = q == j ? T(1) : T(0);
#pragma unroll
    for (int p = 0; p < 15; ++p) {
      const T xp = x[p];
#pragma unroll
      for (int q = p + 1; q < 16; ++q) x[q] = x[q] - lb[q * LD + p] * xp;
    }
    T* mb = Ms + 16 * b * LD + 16 * b;
#pragma unroll
    for (int q = 0; q < 16; ++q) mb[q * LD + j] = x[q];
  }
  __syncthreads();
  // 16-row pairs (rows 16..31 x columns 0..15, rows 48..63 x columns 32..47):
  // waves 0 and 1 form T = L21 M11 (over L21, which only this wave reads),
  // then M21 = -M22 T
  if (wave < 2) {
    const int o = 32 * wave;
    auto acc = lds_tile<T, LD, 16>(Ls + (o + 16) * LD + o, Ms + o * LD + o, lane);
    lds_store_tile<T, LD>(acc, Ls + (o + 16) * LD + o, T(1), lane);
  }
  __syncthreads();
  if (wave < 2) {
    const int o = 32 * wave;
    auto acc = lds_tile<T, LD, 16>(Ms + (o + 16) * LD + o + 16, Ls + (o + 16) * LD + o, lane);
    lds_store_tile<T, LD>(acc, Ms + (o + 16) * LD + o, T(-1), lane);
  }
  __syncthreads();
  // the 32-row pair (rows 32..63 x columns 0..31): wave w owns 16 x 16 tile
  // (w / 2, w % 2) of T = L21 M11 (stored over L21 once every wave has read
  // it) and then of M21 = -M22 T
  const int ti = wave >> 1, tj = wave & 1;
  {
    auto acc = lds_tile<T, LD, 32>(Ls + (32 + 16 * ti) * LD, Ms + 16 * tj, lane);
    __syncthreads();
    lds_store_tile<T, LD>(acc, Ls + (32 + 16 * ti) * LD + 16 * tj, T(1), lane);
  }
  __syncthreads();
  {
    // row block ti of the lower-triangular M22 has nonzeros in columns < 16 (ti + 1)
    typename Mfma<T>::acc_t acc;
    if (ti == 0) acc = lds_tile<T, LD, 16>(Ms + 32 * LD + 32, Ls + 32 * LD + 16 * tj, lane);
    else acc = lds_tile<T, LD, 32>(Ms + 48 * LD + 32, Ls + 32 * LD + 16 * tj, lane);
    lds_store_tile<T, LD>(acc, Ms + (32 + 16 * ti) * LD + 16 * tj, T(-1), lane);
  }
  // the B tile over L once every wave has read its T tile
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = t + 256 * u, i = e >> 6, j = e & 63;
    Ls[i * LD + j] = bv[u];
  }
  __syncthreads();
  // M B: wave w owns columns [16w, 16w + 16) of the tile, all four row blocks
  const int col = cb + 16 * wave + (lane & 15);
  auto put = [&](const typename Mfma<T>::acc_t& acc, int rb) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 16 * rb + Mfma<T>::row(lane >> 4, v);
      if (i < s && col < nc) A[(int64_t)rowp[i] * n + c0 + col] = acc[v];
    }
  };
  put(lds_tile<T, LD, 16>(Ms, Ls + 16 * wave, lane), 0);
  put(lds_tile<T, LD, 32>(Ms + 16 * LD, Ls + 16 * wave, lane), 1);
  put(lds_tile<T, LD, 48>(Ms + 32 * LD, Ls + 16 * wave, lane), 2);
  put(lds_tile<T, LD, 64>(Ms + 48 * LD, Ls + 16 * wave, lane), 3);
}

// Packed row order from a factorisation whose rows never moved: row p of
// columns [c0, c0+nc) becomes physical row perm[p]'s, through scratch (phase 0
// gathers into scratch, phase 1 writes back). 16 rows x 256 columns per
// workgroup.
template <class T>
__global__ __launch_bounds__(256) void gather_rows_kernel(T* Abase, int64_t strideA, int n, int c0, int nc,
                                                          const int32_t* permb, T* scratch, int64_t scr_stride,
                                                          int phase) {
  const int64_t mat = blockIdx.z;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= nc) return;
  const int p0 = blockIdx.y * 16;
  T* A = Abase + mat * strideA;
  T* sc = scratch + mat * scr_stride;
  const int32_t* perm = permb + mat * n;
  T buf[16];
  if (phase == 0) {
#pragma unroll
    for (int q = 0; q < 16; ++q) buf[q] = A[(int64_t)perm[min(p0 + q, n - 1)] * n + c0 + c];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (p0 + q < n) sc[(int64_t)(p0 + q) * nc + c] = buf[q];
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) buf[q] = sc[(int64_t)min(p0 + q, n - 1) * nc + c];
#pragma unroll
    for (int q = 0; q < 16; ++q)
      if (p0 + q < n) A[(int64_t)(p0 + q) * n + c0 + c] = buf[q];
  }
}

__global__ void init_factor_kernel(int64_t count, int n, int32_t* perm, int32_t* info) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count * n) perm[i] = (int32_t)(i % n);
  if (i < count) info[i] = 0;
}

// Smallest inner dimension of a trailing update that takes the split-BF16 path
// in the BF16 precision (NUMPP_HIP_BF16_MINK, default 256).
int bf16_min_k() {
  static const int v = [] {
    const char* e = std::getenv("NUMPP_HIP_BF16_MINK");
    const int x = e ? std::atoi(e) : 256;
    return x > 0 ? x : 256;
  }();
  return v;
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

  static int half(int w) {
    int h = (w / 2 + LEAF - 1) / LEAF * LEAF;
    return h >= w ? w - LEAF : h;
  }

  int leaf(int k0, int w) {
    LeafArgs<T> la{A, strideA, n, k0, w, ipiv, perm, info, mask};
    const int m = n - k0;
    const size_t shm = (size_t)m * sizeof(int);
    dim3 grid((unsigned)count);
    ProfScope ps(kProfPanel, (double)count * m * w * w, 2.0 * sizeof(T) * count * m * w, s);
    // Many matrices (at least four per CU): one or two waves per matrix with
    // one row per lane up to 128 rows and two up to 256; the per-column pivot
    // search, pivot-row exchange and barrier then cost one or two waves'
    // instructions instead of four, and the other matrices' waves hide the
    // column chain's latency.
    static const int thin_env = [] {
      const char* e = std::getenv("NUMPP_HIP_LEAF_THIN");
      return e ? std::atoi(e) : 1;
    }();
    // (NUMPP_HIP_LEAF_THIN=2: thin panels up to 256 rows for any stack size)
    const bool thin = thin_env != 0 && (thin_env == 2 || count >= 4 * (int64_t)std::max(cu_count(), 1)) && m <= 256;
    NPPH_TRY(thin ? leaf_launch_thin(la, m, grid, shm, s) : leaf_launch_wide(la, m, grid, shm, s));
    NPPH_HIP(hipGetLastError());
    return NPPH_OK;
  }

  // C <- C - A B with the rows of A and C at positions [ra, ra+M) and the rows
  // of B at positions [rb, rb+K) (all through perm); a, b, c point at column 0
  // of their operand.
  int gemm_update(int64_t M, int64_t N, int64_t K, int ra, T* a, int rb, T* b, T* c) {
    if (M <= 0 || N <= 0 || K <= 0) return NPPH_OK;
    if constexpr (sizeof(T) == 4) {
      // split-BF16 for the deep products; the shallow ones (which would pay two
      // split passes for little MFMA work) stay on the FP32 MFMA path, which is
      // at least as accurate
      if (bf16_terms > 0 && K >= bf16_min_k() && M >= 128 && N >= 128)
        return gemm_bf16split_update(M, N, K, (const float*)a, n, strideA, (const float*)b, n, strideA, (float*)c, n,
                                     strideA, count, bf16_terms, s,
                                     RowMaps{perm + ra, perm + rb, perm + ra, (int64_t)n});
    }
    GemmArgs<T> g{M, N, K, T(-1), T(1), a, n, strideA, b, n, strideA, c, n, strideA, 0, 0, 0};
    g.maps = RowMaps{perm + ra, perm + rb, perm + ra, (int64_t)n};
    return gemm_launch<T>(g, count, s);
  }

  // X <- L^-1 X for the unit-lower block at (r0, r0) of size sz and columns [c0, c0+nc).
  int trsm(int r0, int sz, int c0, int nc) {
    if (nc <= 0) return NPPH_OK;
    if (sz <= LEAF) {
      ProfScope ps(kProfMoveTrsm, (double)count * sz * sz * nc, 2.0 * sizeof(T) * count * (double)sz * nc, s);
      hipLaunchKernelGGL((trsm_inv_kernel<T>), dim3((unsigned)ceil_div(nc, 64), (unsigned)count), dim3(256), 0, s, A,
                         strideA, n, r0, sz, c0, nc, perm, mask);
      NPPH_HIP(hipGetLastError());
      return NPPH_OK;
    }
    const int s1 = half(sz);
    NPPH_TRY(trsm(r0, s1, c0, nc));
    NPPH_TRY(gemm_update(sz - s1, nc, s1, r0 + s1, A + r0, r0, A + c0, A + c0));
    return trsm(r0 + s1, sz - s1, c0, nc);
  }

  int getrf(int k0, int w) {
    if (w <= LEAF) return leaf(k0, w);
    const int w1 = half(w), w2 = w - w1;
    NPPH_TRY(getrf(k0, w1));
    NPPH_TRY(trsm(k0, w1, k0 + w1, w2));
    NPPH_TRY(gemm_update(n - k0 - w1, w2, w1, k0 + w1, A + k0, k0, A + k0 + w1, A + k0 + w1));
    return getrf(k0 + w1, w2);
  }
};


// Row-gather staging per matrix (unscramble_rows): half of the columns at a
// time, at most n x (n/2 + 1) elements.
int64_t gather_stride(int n) { return (int64_t)n * (n / 2 + LEAF); }

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
  {
    const int64_t tot = count * n;
    hipLaunchKernelGGL(init_factor_kernel, dim3((unsigned)ceil_div(tot, 256)), dim3(256), 0, s, count, n, perm, info);
    NPPH_HIP(hipGetLastError());
  }
  return r.getrf(0, n);
}

// Concurrent factorisation groups. The panels and triangular leaves of one
// stack run one workgroup (or a few) per matrix through a chain of dependent
// column steps, so while they run most CUs idle. A large stack of large
// systems is therefore split into G groups of matrices, each factorised on a
// stream of its own (forked from s and joined back onto it by events), so that
// one group's panels overlap the other groups' MFMA updates. The streams and
// events are the device's own, shared by every caller under one mutex (the
// mutex covers only the issue, not the execution).
constexpr int kMaxGroups = 4;
struct GroupStreams {
  // held for an issue, and for the whole capture of a grouped factorisation
  // (its branch streams belong to that capture until it ends, so no other
  // thread may queue work on them meanwhile)
  std::recursive_mutex mu;
  hipStream_t st[kMaxGroups] = {};
  hipEvent_t fork = nullptr, join[kMaxGroups] = {};
};

int group_streams(GroupStreams** out) {
  static std::mutex mu;
  static std::vector<GroupStreams*> per_dev;
  int dev = 0;
  NPPH_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
  if (!per_dev[dev]) {
    auto* g = new GroupStreams();
    for (int i = 0; i < kMaxGroups; ++i) {
      NPPH_HIP(hipStreamCreateWithFlags(&g->st[i], hipStreamNonBlocking));
      NPPH_HIP(hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming));
    }
    NPPH_HIP(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
    per_dev[dev] = g;
  }
  *out = per_dev[dev];
  return NPPH_OK;
}

// Groups for a stack: NUMPP_HIP_LU_GROUPS (1 = off) when set, else 4 for at
// least 16 FP64 matrices of order >= 1024 (64 x 2048 FP64 solve: 17.35 ->
// 16.55 ms). FP32 keeps one group: measured on the same device, four groups
// took its 64 x 2048 refined solve from 12.22 to 12.36-12.62 ms (its cheaper
// updates leave less to hide the panels behind, and the groups' shorter
// products lose more to launch latency than the overlap wins). The split-BF16 update
// keeps one group (its split planes live in the caller's workspace slots).
int lu_groups(int64_t count, int n, size_t elem, int bf16_terms) {
  static const int env = [] {
    const char* e = std::getenv("NUMPP_HIP_LU_GROUPS");
    return e ? std::atoi(e) : 0;
  }();
  if (bf16_terms > 0) return 1;
  int g = env > 0 ? env : (elem == 8 && count >= 16 && n >= 1024 ? 4 : 1);
  g = std::min(g, kMaxGroups);
  while (g > 1 && count < 2 * g) --g;
  return g;
}

template <class T>
int getrf_grouped(int G, int64_t count, int n, T* A, int64_t strideA, int32_t* ipiv, int32_t* perm, int32_t* info,
                  const int32_t* mask, hipStream_t s) {
  GroupStreams* gs = nullptr;
  NPPH_TRY(group_streams(&gs));
  std::lock_guard<std::recursive_mutex> lk(gs->mu);
  NPPH_HIP(hipEventRecord(gs->fork, s));
  for (int g = 0; g < G; ++g) NPPH_HIP(hipStreamWaitEvent(gs->st[g], gs->fork, 0));
  int st = NPPH_OK;
  for (int g = 0; g < G && st == NPPH_OK; ++g) {
    const int64_t c0 = count * g / G, c = count * (g + 1) / G - c0;
    st = getrf_group<T>(c, n, A + c0 * strideA, strideA, ipiv + c0 * n, perm + c0 * n, info + c0, 0,
                        mask ? mask + c0 : nullptr, gs->st[g]);
  }
  // every branch joins back, also after a failure, so a capture ends cleanly
  for (int g = 0; g < G; ++g) {
    NPPH_HIP(hipEventRecord(gs->join[g], gs->st[g]));
    NPPH_HIP(hipStreamWaitEvent(s, gs->join[g], 0));
  }
  return st;
}

}  // namespace

int max_blocked_n() { return 8 * PT; }

size_t blocked_getrf_scratch_bytes(int64_t count, int n, size_t elem) {
  return (elem * (size_t)count * gather_stride(n) + 255) & ~(size_t)255;
}

template <class T>
int unscramble_rows(int64_t count, int n, T* A, int64_t strideA, const int32_t* perm, hipStream_t s) {
  if (count <= 0 || n <= 0) return NPPH_OK;
  if (count > 65535) { set_error("unscramble: count > 65535"); return NPPH_ERR_INVALID; }
  T* scr;
  NPPH_TRY(workspace(kWsMove, sizeof(T) * (size_t)count * gather_stride(n), (void**)&scr));
  const int h = n / 2;
  for (int part = 0; part < 2; ++part) {
    const int c0 = part == 0 ? 0 : h, nc = part == 0 ? h : n - h;
    if (nc <= 0) continue;
    dim3 grid((unsigned)ceil_div(nc, 256), (unsigned)ceil_div(n, 16), (unsigned)count);
    for (int phase = 0; phase < 2; ++phase) {
      hipLaunchKernelGGL(gather_rows_kernel<T>, grid, dim3(256), 0, s, A, strideA, n, c0, nc, perm, scr,
                         gather_stride(n), phase);
      NPPH_HIP(hipGetLastError());
    }
  }
  return NPPH_OK;
}
template int unscramble_rows<double>(int64_t, int, double*, int64_t, const int32_t*, hipStream_t);
template int unscramble_rows<float>(int64_t, int, float*, int64_t, const int32_t*, hipStream_t);

namespace {

// Cached hipGraphs of whole factorisations, keyed by shape, pointers and the
// workspace generation. A new key runs eagerly once (sizing every workspace);
// its second call is captured and every later call replays the graph.
struct GraphKey {
  int tsize;
  int64_t count;
  int n, bf16;
  const void *A, *ipiv, *perm, *info, *mask;
  int64_t strideA;
  uint64_t gen;
  hipStream_t s;
  bool operator==(const GraphKey& o) const {
    return tsize == o.tsize && count == o.count && n == o.n && bf16 == o.bf16 && A == o.A &&
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
  // A caller with a private workspace set (a serving worker, a bucket of an
  // uneven batch) already overlaps its factorisations with other streams' and
  // issues on its own stream alone; so do callers with caller-provided scratch
  // or an outer capture.
  const bool own_streams = ws_private() || plain_issue() || stream_capturing(s);
  const int G = own_streams ? 1 : lu_groups(count, n, sizeof(T), bf16_terms);
  auto issue = [&] {
    if (G > 1) return getrf_grouped<T>(G, count, n, A, strideA, ipiv, perm, info, mask, s);
    return getrf_group<T>(count, n, A, strideA, ipiv, perm, info, bf16_terms, mask, s);
  };
  // Those callers also get a plain issue on their stream instead of a cached
  // graph; so does per-launch profiling (a replayed graph has no per-kernel
  // events).
  if (own_streams || !graphs_enabled() || prof_enabled()) return issue();
  GraphKey key{(int)sizeof(T), count, n, bf16_terms, A, ipiv, perm, info, mask, strideA, workspace_generation(), s};
  std::lock_guard<std::mutex> lk(g_graph_mu);
  for (auto& e : g_graphs)
    if (e.key == key) {
      NPPH_HIP(hipGraphLaunch(e.exec, s));
      return NPPH_OK;
    }
  bool seen = false;
  for (auto& k : g_seen) seen = seen || (k == key);
  if (!seen) {
    const int st = issue();
    if (st == NPPH_OK) {
      key.gen = workspace_generation();  // the eager run may have grown the workspaces
      if (g_seen.size() >= 4 * kMaxGraphs) g_seen.erase(g_seen.begin());
      g_seen.push_back(key);
    }
    return st;
  }
  std::unique_lock<std::recursive_mutex> glk;
  if (G > 1) {
    GroupStreams* gs = nullptr;
    NPPH_TRY(group_streams(&gs));
    glk = std::unique_lock<std::recursive_mutex>(gs->mu);
  }
  NPPH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int st = issue();
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  if (glk.owns_lock()) glk.unlock();
  if (st != NPPH_OK || ec != hipSuccess || workspace_generation() != key.gen) {
    if (graph) (void)hipGraphDestroy(graph);
    if (st != NPPH_OK) return st;
    // capture failed: fall back to an eager issue of the same work
    return issue();
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
