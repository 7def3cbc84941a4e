// Shared definitions for the numpp HIP backend (gfx950 / MI355X).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstddef>

#include "numpp_hip.h"

namespace npph {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kWave = 64;

// 16 x 16 x 4 MFMA in FP64 or FP32. Operands: one A element and one B element
// per lane (A[i][k]: lane = i + 16k, B[k][j]: lane = j + 16k); result: four
// C elements per lane, column lane % 16, rows row(lane / 16, v), v = 0..3
// (lane/16 + 4v for the FP64 instruction, 4 (lane/16) + v for the FP32 one).
template <class T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// ---- error reporting (thread-local message, C-ABI status codes) ----
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define NPPH_HIP(x)                                                        \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) return ::npph::hip_fail(e_, #x, __FILE__, __LINE__); \
  } while (0)

#define NPPH_TRY(x)               \
  do {                            \
    int s_ = (x);                 \
    if (s_ != NPPH_OK) return s_; \
  } while (0)

// ---- device state ----
// One process-wide context per device, created once by ensure_init(): device
// probe (gfx950 check) and a default stream for host-pointer calls.
struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};

int ensure_init();                // NPPH_OK or NPPH_ERR_NO_DEVICE
hipStream_t default_stream();     // library-owned stream for host-pointer calls
int cu_count();                   // compute units on the device
enum WsSlot { kWsA = 0, kWsB, kWsPiv, kWsInfo, kWsLo, kWsR, kWsX, kWsNorm, kWsMisc, kWsMove, kWsPerm, kWsD, kWsNet, kWsFlags, kWsPart, kWsSync, kWsSplitA, kWsSplitB, kWsCount };

// Scratch memory. Every entry point runs inside a CallScope for the stream it
// issues on. The scope binds that stream's workspace set to the calling thread
// and holds the set's lock until the entry point returns, so
//   - work issued on one stream reuses that stream's buffers in stream order,
//   - calls on different streams never share a buffer,
//   - two threads issuing on the same stream take turns (a host-pointer call
//     holds the lock through its final synchronisation),
//   - growing a slot waits only for the set's own stream before freeing.
// workspace() returns a device buffer of at least `bytes` from a named slot of
// the bound set (grow-only; freed when the stream is destroyed through
// npph_stream_destroy, else at process exit).
int workspace(int slot, size_t bytes, void** out);
// Incremented whenever a workspace slot is reallocated; a captured hipGraph
// that baked in workspace addresses is stale once this changes.
uint64_t workspace_generation();
// True while `s` is being captured into a hipGraph (no allocation, no sync).
bool stream_capturing(hipStream_t s);

struct WsSet;
// A private set (a serving worker's): bound with ws_bind for the worker's
// lifetime; CallScopes on that thread then leave it in place.
WsSet* ws_create(hipStream_t owner);
void ws_destroy(WsSet* w);
void ws_bind(WsSet* w);
WsSet* ws_bound();  // the set bound on the calling thread (nullptr: none)
bool ws_private();  // the calling thread has a private set bound
// Frees the per-stream set of `s` (after synchronising s).
void ws_release_stream(hipStream_t s);

class CallScope {
 public:
  explicit CallScope(hipStream_t s);
  ~CallScope();
  CallScope(const CallScope&) = delete;
  CallScope& operator=(const CallScope&) = delete;

 private:
  WsSet* set_ = nullptr;
};

// Caller-provided scratch in place of the workspace slots (the allocation-free
// entry point): while a ScratchArena is bound on a thread, workspace() carves
// 256-byte-aligned regions out of [base, base + bytes) instead of allocating
// (a slot asked for again gets its earlier region when that is large enough),
// and fails with NPPH_ERR_INVALID when the arena is full.
struct ScratchArena {
  char* base = nullptr;
  size_t bytes = 0;
  size_t used = 0;
  size_t off[kWsCount] = {};
  size_t len[kWsCount] = {};
};
void arena_bind(ScratchArena* a);  // nullptr unbinds
bool arena_bound();

// While a PlainIssue is alive on a thread, large factorisations on that thread
// are issued as plain launches on the caller's stream (no cached graphs): a
// caller that runs work eagerly and then captures the same work needs both
// runs to take the same path and touch the same workspaces.
struct PlainIssue {
  PlainIssue();
  ~PlainIssue();
  PlainIssue(const PlainIssue&) = delete;
  PlainIssue& operator=(const PlainIssue&) = delete;
};
bool plain_issue();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- wave-level helpers ----
__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, kWave); }
__device__ __forceinline__ int shfl_xor_i(int v, int m) { return __shfl_xor(v, m, kWave); }

__device__ __forceinline__ double readlane_d(double v, int lane) {
  int2 u = __builtin_bit_cast(int2, v);
  u.x = __builtin_amdgcn_readlane(u.x, lane);
  u.y = __builtin_amdgcn_readlane(u.y, lane);
  return __builtin_bit_cast(double, u);
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
template <class T>
__device__ __forceinline__ T readlane(T v, int lane);
template <>
__device__ __forceinline__ double readlane<double>(double v, int lane) { return readlane_d(v, lane); }
template <>
__device__ __forceinline__ float readlane<float>(float v, int lane) { return readlane_f(v, lane); }

// Pivot choice with NumPP's rule (linalg.cpp:55-56): the largest |value|, and
// among equal magnitudes the one at the smallest current row position. Returns
// true when (v1, p1) beats (v0, p0).
__device__ __forceinline__ bool piv_better(double v1, int p1, double v0, int p0) {
  return v1 > v0 || (v1 == v0 && p1 < p0);
}
__device__ __forceinline__ bool piv_better(float v1, int p1, float v0, int p0) {
  return v1 > v0 || (v1 == v0 && p1 < p0);
}

// Pivot-search key of entry a at current position pos for step k, matching
// NumPP's search (linalg.cpp:55-57): it starts from row k and moves to a later
// row only when that row's |a| is strictly larger. A NaN at row k therefore is
// never replaced (key +inf, and on a tie the smallest position wins), and a NaN
// below it is never chosen (key -1, below every real candidate).
template <class V>
__device__ __forceinline__ V piv_key(V a, int pos, int k) {
  return a != a ? (pos == k ? V(INFINITY) : V(-1)) : fabs(a);
}

// Value of lane (this lane's position in a DPP pattern CTRL); all patterns used
// stay inside a 16-lane row.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f(double v) {
  int2 u = __builtin_bit_cast(int2, v);
  u.x = __builtin_amdgcn_mov_dpp(u.x, CTRL, 0xF, 0xF, false);
  u.y = __builtin_amdgcn_mov_dpp(u.y, CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, u);
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}

// Wave-wide maximum of a value that is never NaN, and minimum of an int: within
// each 16-lane row, DPP swaps with the neighbouring lane, the neighbouring pair,
// the mirrored half-row and the mirrored row (quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror, row_mirror) leave the row's result in all of its
// lanes; the four row results are then combined with v_readlane. One compare
// per stage.
template <class V>
__device__ __forceinline__ V wave_max(V v) {
  V o;
  o = dpp_f<0xB1>(v); v = o > v ? o : v;
  o = dpp_f<0x4E>(v); v = o > v ? o : v;
  o = dpp_f<0x141>(v); v = o > v ? o : v;
  o = dpp_f<0x140>(v); v = o > v ? o : v;
  const V r0 = readlane<V>(v, 0), r1 = readlane<V>(v, 16), r2 = readlane<V>(v, 32), r3 = readlane<V>(v, 48);
  const V a = r1 > r0 ? r1 : r0, b = r3 > r2 ? r3 : r2;
  return b > a ? b : a;
}
__device__ __forceinline__ int wave_min_i(int v) {
  v = min(v, dpp_i<0xB1>(v));
  v = min(v, dpp_i<0x4E>(v));
  v = min(v, dpp_i<0x141>(v));
  v = min(v, dpp_i<0x140>(v));
  return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// The same argmax over lanes 0..7 only (the other lanes' values are ignored):
// three DPP stages (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror)
// leave the result of lanes 0..7 in lane 0.
template <class V>
__device__ __forceinline__ void lanes8_argmax(V& v, int& pos) {
  const int lane = threadIdx.x & 63;
  V x = v, o;
  o = dpp_f<0xB1>(x); x = o > x ? o : x;
  o = dpp_f<0x4E>(x); x = o > x ? o : x;
  o = dpp_f<0x141>(x); x = o > x ? o : x;
  const V wm = readlane<V>(x, 0);
  const uint64_t m = __ballot(lane < 8 && v == wm);
  int wp;
  if (__popcll(m) == 1) {
    wp = __builtin_amdgcn_readlane(pos, (int)__builtin_ctzll(m));
  } else {
    int q = (lane < 8 && v == wm) ? pos : 0x7fffffff;
    q = min(q, dpp_i<0xB1>(q));
    q = min(q, dpp_i<0x4E>(q));
    q = min(q, dpp_i<0x141>(q));
    wp = __builtin_amdgcn_readlane(q, 0);
  }
  v = wm;
  pos = wp;
}

// Wave-wide argmax of (v, pos) with NumPP's pivot rule (linalg.cpp:55-56): the
// largest v, and among equal values the smallest position. Every lane receives
// the winner. v is never NaN (pivot keys map NaN to +inf or -1, piv_key), so
// the maximum is exact and order-independent; the lanes holding it are found
// with a ballot, and only when several hold it (ties) is their minimum position
// reduced as well.
template <class V>
__device__ __forceinline__ void wave_argmax(V& v, int& pos) {
  if constexpr (sizeof(V) == 8) {
    // FP64 values are ranked by their FP32 roundings first (half the DPP
    // traffic and compares): rounding to nearest never reverses an order, so a
    // single lane holding the FP32 maximum holds the FP64 maximum too. Only
    // when several lanes share the FP32 maximum (equal or nearly equal values)
    // is the exact FP64 argmax below run.
    const float f = (float)v;
    const float fm = wave_max(f);
    const uint64_t mf = __ballot(f == fm);
    if (__popcll(mf) == 1) {
      const int wl = (int)__builtin_ctzll(mf);
      v = readlane<V>(v, wl);
      pos = __builtin_amdgcn_readlane(pos, wl);
      return;
    }
  }
  const V wm = wave_max(v);
  const uint64_t m = __ballot(v == wm);
  int wp;
  if (__popcll(m) == 1) wp = __builtin_amdgcn_readlane(pos, (int)__builtin_ctzll(m));
  else wp = wave_min_i(v == wm ? pos : 0x7fffffff);
  v = wm;
  pos = wp;
}

}  // namespace npph
