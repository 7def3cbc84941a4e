// Uneven batches: many independent systems of different order n and nrhs in
// one call.
//
// A plan sorts the systems into size buckets. Each bucket holds a uniform
// padded stack: a system of order n < N is embedded as diag(A, I) of order N
// with zero right-hand-side rows, which leaves the leading factorisation, its
// pivots, its exact-zero test and its solution unchanged (the padded rows and
// columns never enter the leading block's arithmetic). Buckets then run the
// uniform kernels: one wavefront per system for N <= 64, the blocked LU with
// MFMA updates above. The padding granularity per size range is picked by a
// measured autotuner (per-system time of each candidate bucket order on a
// calibration stack, cached to a JSON file), so the choice depends only on the
// system's (n, nrhs) and the precision, never on the rest of the batch.
//
// Packed layout (host-visible contract): system i's A starts at element
// sum_{j<i} n_j^2, its B and X at sum_{j<i} n_j*nrhs_j, its berr at
// sum_{j<i} nrhs_j; all row-major.
//
// Every precision reports the normwise backward error of each result (FP64:
// LU, getrs and one residual; FP32/BF16: refinement).
//
// Buckets run concurrently: each is assigned to one of up to four branch
// streams (largest estimated work first, onto the least-loaded branch) and
// has its own workspace set, so buckets share no scratch and their launches
// overlap on the device. The refinement of every bucket runs a fixed budget of
// test + correction iterations with no host read-back, so the whole execution
// (fork, per-bucket pack, LU, sweeps, unpack, join) is one stream-ordered
// sequence that is captured once into a hipGraph and replayed. After it, one
// read-back of each bucket's counts tells which buckets still have systems
// iterating (or stopped, for the FP64 fallback); those alone continue eagerly
// from the same device state, and the budget grows for later captures.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "lu.hpp"
#include "prof.hpp"

namespace npph {

int refined_impl(const npph_refine_opts& o, int64_t count, int n, int nrhs, const double* A, const double* B,
                 double* X, double* berr, int32_t* iters_out, int32_t* info_out, hipStream_t s);

namespace {

struct Bucket {
  int N = 0, R = 0;  // padded order, padded nrhs (the bucket's widest)
  int br = 0;        // branch stream
  double cost = 0.0; // estimated work (microseconds)
  WsSet* ws = nullptr;  // the bucket's own workspace set
  RefineRun run;     // the refined solve's state between capture and read-back
  int it_after = 0;  // run.it after the budgeted sweeps
  std::vector<int32_t> ids;
  int32_t* d_ids = nullptr;
  double* A = nullptr;  // cnt x N x N
  double* B = nullptr;  // cnt x N x R
  double* X = nullptr;  // cnt x N x R
  double* berr = nullptr;
  int32_t* iters = nullptr;
  int32_t* info = nullptr;
};

struct Plan {
  int64_t count = 0;
  npph_refine_opts opts{};
  std::vector<int32_t> n, nrhs;
  int32_t *d_n = nullptr, *d_nrhs = nullptr;
  int64_t *d_aoff = nullptr, *d_boff = nullptr, *d_coff = nullptr;
  std::vector<Bucket> buckets;
  std::vector<void*> allocs;
  std::vector<hipStream_t> branches;
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_chain = nullptr;  // per-launch profiling: buckets one after another
  int32_t* d_rem = nullptr;  // per bucket: systems still iterating, stopped without convergence
  int32_t* h_rem = nullptr;  // pinned host copy
  int budget = 0;            // test + correction iterations issued without read-back
  int gBudget = -1;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  const void* gA = nullptr;
  const void* gB = nullptr;
  const void* gX = nullptr;
  const void* gI = nullptr;
  const void* gE = nullptr;
  const void* gT = nullptr;
  hipStream_t gS = nullptr;
  uint64_t gGen = ~0ull;
  bool warmed = false;
};

// ---- pack / unpack ----

// Packs one bucket and prepares its factorisation in the same pass: the working
// copy the LU overwrites (FP64, or rounded to FP32 for the mixed-precision
// factorisation), ||A||_inf of the system itself (its n rows) and, when Ap is
// non-null (orders above 256), a padded FP64 copy for the residual. A
// workgroup takes 16 rows of one system (grid: systems x ceil(N / 16)), a wave
// four of them (lanes over columns), and the same rows of B's padded copy
// (zero rows and columns); row sums meet in anorm by atomic maximum (anorm
// starts at zero).
constexpr int PACK_ROWS = 16;
template <class LO>
__global__ __launch_bounds__(256) void pack_prep_kernel(int N, int R, const int32_t* ids, const int32_t* nn,
                                                        const int32_t* nr, const int64_t* aoff, const int64_t* boff,
                                                        const double* A, const double* B, double* Ap, double* Bp,
                                                        LO* Lo, double* anorm) {
  const int64_t k = blockIdx.x;
  const int i0 = blockIdx.y * PACK_ROWS;
  const int id = ids[k];
  const int n = nn[id], m = nr[id];
  const double* __restrict__ a = A + aoff[id];
  const double* __restrict__ b = B + boff[id];
  double* __restrict__ ap = Ap ? Ap + k * (int64_t)N * N : nullptr;
  LO* __restrict__ lo = Lo + k * (int64_t)N * N;
  double* __restrict__ bp = Bp + k * (int64_t)N * R;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double amax = 0.0;
#pragma unroll
  for (int rr = 0; rr < PACK_ROWS / 4; ++rr) {
    const int i = i0 + 4 * wave + rr;
    if (i >= N) break;  // uniform over the wave
    const double* __restrict__ ar = a + (int64_t)min(i, n - 1) * n;
    double s = 0.0;
#pragma unroll 4
    for (int j = lane; j < N; j += 64) {
      const bool in = i < n && j < n;
      const double v = in ? ar[min(j, n - 1)] : (i == j ? 1.0 : 0.0);
      s += in ? fabs(v) : 0.0;
      if (ap) ap[(int64_t)i * N + j] = v;
      lo[(int64_t)i * N + j] = (LO)v;
    }
#pragma unroll
    for (int q = 32; q >= 1; q >>= 1) s += __shfl_xor(s, q, 64);
    amax = fmax(amax, s);
  }
  if (lane == 0 && i0 + 4 * wave < n)
    atomicMax(reinterpret_cast<unsigned long long*>(anorm + k), (unsigned long long)__double_as_longlong(amax));
  const int rows = min(PACK_ROWS, N - i0);
  for (int e = t; e < rows * R; e += 256) {
    const int i = i0 + e / R, c = e % R;
    bp[(int64_t)i * R + c] = (i < n && c < m) ? b[(int64_t)i * m + c] : 0.0;
  }
}

__global__ void zero_f64_kernel(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ __launch_bounds__(256) void unpack_kernel(int N, int R, const int32_t* ids, const int32_t* nn,
                                                     const int32_t* nr, const int64_t* boff, const int64_t* coff,
                                                     const double* Xp, const int32_t* infop, const double* berrp,
                                                     const int32_t* itersp, double* X, int32_t* info, double* berr,
                                                     int32_t* iters) {
  const int64_t k = blockIdx.x;
  const int id = ids[k];
  const int n = nn[id], m = nr[id];
  const double* xp = Xp + k * (int64_t)N * R;
  double* x = X + boff[id];
  for (int e = threadIdx.x; e < n * m; e += 256) {
    const int i = e / m, c = e - i * m;
    x[e] = xp[(int64_t)i * R + c];
  }
  if (threadIdx.x == 0) {
    if (info) info[id] = infop[k];
    if (iters) iters[id] = itersp ? itersp[k] : 0;
  }
  if (berr && berrp)
    for (int c = threadIdx.x; c < m; c += 256) berr[coff[id] + c] = berrp[k * (int64_t)R + c];
}

// ---- autotuner: per-system time of a uniform padded bucket ----
std::mutex g_tune_mu;
std::map<std::string, double> g_tune;  // key "prec:N:R" -> microseconds per system
bool g_tune_loaded = false;

std::string tune_path() {
  if (const char* e = std::getenv("NUMPP_HIP_TUNE_CACHE")) return e;
  const char* home = std::getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/numpp_hip_tune.json";
}

void tune_load() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  std::ifstream f(tune_path());
  if (!f) return;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string t = ss.str();
  size_t p = 0;
  while ((p = t.find('"', p)) != std::string::npos) {
    const size_t q = t.find('"', p + 1);
    if (q == std::string::npos) break;
    const std::string key = t.substr(p + 1, q - p - 1);
    const size_t c = t.find(':', q);
    if (c == std::string::npos) break;
    g_tune[key] = std::atof(t.c_str() + c + 1);
    p = t.find_first_of(",}", c);
    if (p == std::string::npos) break;
  }
}

// mkdir -p of the cache file's directory (no shell)
void make_dirs(const std::string& dir) {
  for (size_t p = 1; p <= dir.size(); ++p)
    if (p == dir.size() || dir[p] == '/') (void)::mkdir(dir.substr(0, p).c_str(), 0755);
}

void tune_save() {
  const std::string path = tune_path();
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos && slash > 0) make_dirs(path.substr(0, slash));
  std::ofstream f(path);
  if (!f) return;
  f << "{";
  bool first = true;
  for (auto& kv : g_tune) {
    f << (first ? "" : ",") << "\n  \"" << kv.first << "\": " << kv.second;
    first = false;
  }
  f << "\n}\n";
}

__global__ void calib_fill_kernel(int64_t total, int N, double* A, double* B, int R) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) {
    const int64_t w = e % ((int64_t)N * N);
    const int i = (int)(w / N), j = (int)(w % N);
    // deterministic, diagonally weighted test matrix
    const uint64_t h = (uint64_t)e * 0x9E3779B97F4A7C15ull;
    const double u = ((h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
    A[e] = u + (i == j ? (double)N : 0.0);
  }
  if (e < total / N * R) B[e] = 1.0;
}

// Measured microseconds per system for a uniform stack of order N.
int tune_cost(const npph_refine_opts& o, int N, int R, double* us, hipStream_t s) {
  char key[64];
  std::snprintf(key, sizeof key, "p%d:%d:%d", o.precision, N, R);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune_load();
    auto it = g_tune.find(key);
    if (it != g_tune.end()) { *us = it->second; return NPPH_OK; }
  }
  const int64_t cnt = N <= 64 ? 1024 : (N <= 128 ? 256 : 128);
  double *A, *A0, *B, *X, *E;
  int32_t *I, *F;
  NPPH_HIP(hipMalloc(&A, sizeof(double) * cnt * N * N));
  NPPH_HIP(hipMalloc(&A0, sizeof(double) * cnt * N * N));
  NPPH_HIP(hipMalloc(&B, sizeof(double) * cnt * N * R));
  NPPH_HIP(hipMalloc(&X, sizeof(double) * cnt * N * R));
  NPPH_HIP(hipMalloc(&E, sizeof(double) * cnt * R));
  NPPH_HIP(hipMalloc(&I, sizeof(int32_t) * cnt));
  NPPH_HIP(hipMalloc(&F, sizeof(int32_t) * cnt));
  const int64_t total = cnt * N * N;
  hipLaunchKernelGGL(calib_fill_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s, total, N, A0, B, R);
  hipEvent_t e0, e1;
  NPPH_HIP(hipEventCreate(&e0));
  NPPH_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int st = NPPH_OK;
  for (int rep = 0; rep < 3 && st == NPPH_OK; ++rep) {
    NPPH_HIP(hipMemcpyAsync(A, A0, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
    NPPH_HIP(hipEventRecord(e0, s));
    st = refined_impl(o, cnt, N, R, A, B, X, E, I, F, s);
    NPPH_HIP(hipEventRecord(e1, s));
    NPPH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    NPPH_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0) best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(A); (void)hipFree(A0); (void)hipFree(B); (void)hipFree(X); (void)hipFree(E);
  (void)hipFree(I); (void)hipFree(F);
  if (st != NPPH_OK) return st;
  *us = 1e3 * best / cnt;
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[key] = *us;
  tune_save();
  return NPPH_OK;
}

int round_up(int v, int g) { return (v + g - 1) / g * g; }
// Orders up to 64 pad to a multiple of 8 (the small-order kernels take any
// n <= 64; their cost follows the order they are given).
int small_order(int n) { return round_up(n, 8); }

// Bucket order for a system of order n under padding granularity G (n > 64).
int bucket_order(int n, int G) { return n <= 64 ? small_order(n) : round_up(n, G); }

template <class T>
int dalloc(Plan* p, T** out, size_t count) {
  void* q = nullptr;
  NPPH_HIP(hipMalloc(&q, sizeof(T) * (count ? count : 1)));
  p->allocs.push_back(q);
  *out = static_cast<T*>(q);
  return NPPH_OK;
}

int plan_build(Plan* p, int autotune, hipStream_t s) {
  const int64_t count = p->count;
  int maxr = 1;
  for (int64_t i = 0; i < count; ++i) maxr = std::max(maxr, (int)p->nrhs[i]);
  const int R = maxr <= 8 ? std::max(1, maxr) : round_up(maxr, 8);
  // Candidate padding granularities for n > 64; pick the one with the least
  // measured total time (sum over buckets of count x per-system time).
  std::vector<int> cands = {32, 64};
  int bestG = 32;
  if (const char* e = std::getenv("NUMPP_HIP_VBATCH_G")) {
    const int g = std::atoi(e);
    if (g == 16 || g == 32 || g == 64) cands = {g};
    bestG = cands[0];
  }
  if (autotune && cands.size() > 1) {
    double bestT = 1e300;
    for (int G : cands) {
      std::map<int, int64_t> hist;
      for (int64_t i = 0; i < count; ++i) hist[bucket_order(p->n[i], G)]++;
      double tot = 0.0;
      for (auto& kv : hist) {
        double us = 0.0;
        NPPH_TRY(tune_cost(p->opts, kv.first, R, &us, s));
        tot += us * kv.second;
      }
      if (tot < bestT) { bestT = tot; bestG = G; }
    }
  }
  std::map<int, std::vector<int32_t>> by;
  for (int64_t i = 0; i < count; ++i) by[bucket_order(p->n[i], bestG)].push_back((int32_t)i);
  // Buckets with their estimated work; a bucket holding more than a quarter
  // of the whole is cut into equal parts (each at least 256 systems), so the
  // branches can share the largest orders. Kernels work per system, so where a
  // system lands does not change its result.
  struct Part {
    int N;
    std::vector<int32_t> ids;
    double cost;
  };
  std::vector<Part> parts;
  double total = 0.0;
  for (auto& kv : by) {
    double us = 0.0;
    if (autotune) NPPH_TRY(tune_cost(p->opts, kv.first, R, &us, s));
    const double per = autotune ? us : (double)kv.first * kv.first * (kv.first + 64.0) * 1e-6;
    parts.push_back({kv.first, kv.second, per * (double)kv.second.size()});
    total += parts.back().cost;
  }
  std::vector<Part> cut;
  for (auto& q : parts) {
    const int64_t c = (int64_t)q.ids.size();
    int pieces = (int)std::min<int64_t>((int64_t)std::ceil(q.cost / (0.25 * total)), c / 256);
    pieces = std::max(1, pieces);
    for (int k = 0; k < pieces; ++k) {
      const int64_t lo = c * k / pieces, hi = c * (k + 1) / pieces;
      cut.push_back({q.N, std::vector<int32_t>(q.ids.begin() + lo, q.ids.begin() + hi),
                     q.cost * (double)(hi - lo) / (double)c});
    }
  }
  for (auto& q : cut) {
    Bucket b;
    b.N = q.N;
    b.ids = std::move(q.ids);
    int rb = 1;
    for (int32_t id : b.ids) rb = std::max(rb, (int)p->nrhs[id]);
    b.R = rb;
    const size_t c = b.ids.size();
    b.cost = q.cost;
    NPPH_TRY(dalloc(p, &b.d_ids, c));
    // orders up to 256 keep their matrices packed (the residual reads them in
    // place); larger ones get a padded FP64 copy
    if (b.N > 256) NPPH_TRY(dalloc(p, &b.A, c * b.N * b.N));
    NPPH_TRY(dalloc(p, &b.B, c * b.N * b.R));
    NPPH_TRY(dalloc(p, &b.X, c * b.N * b.R));
    NPPH_TRY(dalloc(p, &b.berr, c * b.R));
    NPPH_TRY(dalloc(p, &b.iters, c));
    NPPH_TRY(dalloc(p, &b.info, c));
    NPPH_HIP(hipMemcpyAsync(b.d_ids, b.ids.data(), sizeof(int32_t) * c, hipMemcpyHostToDevice, s));
    p->buckets.push_back(std::move(b));
  }
  // branches: up to four streams (the device's hardware queues per process),
  // buckets placed largest first onto the least-loaded one
  const int nbr = std::min<int>(4, (int)p->buckets.size());
  std::vector<size_t> order(p->buckets.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return p->buckets[x].cost > p->buckets[y].cost; });
  std::vector<double> load(nbr, 0.0);
  for (size_t i : order) {
    const int br = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    p->buckets[i].br = br;
    load[br] += p->buckets[i].cost;
  }
  for (int i = 0; i < nbr; ++i) {
    hipStream_t bs = nullptr;
    NPPH_HIP(hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
    p->branches.push_back(bs);
    hipEvent_t e = nullptr;
    NPPH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    p->ev_join.push_back(e);
  }
  NPPH_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
  NPPH_HIP(hipEventCreateWithFlags(&p->ev_chain, hipEventDisableTiming));
  for (auto& b : p->buckets) b.ws = ws_create(p->branches[b.br]);
  NPPH_TRY(dalloc(p, &p->d_rem, 2 * p->buckets.size()));
  void* hp = nullptr;
  NPPH_HIP(hipHostMalloc(&hp, sizeof(int32_t) * 2 * p->buckets.size(), hipHostMallocDefault));
  p->h_rem = static_cast<int32_t*>(hp);
  const bool bf16 = p->opts.precision == NPPH_PREC_BF16;
  // the budget grows after an execution that needed more (finish_plan)
  p->budget = p->opts.precision == NPPH_PREC_FP64 ? 0 : (bf16 ? 3 : 1);
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}

// Releases what plan_build made beyond the plain allocations.
void plan_release(Plan* p) {
  for (auto& b : p->buckets)
    if (b.ws) ws_destroy(b.ws);
  for (hipStream_t bs : p->branches) (void)hipStreamDestroy(bs);
  for (hipEvent_t e : p->ev_join) (void)hipEventDestroy(e);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_chain) (void)hipEventDestroy(p->ev_chain);
  if (p->h_rem) (void)hipHostFree(p->h_rem);
  for (void* q : p->allocs) (void)hipFree(q);
}

// Binds a bucket's workspace set for the scope, restoring the caller's.
struct BucketScope {
  WsSet* outer;
  explicit BucketScope(WsSet* w) : outer(ws_bound()) { ws_bind(w); }
  ~BucketScope() { ws_bind(outer); }
};

int unpack(Plan* p, Bucket& b, double* X, double* berr, int32_t* iters, int32_t* info, hipStream_t bs) {
  hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)b.ids.size()), dim3(256), 0, bs, b.N, b.R, b.d_ids, p->d_n,
                     p->d_nrhs, p->d_boff, p->d_coff, b.X, b.info, b.berr, b.iters, X, info, berr, iters);
  NPPH_HIP(hipGetLastError());
  return NPPH_OK;
}

// One bucket's stream-ordered work on its branch: pack, factorise, solve
// (FP64) or sweep within the budget (the other precisions), unpack.
int issue_bucket(Plan* p, size_t bi, const double* A, const double* B, double* X, double* berr, int32_t* iters,
                 int32_t* info, hipStream_t bs) {
  Bucket& b = p->buckets[bi];
  BucketScope scope(b.ws);
  const int64_t c = (int64_t)b.ids.size();
  void* buf;
  NPPH_TRY(workspace(kWsLo, refine_factor_bytes(p->opts.precision, c, b.N), &buf));
  RefineFactors f;
  refine_factor_carve(p->opts.precision, c, b.N, buf, f);
  hipLaunchKernelGGL(zero_f64_kernel, dim3((unsigned)ceil_div(c, 256)), dim3(256), 0, bs, f.anorm, c);
  const dim3 pgrid((unsigned)c, (unsigned)ceil_div(b.N, PACK_ROWS));
  if (p->opts.precision == NPPH_PREC_FP64)
    hipLaunchKernelGGL(pack_prep_kernel<double>, pgrid, dim3(256), 0, bs, b.N, b.R, b.d_ids, p->d_n,
                       p->d_nrhs, p->d_aoff, p->d_boff, A, B, b.A, b.B, f.lu64, f.anorm);
  else
    hipLaunchKernelGGL(pack_prep_kernel<float>, pgrid, dim3(256), 0, bs, b.N, b.R, b.d_ids, p->d_n,
                       p->d_nrhs, p->d_aoff, p->d_boff, A, B, b.A, b.B, f.lu32, f.anorm);
  NPPH_HIP(hipGetLastError());
  NPPH_TRY(refine_factor(p->opts, b.A, f, bs, true));
  RefineRun& r = b.run;
  r = RefineRun();
  r.o = p->opts;
  r.A = b.A;
  if (!b.A) r.var = PackedA{b.d_ids, p->d_n, p->d_aoff, A};
  r.f = f;
  r.nrhs = b.R;
  r.B = b.B;
  r.X = b.X;
  r.berr_out = b.berr;
  r.iters_out = b.iters;
  r.info_out = b.info;
  NPPH_TRY(refine_begin(r, bs));
  if (p->opts.precision != NPPH_PREC_FP64) {
    NPPH_TRY(refine_steps(r, p->budget, bs));
    NPPH_HIP(hipMemcpyAsync(p->d_rem + 2 * bi, r.ctr, 2 * sizeof(int32_t), hipMemcpyDeviceToDevice, bs));
  }
  b.it_after = r.it;
  NPPH_TRY(refine_end(r, 0, bs));
  return unpack(p, b, X, berr, iters, info, bs);
}

// The stream-ordered part of an execution (what a graph captures): fork onto
// the branches, every bucket, join back onto s.
int run_plan(Plan* p, const double* A, const double* B, double* X, double* berr, int32_t* iters, int32_t* info,
             hipStream_t s) {
  NPPH_HIP(hipEventRecord(p->ev_fork, s));
  for (hipStream_t bs : p->branches) NPPH_HIP(hipStreamWaitEvent(bs, p->ev_fork, 0));
  // Under per-launch profiling the buckets run one after another, so each
  // launch's event interval holds that launch alone.
  const bool serial = prof_enabled();
  int st = NPPH_OK;
  for (size_t i = 0; i < p->buckets.size() && st == NPPH_OK; ++i) {
    hipStream_t bs = p->branches[p->buckets[i].br];
    if (serial && i > 0) NPPH_HIP(hipStreamWaitEvent(bs, p->ev_chain, 0));
    st = issue_bucket(p, i, A, B, X, berr, iters, info, bs);
    if (serial) NPPH_HIP(hipEventRecord(p->ev_chain, bs));
  }
  // join every branch even after a failure, so a capture can end cleanly
  for (size_t k = 0; k < p->branches.size(); ++k) {
    NPPH_HIP(hipEventRecord(p->ev_join[k], p->branches[k]));
    NPPH_HIP(hipStreamWaitEvent(s, p->ev_join[k], 0));
  }
  return st;
}

// After run_plan (eager or replayed): buckets whose budget left systems
// iterating continue with read-backs; buckets with systems stopped without
// convergence get the FP64 fallback; both unpack again. The budget of later
// captures grows to cover what was needed. Synchronises s (refining
// precisions only).
int finish_plan(Plan* p, double* X, double* berr, int32_t* iters, int32_t* info, hipStream_t s) {
  if (p->opts.precision == NPPH_PREC_FP64) return NPPH_OK;
  const size_t nb = p->buckets.size();
  NPPH_HIP(hipMemcpyAsync(p->h_rem, p->d_rem, sizeof(int32_t) * 2 * nb, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  std::vector<bool> used(p->branches.size(), false);
  int need = p->budget;
  for (size_t i = 0; i < nb; ++i) {
    Bucket& b = p->buckets[i];
    const int32_t act = p->h_rem[2 * i], fb = p->h_rem[2 * i + 1];
    if (act <= 0 && !(fb > 0 && p->opts.fallback_fp64)) continue;
    hipStream_t bs = p->branches[b.br];
    BucketScope scope(b.ws);
    b.run.it = b.it_after;  // the state the budgeted sweeps left (also after a replay)
    b.run.eta_issued = true;
    int32_t n_fb = fb;
    if (act > 0) {
      NPPH_TRY(refine_loop(b.run, &n_fb, bs));
      need = std::max(need, b.run.it);  // the loop ended after the test of iteration `it`
    }
    NPPH_TRY(refine_end(b.run, n_fb, bs));
    NPPH_TRY(unpack(p, b, X, berr, iters, info, bs));
    used[b.br] = true;
  }
  for (size_t k = 0; k < p->branches.size(); ++k)
    if (used[k]) {
      NPPH_HIP(hipEventRecord(p->ev_join[k], p->branches[k]));
      NPPH_HIP(hipStreamWaitEvent(s, p->ev_join[k], 0));
    }
  const int cap = (p->opts.max_iter > 0 ? p->opts.max_iter : 10) + 1;
  p->budget = std::min(cap, need);
  return NPPH_OK;
}

}  // namespace

}  // namespace npph

using namespace npph;

extern "C" {

int npph_vbatch_plan_create(int64_t count, const int32_t* n, const int32_t* nrhs, const npph_refine_opts* opts,
                            int autotune, void** plan_out) {
  NPPH_TRY(ensure_init());
  if (count <= 0 || !n || !nrhs || !opts || !plan_out) { set_error("vbatch: bad arguments"); return NPPH_ERR_INVALID; }
  if (count > 65535 * 16) { set_error("vbatch: count too large"); return NPPH_ERR_INVALID; }
  for (int64_t i = 0; i < count; ++i) {
    if (n[i] < 1 || n[i] > max_blocked_n() || nrhs[i] < 1 || nrhs[i] > 4096) {
      set_error("vbatch: system %ld has n=%d nrhs=%d", (long)i, n[i], nrhs[i]);
      return NPPH_ERR_INVALID;
    }
  }
  Plan* p = new Plan;
  p->count = count;
  p->opts = *opts;
  p->n.assign(n, n + count);
  p->nrhs.assign(nrhs, nrhs + count);
  std::vector<int64_t> aoff(count), boff(count), coff(count);
  int64_t a = 0, b = 0, c = 0;
  for (int64_t i = 0; i < count; ++i) {
    aoff[i] = a; boff[i] = b; coff[i] = c;
    a += (int64_t)n[i] * n[i];
    b += (int64_t)n[i] * nrhs[i];
    c += nrhs[i];
  }
  hipStream_t s = default_stream();
  CallScope cs(s);
  int st = NPPH_OK;
  auto up = [&](auto** dst, const auto* src) {
    if (st != NPPH_OK) return;
    st = dalloc(p, dst, count);
    if (st == NPPH_OK && hipMemcpyAsync(*dst, src, sizeof(**dst) * count, hipMemcpyHostToDevice, s) != hipSuccess) {
      set_error("vbatch: upload failed");
      st = NPPH_ERR_HIP;
    }
  };
  up(&p->d_n, n);
  up(&p->d_nrhs, nrhs);
  up(&p->d_aoff, aoff.data());
  up(&p->d_boff, boff.data());
  up(&p->d_coff, coff.data());
  if (st == NPPH_OK) st = plan_build(p, autotune, s);
  if (st != NPPH_OK) {
    (void)hipStreamSynchronize(s);
    plan_release(p);
    delete p;
    return st;
  }
  *plan_out = p;
  return NPPH_OK;
}

int npph_vbatch_plan_destroy(void* plan) {
  NPPH_TRY(ensure_init());
  Plan* p = static_cast<Plan*>(plan);
  if (!p) return NPPH_OK;
  NPPH_HIP(hipDeviceSynchronize());
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  plan_release(p);
  delete p;
  return NPPH_OK;
}

// Bucket summary: nbuckets on return; order/count/padded-rhs arrays filled up
// to `cap` entries when non-null.
int npph_vbatch_plan_info(void* plan, int* nbuckets, int32_t* order, int64_t* counts, int32_t* rpad, int cap) {
  Plan* p = static_cast<Plan*>(plan);
  if (!p || !nbuckets) { set_error("vbatch: bad plan"); return NPPH_ERR_INVALID; }
  *nbuckets = (int)p->buckets.size();
  for (int i = 0; i < (int)p->buckets.size() && i < cap; ++i) {
    if (order) order[i] = p->buckets[i].N;
    if (counts) counts[i] = (int64_t)p->buckets[i].ids.size();
    if (rpad) rpad[i] = p->buckets[i].R;
  }
  return NPPH_OK;
}

int npph_vbatch_execute(void* plan, const double* A, const double* B, double* X, double* berr, int32_t* iters,
                        int32_t* info, int use_graph, void* stream) {
  NPPH_TRY(ensure_init());
  Plan* p = static_cast<Plan*>(plan);
  if (!p || !A || !B || !X) { set_error("vbatch: bad arguments"); return NPPH_ERR_INVALID; }
  hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  CallScope cs(s);
  // First call: run eagerly so every bucket's workspace set reaches its final
  // size; graphs are captured from the second call on.
  if (!use_graph || !p->warmed) {
    NPPH_TRY(run_plan(p, A, B, X, berr, iters, info, s));
    p->warmed = true;
    return finish_plan(p, X, berr, iters, info, s);
  }
  bool stale = !p->exec || p->gA != A || p->gB != B || p->gX != X || p->gI != info || p->gE != berr ||
               p->gT != iters || p->gS != s || p->gGen != workspace_generation() || p->gBudget != p->budget;
  if (stale) {
    if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
    if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
    const uint64_t gen = workspace_generation();
    NPPH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = run_plan(p, A, B, X, berr, iters, info, s);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (st != NPPH_OK) { if (g) (void)hipGraphDestroy(g); return st; }
    NPPH_HIP(ec);
    if (workspace_generation() != gen) {
      (void)hipGraphDestroy(g);
      set_error("vbatch: workspace grew during capture");
      return NPPH_ERR_INVALID;
    }
    p->graph = g;
    NPPH_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    p->gA = A; p->gB = B; p->gX = X; p->gI = info; p->gE = berr; p->gT = iters; p->gS = s; p->gGen = gen;
    p->gBudget = p->budget;
  }
  NPPH_HIP(hipGraphLaunch(p->exec, s));
  return finish_plan(p, X, berr, iters, info, s);
}

}  // extern "C"
