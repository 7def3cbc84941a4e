// Serving loop: solve requests of different n and nrhs as they arrive.
//
// - Requests enter a FIFO queue (npph_server_submit) and are served by W worker
//   threads, each with its own HIP stream (default priority) and its own
//   workspace set, so the work of concurrently admitted requests overlaps on
//   the device.
// - Admission: a request starts only when its device bytes (A copy, factors,
//   B/X, workspace estimate) fit in the byte budget next to everything running
//   and every resident factorisation; least-recently-used resident factors that
//   no running request uses are evicted to make room. A request that cannot fit
//   even on an empty device fails with NPPH_ERR_BUDGET.
// - Continuous batching: every request is solved at a padded order N
//   (bucket_order: diag(A, I), B with zero rows). When a worker takes a head
//   request that factorises, it also takes queued factorising requests of the
//   same N (up to max_batch, budget permitting; requests that keep their
//   factors included) and solves them as one stack; right-hand sides are
//   padded with zero columns to the widest request. Every kernel works per
//   system and per column and ||A||_inf covers the caller's rows only, so a
//   result is bitwise the same alone and in the mix.
// - Resident factors: a request with keep = 1 leaves (padded A, factors) on the
//   device under its key (a copy of its slice when it ran in a group); a later
//   request with the same key skips the factorisation and only runs the
//   refined solve with new right-hand sides. A reuse request
//   whose key is still being factorised by an earlier request (queued or
//   running) is held back and queued again as soon as the factors are resident.
// Host buffers passed to submit must stay valid until wait returns.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "lu.hpp"

namespace npph {
namespace {

using Clock = std::chrono::steady_clock;

struct Resident {
  uint64_t key = 0;
  int n = 0, N = 0;      // caller's order, padded order
  double* A = nullptr;  // padded FP64 copy (residual)
  void* fbuf = nullptr;  // factors
  RefineFactors f;
  size_t bytes = 0;
  int refs = 0;
  std::list<uint64_t>::iterator lru;
};

struct Request {
  int64_t id = 0;
  int n = 0, N = 0, nrhs = 0;  // N: padded order (bucket_order)
  const double* A = nullptr;
  const double* B = nullptr;
  double* X = nullptr;
  double* berr = nullptr;
  uint64_t key = 0;
  int keep = 0;
  int32_t info = 0, iters = 0;
  int status = NPPH_OK;
  bool done = false;
  size_t bytes = 0;
  bool batched = false;
  Clock::time_point t_submit, t_start, t_done;
};

struct Server {
  npph_serve_config cfg{};
  npph_refine_opts opts{};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> queue;
  std::unordered_map<int64_t, Request*> reqs;
  std::unordered_map<uint64_t, Resident*> cache;
  std::unordered_map<uint64_t, int> pending_keep;  // key -> in-flight requests that will keep it
  std::unordered_map<uint64_t, std::vector<Request*>> deferred;  // reuse requests waiting for their key
  std::list<uint64_t> lru;  // front = most recent
  std::vector<std::thread> workers;
  bool stop = false;
  int64_t next_id = 1;
  size_t budget = 0, used = 0, peak = 0;
  int64_t completed = 0, batched = 0, fallbacks = 0, evictions = 0;
};

size_t ws_estimate(int n, int nrhs) {
  // refined solve + fallback: R, Rlo, D (nrhs columns), an FP64 copy of A and
  // B for the fallback, the blocked LU's move scratch and bookkeeping.
  const size_t nn = (size_t)n * n, nb = (size_t)n * nrhs;
  return nn * 8 + nb * 8 * 3 + nb * 4 * 2 + (size_t)n * 128 * 8 + 64 * 1024;
}

// Device bytes of one request of padded order N with R right-hand-side
// columns (the group's widest).
size_t request_bytes(const Server& sv, int N, int R, bool reuse) {
  size_t b = 2 * (size_t)N * R * 8 + ws_estimate(N, R);
  // padded A, factors, and the packed staging of A and B
  if (!reuse) b += (size_t)N * N * 8 + refine_factor_bytes(sv.opts.precision, 1, N) + (size_t)N * (N + R) * 8;
  return b;
}

// Evicts unused resident factors (LRU first) until `need` more bytes fit.
// Caller holds mu. Returns true when the request fits.
bool make_room(Server& sv, size_t need, std::vector<Resident*>& freed) {
  while (sv.used + need > sv.budget) {
    Resident* victim = nullptr;
    for (auto it = sv.lru.rbegin(); it != sv.lru.rend(); ++it) {
      Resident* r = sv.cache[*it];
      if (r->refs == 0) { victim = r; break; }
    }
    if (!victim) return false;
    sv.lru.erase(victim->lru);
    sv.cache.erase(victim->key);
    sv.used -= victim->bytes;
    ++sv.evictions;
    freed.push_back(victim);
  }
  return true;
}

// A request that was to keep factors under `key` has finished (or failed);
// reuse requests held for that key are queued again once the factors are
// resident, or failed when no other in-flight request will provide them.
// Caller holds mu.
void release_pending(Server& sv, uint64_t key) {
  auto pk = sv.pending_keep.find(key);
  if (pk != sv.pending_keep.end() && --pk->second <= 0) sv.pending_keep.erase(pk);
  auto d = sv.deferred.find(key);
  if (d == sv.deferred.end()) return;
  if (sv.cache.count(key)) {
    for (auto q = d->second.rbegin(); q != d->second.rend(); ++q) sv.queue.push_front(*q);
    sv.deferred.erase(d);
  } else if (!sv.pending_keep.count(key)) {
    for (auto* q : d->second) {
      q->status = NPPH_ERR_NOT_FOUND;
      q->done = true;
      q->t_start = q->t_done = Clock::now();
    }
    sv.deferred.erase(d);
  }
}

void free_resident(Resident* r, hipStream_t s) {
  if (r->A) (void)hipFreeAsync(r->A, s);
  if (r->fbuf) (void)hipFreeAsync(r->fbuf, s);
  delete r;
}

// Stream-ordered device buffers of one request, freed (stream-ordered) when
// the holder goes out of scope, on every return path, unless handed over with
// release().
struct AsyncBufs {
  hipStream_t s;
  std::vector<void*> held;
  explicit AsyncBufs(hipStream_t st) : s(st) {}
  AsyncBufs(const AsyncBufs&) = delete;
  AsyncBufs& operator=(const AsyncBufs&) = delete;
  template <class T>
  int get(T** out, size_t bytes) {
    void* p = nullptr;
    NPPH_HIP(hipMallocAsync(&p, bytes ? bytes : 256, s));
    held.push_back(p);
    *out = static_cast<T*>(p);
    return NPPH_OK;
  }
  void release(const void* p) {
    for (auto& q : held)
      if (q == p) q = nullptr;
  }
  void free_all() {
    for (void* q : held)
      if (q) (void)hipFreeAsync(q, s);
    held.clear();
  }
  ~AsyncBufs() { free_all(); }
};

// Padded order of a request. Requests whose orders share a bucket are
// factorised together as diag(A, I) of the bucket's order: powers of two up to
// 64 (the small-order LU's sizes), then multiples of 32 up to 256, of 64 up to
// 1024 (the recursive LU's leaf width) and of 128 above. Every request is
// padded, alone or in a group, so its result does not depend on its company.
int bucket_order(int n) {
  int N;
  if (n <= 64) {
    N = 8;
    while (N < n) N <<= 1;
  } else if (n <= 256) {
    N = (n + 31) & ~31;
  } else if (n <= 1024) {
    N = (n + 63) & ~63;
  } else {
    N = (n + 127) & ~127;
  }
  return std::max(n, std::min(N, max_blocked_n()));
}

// Pads one system: src holds the caller's n x n matrix (contiguous, as copied
// from the host); A (N x N FP64) receives diag(A, I), the LU's working copy
// `lo` (FP64, or FP32 for the mixed precisions) the same values, and
// ||A||_inf is taken over the caller's n rows only (*anorm starts at zero).
// One wave per row of the padded matrix.
template <class LO>
__global__ __launch_bounds__(256) void pad_prep_kernel(int N, int n, const double* __restrict__ src,
                                                       double* __restrict__ A, LO* __restrict__ lo, double* anorm) {
  const int i = (int)(((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  const double* __restrict__ srow = src + (int64_t)min(i, n - 1) * n;
  double* __restrict__ a = A + (int64_t)i * N;
  LO* __restrict__ o = lo + (int64_t)i * N;
  double s = 0.0;
  for (int j = lane; j < N; j += 64) {
    const bool in = i < n && j < n;
    const double v = in ? srow[j] : (i == j ? 1.0 : 0.0);
    s += in ? fabs(v) : 0.0;
    a[j] = v;
    o[j] = (LO)v;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  if (lane == 0 && i < n)
    atomicMax(reinterpret_cast<unsigned long long*>(anorm), (unsigned long long)__double_as_longlong(s));
}

// dst (rows x dcols) <- src (n x m, contiguous) in its leading block, zeros
// elsewhere; with rows = n and dcols = m it is a plain copy.
__global__ void pad_block_kernel(int rows, int dcols, int n, int m, const double* __restrict__ src,
                                 double* __restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * dcols) return;
  const int i = (int)(e / dcols), c = (int)(e - (int64_t)i * dcols);
  dst[e] = (i < n && c < m) ? src[(int64_t)i * m + c] : 0.0;
}

// dst (n x m, contiguous) <- the leading n x m block of src (rows at pitch scols).
__global__ void unpad_block_kernel(int n, int m, int scols, const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * m) return;
  const int i = (int)(e / m), c = (int)(e - (int64_t)i * m);
  dst[e] = src[(int64_t)i * scols + c];
}

// Solves one request with resident factors (of the request's padded order).
int serve_reuse(Server& sv, Request* r, Resident* res, hipStream_t s) {
  AsyncBufs mem(s);
  double *dB, *dX, *dE;
  int32_t *dI, *dF;
  const size_t nb = (size_t)r->n * r->nrhs, Nb = (size_t)res->N * r->nrhs;
  NPPH_TRY(mem.get(&dB, Nb * 8));
  NPPH_TRY(mem.get(&dX, Nb * 8));
  NPPH_TRY(mem.get(&dE, (size_t)r->nrhs * 8));
  NPPH_TRY(mem.get(&dI, 4));
  NPPH_TRY(mem.get(&dF, 4));
  // row-major: the caller's n rows are the leading n * nrhs elements
  if (Nb > nb) NPPH_HIP(hipMemsetAsync(dB + nb, 0, (Nb - nb) * 8, s));
  NPPH_HIP(hipMemcpyAsync(dB, r->B, nb * 8, hipMemcpyHostToDevice, s));
  NPPH_TRY(refine_solve(sv.opts, res->A, res->f, r->nrhs, dB, dX, dE, dI, dF, s));
  NPPH_HIP(hipMemcpyAsync(r->X, dX, nb * 8, hipMemcpyDeviceToHost, s));
  if (r->berr) NPPH_HIP(hipMemcpyAsync(r->berr, dE, (size_t)r->nrhs * 8, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(&r->iters, dI, 4, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(&r->info, dF, 4, hipMemcpyDeviceToHost, s));
  mem.free_all();
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}

// Factorises and solves a group of requests of one padded order N as one
// stack; right-hand sides are padded with zero rows and with zero columns to
// the widest request. Requests with keep = 1 hand their (A, factors) to the
// cache: a lone request hands over its buffers, a request in a group gets a
// copy of its slice, so the group's buffers can go.
int serve_group(Server& sv, std::vector<Request*>& g, hipStream_t s, std::vector<Resident*>& kept) {
  const int N = g[0]->N;
  const int64_t k = (int64_t)g.size();
  const int prec = sv.opts.precision;
  int R = 1;
  for (auto* r : g) R = std::max(R, r->nrhs);
  const size_t NN = (size_t)N * N, NR = (size_t)N * R;
  AsyncBufs mem(s);
  double *dA, *dB, *dX, *dE;
  int32_t *dI, *dF;
  void* fbuf;
  NPPH_TRY(mem.get(&dA, k * NN * 8));
  NPPH_TRY(mem.get(&dB, k * NR * 8));
  NPPH_TRY(mem.get(&dX, k * NR * 8));
  NPPH_TRY(mem.get(&dE, k * R * 8));
  NPPH_TRY(mem.get(&dI, k * 4));
  NPPH_TRY(mem.get(&dF, k * 4));
  NPPH_TRY(mem.get(&fbuf, refine_factor_bytes(prec, k, N)));
  // Host data moves only by contiguous copies (a pitched copy from pageable
  // host memory is issued row by row): each request's A and B land packed in
  // a staging buffer and are padded on the device; X goes back the same way.
  std::vector<size_t> soff(k + 1, 0);
  for (int64_t i = 0; i < k; ++i)
    soff[i + 1] = soff[i] + (size_t)g[i]->n * g[i]->n + (size_t)g[i]->n * g[i]->nrhs;
  double* stage;
  NPPH_TRY(mem.get(&stage, soff[k] * 8));
  RefineFactors f;
  refine_factor_carve(prec, k, N, fbuf, f);
  NPPH_HIP(hipMemsetAsync(f.anorm, 0, (size_t)k * 8, s));
  for (int64_t i = 0; i < k; ++i) {
    Request* r = g[i];
    const int n = r->n;
    const size_t nn = (size_t)n * n, nb = (size_t)n * r->nrhs;
    double* sa = stage + soff[i];
    double* sb = sa + nn;
    NPPH_HIP(hipMemcpyAsync(sa, r->A, nn * 8, hipMemcpyHostToDevice, s));
    NPPH_HIP(hipMemcpyAsync(sb, r->B, nb * 8, hipMemcpyHostToDevice, s));
    const dim3 grid((unsigned)ceil_div((int64_t)N * 64, 256));
    if (prec == NPPH_PREC_FP64)
      hipLaunchKernelGGL(pad_prep_kernel<double>, grid, dim3(256), 0, s, N, n, sa, dA + i * NN, f.lu64 + i * NN,
                         f.anorm + i);
    else
      hipLaunchKernelGGL(pad_prep_kernel<float>, grid, dim3(256), 0, s, N, n, sa, dA + i * NN, f.lu32 + i * NN,
                         f.anorm + i);
    hipLaunchKernelGGL(pad_block_kernel, dim3((unsigned)ceil_div((int64_t)NR, 256)), dim3(256), 0, s, N, R, n,
                       r->nrhs, sb, dB + i * NR);
    NPPH_HIP(hipGetLastError());
  }
  NPPH_TRY(refine_factor(sv.opts, dA, f, s, /*prepped=*/true));
  NPPH_TRY(refine_solve(sv.opts, dA, f, R, dB, dX, dE, dI, dF, s));
  std::vector<int32_t> iters(k), info(k);
  std::vector<double> berr((size_t)k * R);
  for (int64_t i = 0; i < k; ++i) {
    Request* r = g[i];
    const size_t nb = (size_t)r->n * r->nrhs;
    double* sx = stage + soff[i] + (size_t)r->n * r->n;  // B's staging, no longer read
    hipLaunchKernelGGL(unpad_block_kernel, dim3((unsigned)ceil_div((int64_t)nb, 256)), dim3(256), 0, s, r->n,
                       r->nrhs, R, dX + i * NR, sx);
    NPPH_HIP(hipGetLastError());
    NPPH_HIP(hipMemcpyAsync(r->X, sx, nb * 8, hipMemcpyDeviceToHost, s));
  }
  NPPH_HIP(hipMemcpyAsync(berr.data(), dE, berr.size() * 8, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(iters.data(), dI, k * 4, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(info.data(), dF, k * 4, hipMemcpyDeviceToHost, s));
  // resident copies of kept slices (group) or the buffers themselves (alone)
  const size_t rbytes = NN * 8 + refine_factor_bytes(prec, 1, N);
  std::vector<std::unique_ptr<Resident>> keep;
  AsyncBufs kmem(s);
  for (int64_t i = 0; i < k; ++i) {
    Request* r = g[i];
    if (!(r->keep && r->key)) continue;
    std::unique_ptr<Resident> res(new Resident);
    res->key = r->key;
    res->n = r->n;
    res->N = N;
    res->bytes = rbytes;
    if (k == 1) {
      res->A = dA;
      res->fbuf = fbuf;
      res->f = f;
    } else {
      NPPH_TRY(kmem.get(&res->A, NN * 8));
      NPPH_TRY(kmem.get(&res->fbuf, refine_factor_bytes(prec, 1, N)));
      refine_factor_carve(prec, 1, N, res->fbuf, res->f);
      const RefineFactors& d = res->f;
      NPPH_HIP(hipMemcpyAsync(res->A, dA + i * NN, NN * 8, hipMemcpyDeviceToDevice, s));
      if (prec == NPPH_PREC_FP64)
        NPPH_HIP(hipMemcpyAsync(d.lu64, f.lu64 + i * NN, NN * 8, hipMemcpyDeviceToDevice, s));
      else
        NPPH_HIP(hipMemcpyAsync(d.lu32, f.lu32 + i * NN, NN * 4, hipMemcpyDeviceToDevice, s));
      NPPH_HIP(hipMemcpyAsync(d.ipiv, f.ipiv + i * N, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
      NPPH_HIP(hipMemcpyAsync(d.perm, f.perm + i * N, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
      NPPH_HIP(hipMemcpyAsync(d.info, f.info + i, 4, hipMemcpyDeviceToDevice, s));
      NPPH_HIP(hipMemcpyAsync(d.anorm, f.anorm + i, 8, hipMemcpyDeviceToDevice, s));
    }
    keep.push_back(std::move(res));
  }
  NPPH_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < k; ++i) {
    g[i]->iters = iters[i];
    g[i]->info = info[i];
    if (g[i]->berr) std::copy(berr.begin() + i * R, berr.begin() + i * R + g[i]->nrhs, g[i]->berr);
  }
  for (auto& res : keep) {
    kmem.release(res->A);
    kmem.release(res->fbuf);
    if (k == 1) {
      mem.release(dA);
      mem.release(fbuf);
    }
    kept.push_back(res.release());
  }
  mem.free_all();
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}

void worker_main(Server* sv) {
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return;
  WsSet* ws = ws_create(s);
  ws_bind(ws);
  for (;;) {
    std::vector<Request*> group;
    Resident* reuse = nullptr;
    std::vector<Resident*> freed;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      for (;;) {
        if (sv->stop) break;
        if (sv->queue.empty()) { sv->cv.wait(lk); continue; }
        Request* h = sv->queue.front();
        auto it = h->key ? sv->cache.find(h->key) : sv->cache.end();
        const bool is_reuse = it != sv->cache.end() && it->second->n == h->n;
        h->bytes = request_bytes(*sv, is_reuse ? it->second->N : h->N, h->nrhs, is_reuse);
        if (!is_reuse && !h->A && sv->pending_keep.count(h->key)) {
          // its factors are still being computed: hold it until they are resident
          sv->queue.pop_front();
          sv->deferred[h->key].push_back(h);
          continue;
        }
        if (!is_reuse && !h->A) {
          // the resident factors it referred to were evicted (or never kept)
          sv->queue.pop_front();
          h->status = NPPH_ERR_NOT_FOUND;
          h->done = true;
          h->t_start = h->t_done = Clock::now();
          sv->cv.notify_all();
          continue;
        }
        if (h->bytes > sv->budget) {
          sv->queue.pop_front();
          h->status = NPPH_ERR_BUDGET;
          h->done = true;
          h->t_start = h->t_done = Clock::now();
          if (h->A && h->keep && h->key) release_pending(*sv, h->key);
          sv->cv.notify_all();
          continue;
        }
        // a reuse request pins its factors first, so making room cannot evict them
        if (is_reuse) it->second->refs++;
        if (!make_room(*sv, h->bytes, freed)) {
          if (is_reuse) it->second->refs--;
          sv->cv.wait(lk);
          continue;
        }
        sv->queue.pop_front();
        sv->used += h->bytes;
        group.push_back(h);
        if (is_reuse) {
          reuse = it->second;
          sv->lru.erase(reuse->lru);
          sv->lru.push_front(reuse->key);
          reuse->lru = sv->lru.begin();
        } else {
          // coalesce queued requests that factorise, in the same padded order
          // (requests that keep their factors included); the group's bytes
          // grow with its widest right-hand side
          int R = h->nrhs;
          size_t gbytes = h->bytes;
          for (auto q = sv->queue.begin(); q != sv->queue.end() && (int)group.size() < sv->cfg.max_batch;) {
            Request* c = *q;
            auto cr = c->key ? sv->cache.find(c->key) : sv->cache.end();
            const bool c_reuse = cr != sv->cache.end() && cr->second->n == c->n;
            if (c->N == h->N && c->A && !c_reuse) {
              const int R2 = std::max(R, c->nrhs);
              const size_t total = request_bytes(*sv, h->N, R2, false) * (group.size() + 1);
              if (sv->used + (total - gbytes) <= sv->budget) {
                sv->used += total - gbytes;
                gbytes = total;
                R = R2;
                group.push_back(c);
                q = sv->queue.erase(q);
                continue;
              }
            }
            ++q;
          }
          for (auto* r : group) r->bytes = gbytes / group.size();
        }
        sv->peak = std::max(sv->peak, sv->used);
        break;
      }
      if (sv->stop && group.empty()) break;
    }
    for (Resident* r : freed) free_resident(r, s);
    freed.clear();
    const auto t0 = Clock::now();
    for (auto* r : group) { r->t_start = t0; r->batched = group.size() > 1; }
    int st;
    std::vector<Resident*> kept;
    if (reuse) st = serve_reuse(*sv, group[0], reuse, s);
    else st = serve_group(*sv, group, s, kept);
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (reuse) reuse->refs--;
      for (auto* r : group) {
        sv->used -= r->bytes;
        r->status = st;
        r->done = true;
        r->t_done = Clock::now();
        ++sv->completed;
        if (r->iters < 0) ++sv->fallbacks;
      }
      if (group.size() > 1) sv->batched += (int64_t)group.size();
      for (auto& kr : kept) {
        auto old = sv->cache.find(kr->key);
        if (old != sv->cache.end()) {
          // replace an older factorisation under the same key (not in use: refs checked)
          Resident* o = old->second;
          if (o->refs == 0) {
            sv->lru.erase(o->lru);
            sv->used -= o->bytes;
            sv->cache.erase(old);
            freed.push_back(o);
          }
        }
        if (!sv->cache.count(kr->key)) {
          sv->lru.push_front(kr->key);
          kr->lru = sv->lru.begin();
          sv->cache[kr->key] = kr;
          sv->used += kr->bytes;
          sv->peak = std::max(sv->peak, sv->used);
          kr = nullptr;
        }
      }
      for (auto* r : group)
        if (r->A && r->keep && r->key) release_pending(*sv, r->key);
      sv->cv.notify_all();
    }
    for (Resident* kr : kept)
      if (kr) free_resident(kr, s);
    for (Resident* r : freed) free_resident(r, s);
    (void)hipStreamSynchronize(s);
  }
  ws_bind(nullptr);
  ws_destroy(ws);
  (void)hipStreamDestroy(s);
}

}  // namespace
}  // namespace npph

using namespace npph;

extern "C" {

int npph_server_create(const npph_serve_config* cfg, void** srv) {
  NPPH_TRY(ensure_init());
  if (!cfg || !srv || cfg->budget_bytes <= 0 || cfg->workers < 1 || cfg->workers > 16) {
    set_error("server: bad config");
    return NPPH_ERR_INVALID;
  }
  hipMemPool_t pool;
  int dev = 0;
  NPPH_HIP(hipGetDevice(&dev));
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t thr = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  Server* sv = new Server;
  sv->cfg = *cfg;
  if (sv->cfg.max_batch < 1) sv->cfg.max_batch = 1;
  sv->opts = npph_refine_opts{cfg->precision, cfg->max_iter, cfg->tol, cfg->fallback_fp64, cfg->bf16_terms};
  sv->budget = (size_t)cfg->budget_bytes;
  for (int i = 0; i < cfg->workers; ++i) sv->workers.emplace_back(worker_main, sv);
  *srv = sv;
  return NPPH_OK;
}

int npph_server_destroy(void* srv) {
  Server* sv = static_cast<Server*>(srv);
  if (!sv) return NPPH_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    for (auto* r : sv->queue) {
      r->status = NPPH_ERR_INVALID;
      r->done = true;
    }
    sv->queue.clear();
    for (auto& kv : sv->deferred)
      for (auto* r : kv.second) {
        r->status = NPPH_ERR_INVALID;
        r->done = true;
      }
    sv->deferred.clear();
    sv->cv.notify_all();
  }
  for (auto& t : sv->workers) t.join();
  for (auto& kv : sv->cache) free_resident(kv.second, default_stream());
  (void)hipStreamSynchronize(default_stream());
  for (auto& kv : sv->reqs) delete kv.second;
  delete sv;
  return NPPH_OK;
}

int npph_server_submit(void* srv, int64_t n, int64_t nrhs, const double* A, const double* B, double* X,
                       double* berr, uint64_t key, int keep, int64_t* req_id) {
  Server* sv = static_cast<Server*>(srv);
  if (!sv || n < 1 || n > max_blocked_n() || nrhs < 1 || !B || !X || !req_id) {
    set_error("server submit: bad arguments");
    return NPPH_ERR_INVALID;
  }
  Request* r = new Request;
  r->n = (int)n;
  r->N = bucket_order((int)n);
  r->nrhs = (int)nrhs;
  r->A = A;
  r->B = B;
  r->X = X;
  r->berr = berr;
  r->key = key;
  r->keep = keep;
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (!A && !(key && (sv->cache.count(key) || sv->pending_keep.count(key)))) {
    delete r;
    set_error("server submit: A is null and key %llu is neither resident nor being factorised",
              (unsigned long long)key);
    return NPPH_ERR_NOT_FOUND;
  }
  if (A && keep && key) sv->pending_keep[key]++;
  r->id = sv->next_id++;
  sv->reqs[r->id] = r;
  sv->queue.push_back(r);
  sv->cv.notify_all();
  *req_id = r->id;
  return NPPH_OK;
}

int npph_server_wait(void* srv, int64_t req_id, int32_t* iters, int32_t* info, double* latency_ms,
                     double* queue_ms) {
  Server* sv = static_cast<Server*>(srv);
  if (!sv) { set_error("server wait: null server"); return NPPH_ERR_INVALID; }
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(req_id);
  if (it == sv->reqs.end()) { set_error("server wait: unknown request %ld", (long)req_id); return NPPH_ERR_NOT_FOUND; }
  Request* r = it->second;
  sv->cv.wait(lk, [&] { return r->done; });
  const int st = r->status;
  if (iters) *iters = r->iters;
  if (info) *info = r->info;
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
  sv->reqs.erase(it);
  delete r;
  if (st != NPPH_OK && st != NPPH_ERR_BUDGET) set_error("server: request failed with status %d", st);
  if (st == NPPH_ERR_BUDGET) set_error("server: request exceeds the device byte budget");
  return st;
}

int npph_server_stats(void* srv, int64_t* completed, int64_t* resident, double* bytes_in_use, double* peak_bytes,
                      int64_t* batched, int64_t* fallbacks, int64_t* evictions) {
  Server* sv = static_cast<Server*>(srv);
  if (!sv) { set_error("server stats: null server"); return NPPH_ERR_INVALID; }
  std::lock_guard<std::mutex> lk(sv->mu);
  if (completed) *completed = sv->completed;
  if (resident) *resident = (int64_t)sv->cache.size();
  if (bytes_in_use) *bytes_in_use = (double)sv->used;
  if (peak_bytes) *peak_bytes = (double)sv->peak;
  if (batched) *batched = sv->batched;
  if (fallbacks) *fallbacks = sv->fallbacks;
  if (evictions) *evictions = sv->evictions;
  return NPPH_OK;
}

// Replays a request trace: request i is submitted t_arrive[i] seconds after the
// call starts (a submitter thread sleeps until then), results are collected as
// they complete. Outputs per request: latency_ms (submit -> results on host),
// iters, info; X[i] and berr[i] receive the results. wall_s = first submit to
// last completion.
int npph_server_replay(void* srv, int64_t count, const double* t_arrive, const int32_t* n, const int32_t* nrhs,
                       const double* const* A, const double* const* B, double* const* X, double* const* berr,
                       const uint64_t* key, const int32_t* keep, double* latency_ms, int32_t* iters, int32_t* info,
                       int32_t* status, double* wall_s) {
  Server* sv = static_cast<Server*>(srv);
  if (!sv || count < 0) { set_error("server replay: bad arguments"); return NPPH_ERR_INVALID; }
  std::vector<int64_t> ids(count, -1);
  std::vector<std::atomic<int>> submitted(count);
  for (auto& a : submitted) a.store(0);
  std::atomic<int> submit_err{NPPH_OK};
  const auto t0 = Clock::now();
  std::thread submitter([&] {
    for (int64_t i = 0; i < count; ++i) {
      std::this_thread::sleep_until(t0 + std::chrono::duration_cast<Clock::duration>(
                                             std::chrono::duration<double>(t_arrive[i])));
      int64_t id = -1;
      const int st = npph_server_submit(srv, n[i], nrhs[i], A[i], B[i], X[i], berr ? berr[i] : nullptr,
                                        key ? key[i] : 0, keep ? keep[i] : 0, &id);
      ids[i] = id;
      if (st != NPPH_OK) submit_err.store(st);
      submitted[i].store(st == NPPH_OK ? 1 : 2);
    }
  });
  Clock::time_point last = t0;
  for (int64_t i = 0; i < count; ++i) {
    while (submitted[i].load() == 0) std::this_thread::sleep_for(std::chrono::microseconds(50));
    if (submitted[i].load() == 2) {
      if (status) status[i] = NPPH_ERR_INVALID;
      continue;
    }
    double lat = 0.0;
    const int st = npph_server_wait(srv, ids[i], iters ? iters + i : nullptr, info ? info + i : nullptr, &lat,
                                    nullptr);
    if (status) status[i] = st;
    if (latency_ms) latency_ms[i] = lat;
    last = std::max(last, Clock::now());
  }
  submitter.join();
  if (wall_s) *wall_s = std::chrono::duration<double>(last - t0).count();
  return submit_err.load();
}

}  // extern "C"
