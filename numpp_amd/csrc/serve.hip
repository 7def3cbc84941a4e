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
// - Continuous batching: when a worker takes the head request it also takes
//   queued one-shot requests of the same order n (up to max_batch, budget
//   permitting) and solves them as one stack; right-hand sides are padded with
//   zero columns to the widest request. Every kernel works per system and per
//   column, so a result is bitwise the same alone and in the mix.
// - Resident factors: a request with keep = 1 leaves (A, factors) on the device
//   under its key; a later request with the same key skips the factorisation
//   and only runs the refined solve with new right-hand sides. A reuse request
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
  int n = 0;
  double* A = nullptr;  // FP64 copy (residual)
  void* fbuf = nullptr;  // factors
  RefineFactors f;
  size_t bytes = 0;
  int refs = 0;
  std::list<uint64_t>::iterator lru;
};

struct Request {
  int64_t id = 0;
  int n = 0, nrhs = 0;
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

size_t request_bytes(const Server& sv, int n, int nrhs, bool reuse) {
  size_t b = 2 * (size_t)n * nrhs * 8 + ws_estimate(n, nrhs);
  if (!reuse) b += (size_t)n * n * 8 + refine_factor_bytes(sv.opts.precision, 1, n);
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

// Solves one request with resident factors.
int serve_reuse(Server& sv, Request* r, Resident* res, hipStream_t s) {
  AsyncBufs mem(s);
  double *dB, *dX, *dE;
  int32_t *dI, *dF;
  const size_t nb = (size_t)r->n * r->nrhs;
  NPPH_TRY(mem.get(&dB, nb * 8));
  NPPH_TRY(mem.get(&dX, nb * 8));
  NPPH_TRY(mem.get(&dE, (size_t)r->nrhs * 8));
  NPPH_TRY(mem.get(&dI, 4));
  NPPH_TRY(mem.get(&dF, 4));
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

// Factorises and solves a group of same-order requests as one stack. A
// single request with keep = 1 hands its (A, factors) to the cache.
int serve_group(Server& sv, std::vector<Request*>& g, hipStream_t s, Resident** keep_out) {
  const int n = g[0]->n;
  const int64_t k = (int64_t)g.size();
  int R = 1;
  for (auto* r : g) R = std::max(R, r->nrhs);
  const size_t nn = (size_t)n * n, nbR = (size_t)n * R;
  AsyncBufs mem(s);
  double *dA, *dB, *dX, *dE;
  int32_t *dI, *dF;
  void* fbuf;
  NPPH_TRY(mem.get(&dA, k * nn * 8));
  NPPH_TRY(mem.get(&dB, k * nbR * 8));
  NPPH_TRY(mem.get(&dX, k * nbR * 8));
  NPPH_TRY(mem.get(&dE, k * R * 8));
  NPPH_TRY(mem.get(&dI, k * 4));
  NPPH_TRY(mem.get(&dF, k * 4));
  NPPH_TRY(mem.get(&fbuf, refine_factor_bytes(sv.opts.precision, k, n)));
  for (int64_t i = 0; i < k; ++i) {
    Request* r = g[i];
    NPPH_HIP(hipMemcpyAsync(dA + i * nn, r->A, nn * 8, hipMemcpyHostToDevice, s));
    if (r->nrhs == R) {
      NPPH_HIP(hipMemcpyAsync(dB + i * nbR, r->B, nbR * 8, hipMemcpyHostToDevice, s));
    } else {
      NPPH_HIP(hipMemsetAsync(dB + i * nbR, 0, nbR * 8, s));
      NPPH_HIP(hipMemcpy2DAsync(dB + i * nbR, (size_t)R * 8, r->B, (size_t)r->nrhs * 8, (size_t)r->nrhs * 8, n,
                                hipMemcpyHostToDevice, s));
    }
  }
  RefineFactors f;
  refine_factor_carve(sv.opts.precision, k, n, fbuf, f);
  NPPH_TRY(refine_factor(sv.opts, dA, f, s));
  NPPH_TRY(refine_solve(sv.opts, dA, f, R, dB, dX, dE, dI, dF, s));
  std::vector<int32_t> iters(k), info(k);
  std::vector<double> berr((size_t)k * R);
  for (int64_t i = 0; i < k; ++i) {
    Request* r = g[i];
    NPPH_HIP(hipMemcpy2DAsync(r->X, (size_t)r->nrhs * 8, dX + i * nbR, (size_t)R * 8, (size_t)r->nrhs * 8, n,
                              hipMemcpyDeviceToHost, s));
  }
  NPPH_HIP(hipMemcpyAsync(berr.data(), dE, berr.size() * 8, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(iters.data(), dI, k * 4, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipMemcpyAsync(info.data(), dF, k * 4, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  for (int64_t i = 0; i < k; ++i) {
    g[i]->iters = iters[i];
    g[i]->info = info[i];
    if (g[i]->berr) std::copy(berr.begin() + i * R, berr.begin() + i * R + g[i]->nrhs, g[i]->berr);
  }
  if (k == 1 && g[0]->keep && g[0]->key != 0) {
    Resident* res = new Resident;
    res->key = g[0]->key;
    res->n = n;
    res->A = dA;
    res->fbuf = fbuf;
    res->f = f;
    res->bytes = nn * 8 + refine_factor_bytes(sv.opts.precision, 1, n);
    mem.release(dA);
    mem.release(fbuf);
    *keep_out = res;
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
        h->bytes = request_bytes(*sv, h->n, h->nrhs, is_reuse);
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
        if (!make_room(*sv, h->bytes, freed)) { sv->cv.wait(lk); continue; }
        sv->queue.pop_front();
        sv->used += h->bytes;
        group.push_back(h);
        if (is_reuse) {
          reuse = it->second;
          reuse->refs++;
          sv->lru.erase(reuse->lru);
          sv->lru.push_front(reuse->key);
          reuse->lru = sv->lru.begin();
        } else if (!(h->keep && h->key)) {
          // coalesce queued one-shot requests of the same order
          for (auto q = sv->queue.begin(); q != sv->queue.end() && (int)group.size() < sv->cfg.max_batch;) {
            Request* c = *q;
            if (c->n == h->n && c->A && !(c->keep && c->key) && !(c->key && sv->cache.count(c->key))) {
              const size_t b = request_bytes(*sv, c->n, c->nrhs, false);
              if (sv->used + b <= sv->budget) {
                c->bytes = b;
                sv->used += b;
                group.push_back(c);
                q = sv->queue.erase(q);
                continue;
              }
            }
            ++q;
          }
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
    Resident* kept = nullptr;
    if (reuse) st = serve_reuse(*sv, group[0], reuse, s);
    else st = serve_group(*sv, group, s, &kept);
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
      if (kept) {
        auto old = sv->cache.find(kept->key);
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
        if (!sv->cache.count(kept->key)) {
          sv->lru.push_front(kept->key);
          kept->lru = sv->lru.begin();
          sv->cache[kept->key] = kept;
          sv->used += kept->bytes;
          sv->peak = std::max(sv->peak, sv->used);
          kept = nullptr;
        }
      }
      for (auto* r : group)
        if (r->A && r->keep && r->key) release_pending(*sv, r->key);
      sv->cv.notify_all();
    }
    if (kept) free_resident(kept, s);
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
