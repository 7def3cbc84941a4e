// Per-kernel-class HIP-event timing (see prof.hpp).
#include <atomic>
#include <mutex>
#include <vector>

#include "prof.hpp"

namespace npph {
namespace {
struct Rec {
  int kind;
  double flops, bytes;
  hipEvent_t a, b;
};
std::atomic<int> g_on{0};
std::mutex g_mu;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool prof_enabled() { return g_on.load(std::memory_order_relaxed) != 0; }

void prof_begin(int kind, double flops, double bytes, hipStream_t s, void** token) {
  if (stream_capturing(s)) return;  // no event records inside a captured graph
  if (arena_bound()) return;         // allocation-free calls create no events
  std::lock_guard<std::mutex> g(g_mu);
  Rec r{kind, flops, bytes, take_event(), take_event()};
  if (!r.a || !r.b) return;
  (void)hipEventRecord(r.a, s);
  g_recs.push_back(r);
  *token = reinterpret_cast<void*>(g_recs.size());  // 1-based index
}

void prof_end(void* token, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_mu);
  const size_t i = reinterpret_cast<size_t>(token) - 1;
  if (i < g_recs.size()) (void)hipEventRecord(g_recs[i].b, s);
}

}  // namespace npph

using namespace npph;

extern "C" {

int npph_prof_enable(int on) {
  NPPH_TRY(ensure_init());
  g_on.store(on ? 1 : 0);
  return NPPH_OK;
}

int npph_prof_reset(void) {
  NPPH_TRY(ensure_init());
  std::lock_guard<std::mutex> g(g_mu);
  for (auto& r : g_recs) {
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_recs.clear();
  return NPPH_OK;
}

int npph_prof_get(int kind, double* ms, double* flops, double* bytes, int64_t* launches) {
  NPPH_TRY(ensure_init());
  if (kind < 0 || kind >= kProfKinds) { set_error("prof: bad kind %d", kind); return NPPH_ERR_INVALID; }
  std::lock_guard<std::mutex> g(g_mu);
  double t = 0, f = 0, b = 0;
  int64_t n = 0;
  for (auto& r : g_recs) {
    if (r.kind != kind) continue;
    NPPH_HIP(hipEventSynchronize(r.b));
    float e = 0.f;
    NPPH_HIP(hipEventElapsedTime(&e, r.a, r.b));
    t += e;
    f += r.flops;
    b += r.bytes;
    ++n;
  }
  if (ms) *ms = t;
  if (flops) *flops = f;
  if (bytes) *bytes = b;
  if (launches) *launches = n;
  return NPPH_OK;
}

}  // extern "C"
