// Device probe, error reporting, workspace arena and memory/stream plumbing.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "common.hpp"

namespace npph {

namespace {
thread_local char t_err[512] = "";

struct Context {
  std::once_flag once;
  int status = NPPH_ERR_NO_DEVICE;
  int device = 0;
  int cus = 0;
  char name[256] = "";
  hipStream_t stream = nullptr;
};
Context& ctx() {
  static Context c;
  return c;
}

void do_init() {
  Context& c = ctx();
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    set_error("no HIP device present");
    (void)hipGetLastError();
    return;
  }
  if (hipGetDevice(&c.device) != hipSuccess) { set_error("hipGetDevice failed"); return; }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, c.device) != hipSuccess) { set_error("hipGetDeviceProperties failed"); return; }
  std::snprintf(c.name, sizeof c.name, "%s", p.gcnArchName);
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    set_error("device %s is not gfx950; this backend is built for MI355X only", p.gcnArchName);
    return;
  }
  c.cus = p.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return;
  }
  c.status = NPPH_OK;
}
}  // namespace

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(t_err, sizeof t_err, fmt, ap);
  va_end(ap);
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("%s:%d %s: %s", file, line, what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? NPPH_ERR_OOM : NPPH_ERR_HIP;
}

int ensure_init() {
  Context& c = ctx();
  std::call_once(c.once, do_init);
  if (c.status != NPPH_OK && t_err[0] == '\0') set_error("no usable gfx950 device");
  return c.status;
}

hipStream_t default_stream() { return ctx().stream; }
int cu_count() { return ctx().cus; }

struct WsSet {
  Workspace ws[kWsCount];
  hipStream_t owner = nullptr;
  bool priv = false;
  std::mutex mu;  // held by the CallScope that has the set bound
};

namespace {
std::atomic<uint64_t> g_ws_gen{0};
thread_local WsSet* t_ws = nullptr;
thread_local ScratchArena* t_arena = nullptr;
thread_local int t_plain = 0;

// Per-stream sets, created on first use and kept until the stream is released.
std::mutex g_sets_mu;
std::unordered_map<hipStream_t, WsSet*> g_sets;

WsSet* stream_set(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_sets_mu);
  WsSet*& w = g_sets[s];
  if (!w) {
    w = new WsSet;
    w->owner = s;
  }
  return w;
}

void free_set_buffers(WsSet* w) {
  if (w->owner) (void)hipStreamSynchronize(w->owner);
  for (auto& x : w->ws)
    if (x.ptr) {
      (void)hipFree(x.ptr);
      x.ptr = nullptr;
      x.bytes = 0;
    }
}
}  // namespace

WsSet* ws_create(hipStream_t owner) {
  WsSet* w = new WsSet;
  w->owner = owner;
  w->priv = true;
  return w;
}

void ws_destroy(WsSet* w) {
  if (!w) return;
  free_set_buffers(w);
  delete w;
}

void ws_bind(WsSet* w) { t_ws = w; }
WsSet* ws_bound() { return t_ws; }
bool ws_private() { return t_ws != nullptr && t_ws->priv; }

void ws_release_stream(hipStream_t s) {
  WsSet* w = nullptr;
  {
    std::lock_guard<std::mutex> g(g_sets_mu);
    auto it = g_sets.find(s);
    if (it == g_sets.end()) return;
    w = it->second;
    g_sets.erase(it);
  }
  {
    std::lock_guard<std::mutex> g(w->mu);  // no call is using it any more
    free_set_buffers(w);
  }
  g_ws_gen.fetch_add(1);
  delete w;
}

CallScope::CallScope(hipStream_t s) {
  if (t_ws != nullptr) return;  // a private set, or an enclosing scope
  set_ = stream_set(s);
  set_->mu.lock();
  t_ws = set_;
}

CallScope::~CallScope() {
  if (!set_) return;
  t_ws = nullptr;
  set_->mu.unlock();
}

void arena_bind(ScratchArena* a) { t_arena = a; }
bool arena_bound() { return t_arena != nullptr; }

PlainIssue::PlainIssue() { ++t_plain; }
PlainIssue::~PlainIssue() { --t_plain; }
bool plain_issue() { return t_plain > 0 || t_arena != nullptr; }

uint64_t workspace_generation() { return g_ws_gen.load(); }

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

int workspace(int slot, size_t bytes, void** out) {
  if (slot < 0 || slot >= kWsCount) { set_error("bad workspace slot %d", slot); return NPPH_ERR_INVALID; }
  if (ScratchArena* a = t_arena) {
    if (a->len[slot] >= bytes && a->len[slot] > 0) {
      *out = a->base + a->off[slot];
      return NPPH_OK;
    }
    const size_t want = (bytes + 255) & ~(size_t)255;
    if (a->used + want > a->bytes) {
      set_error("scratch: %zu more bytes needed for slot %d, %zu of %zu in use", want, slot, a->used, a->bytes);
      return NPPH_ERR_INVALID;
    }
    a->off[slot] = a->used;
    a->len[slot] = want;
    a->used += want;
    *out = a->base + a->off[slot];
    return NPPH_OK;
  }
  WsSet* set = t_ws;
  if (!set) { set_error("internal: workspace requested outside a call scope"); return NPPH_ERR_INVALID; }
  Workspace& w = set->ws[slot];
  if (w.bytes < bytes) {
    g_ws_gen.fetch_add(1);
    if (w.ptr) {
      // the previous buffer may still be in use by work queued on the set's stream
      NPPH_HIP(hipStreamSynchronize(set->owner));
      NPPH_HIP(hipFree(w.ptr));
      w.ptr = nullptr;
      w.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    NPPH_HIP(hipMalloc(&w.ptr, want));
    w.bytes = want;
  }
  *out = w.ptr;
  return NPPH_OK;
}

}  // namespace npph

using namespace npph;

static hipStream_t as_stream(void* s) { return s ? static_cast<hipStream_t>(s) : default_stream(); }

extern "C" {

int npph_init(void) { return ensure_init(); }
int npph_available(void) { return ensure_init() == NPPH_OK ? 1 : 0; }
const char* npph_last_error(void) { return t_err; }
const char* npph_device_name(void) {
  ensure_init();
  return ctx().name;
}
int npph_cu_count(void) {
  ensure_init();
  return ctx().cus;
}
int npph_device_id(void) {
  ensure_init();
  return ctx().device;
}

int npph_malloc(void** ptr, uint64_t bytes) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return NPPH_OK;
}
int npph_free(void* ptr) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipFree(ptr));
  return NPPH_OK;
}
int npph_memcpy_htod(void* dst, const void* src, uint64_t bytes, void* stream) {
  NPPH_TRY(ensure_init());
  hipStream_t s = as_stream(stream);
  NPPH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}
int npph_memcpy_dtoh(void* dst, const void* src, uint64_t bytes, void* stream) {
  NPPH_TRY(ensure_init());
  hipStream_t s = as_stream(stream);
  NPPH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  NPPH_HIP(hipStreamSynchronize(s));
  return NPPH_OK;
}
int npph_memcpy_dtod(void* dst, const void* src, uint64_t bytes, void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return NPPH_OK;
}
int npph_memset(void* dst, int value, uint64_t bytes, void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
  return NPPH_OK;
}
int npph_stream_create(void** stream) {
  NPPH_TRY(ensure_init());
  hipStream_t s;
  NPPH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return NPPH_OK;
}
int npph_stream_destroy(void* stream) {
  NPPH_TRY(ensure_init());
  ws_release_stream(static_cast<hipStream_t>(stream));
  NPPH_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return NPPH_OK;
}
int npph_stream_sync(void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipStreamSynchronize(as_stream(stream)));
  return NPPH_OK;
}
int npph_device_sync(void) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipDeviceSynchronize());
  return NPPH_OK;
}
int npph_event_create(void** ev) {
  NPPH_TRY(ensure_init());
  hipEvent_t e;
  NPPH_HIP(hipEventCreate(&e));
  *ev = e;
  return NPPH_OK;
}
int npph_event_destroy(void* ev) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return NPPH_OK;
}
int npph_event_record(void* ev, void* stream) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), as_stream(stream)));
  return NPPH_OK;
}
int npph_event_elapsed_ms(void* start, void* stop, float* ms) {
  NPPH_TRY(ensure_init());
  NPPH_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  NPPH_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return NPPH_OK;
}

}  // extern "C"
