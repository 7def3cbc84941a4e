// Device probe, error reporting, workspace arena and memory/stream plumbing.
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

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
  std::mutex ws_mu;
  Workspace ws[kWsCount];
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
};

namespace {
std::atomic<uint64_t> g_ws_gen{0};
thread_local WsSet* t_ws = nullptr;
}

WsSet* ws_create(hipStream_t owner) {
  WsSet* w = new WsSet;
  w->owner = owner;
  return w;
}

void ws_destroy(WsSet* w) {
  if (!w) return;
  if (w->owner) (void)hipStreamSynchronize(w->owner);
  for (auto& x : w->ws)
    if (x.ptr) (void)hipFree(x.ptr);
  delete w;
}

namespace {
thread_local int t_plain = 0;
}
PlainIssue::PlainIssue() { ++t_plain; }
PlainIssue::~PlainIssue() { --t_plain; }
bool plain_issue() { return t_plain > 0; }

void ws_bind(WsSet* w) { t_ws = w; }
WsSet* ws_current() { return t_ws; }
uint64_t workspace_generation() { return g_ws_gen.load(); }

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

int workspace(int slot, size_t bytes, void** out) {
  Context& c = ctx();
  if (slot < 0 || slot >= kWsCount) { set_error("bad workspace slot %d", slot); return NPPH_ERR_INVALID; }
  std::lock_guard<std::mutex> g(c.ws_mu);
  Workspace& w = t_ws ? t_ws->ws[slot] : c.ws[slot];
  if (w.bytes < bytes) {
    g_ws_gen.fetch_add(1);
    if (w.ptr) {
      // The previous buffer may still be in use by queued work: on the owner
      // stream of a bound set, on any stream for the process-wide set.
      if (t_ws && t_ws->owner) NPPH_HIP(hipStreamSynchronize(t_ws->owner));
      else NPPH_HIP(hipDeviceSynchronize());
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
