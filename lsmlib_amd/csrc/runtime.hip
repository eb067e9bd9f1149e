// runtime.hip -- error side channel, stream selection and library-owned device
// scratch.  The reference's entry points return nothing and check nothing
// (they run serially on host memory); this library keeps their void
// signatures and reports problems through lsmhip_get_last_error() instead.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "lsmhip_core.hpp"

namespace lsmhip {
namespace {
thread_local int g_err = kOk;
thread_local char g_msg[512] = "";
hipStream_t g_stream = nullptr;
bool g_sync = true;

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ws[kMaxDevices] = {};
size_t g_ws_bytes[kMaxDevices] = {};
int* g_flag[kMaxDevices] = {};

int current_device() {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices) return 0;
  return d;
}
}  // namespace

void set_error(int code, const char* fmt, ...) {
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
}

hipStream_t stream() { return g_stream; }
bool synchronous() { return g_sync; }

bool check_hip(hipError_t e, const char* who) {
  if (e == hipSuccess) return true;
  set_error(kErrHip, "%s: %s", who, hipGetErrorString(e));
  return false;
}

bool finish(const char* who) {
  if (!check_hip(hipGetLastError(), who)) return false;
  return !g_sync || check_hip(hipStreamSynchronize(g_stream), who);
}

bool check_box(const Box& outer, const Box& inner, const char* who, const char* what) {
  if (outer.contains(inner)) return true;
  set_error(kErrBox, "%s: %s box [%d,%d]x[%d,%d]x[%d,%d] needs [%d,%d]x[%d,%d]x[%d,%d]", who, what, outer.ilo,
            outer.ihi, outer.jlo, outer.jhi, outer.klo, outer.khi, inner.ilo, inner.ihi, inner.jlo, inner.jhi,
            inner.klo, inner.khi);
  return false;
}

bool check_size(const Box& b, const char* who) {
  if (b.volume() < (1LL << 31)) return true;
  set_error(kErrSize, "%s: array of %lld elements exceeds the 2^31 addressing limit", who, b.volume());
  return false;
}

void* workspace(size_t bytes) {
  std::lock_guard<std::mutex> lock(g_mu);
  int d = current_device();
  if (bytes <= g_ws_bytes[d]) return g_ws[d];
  if (g_ws[d]) {
    (void)hipStreamSynchronize(g_stream);
    (void)hipFree(g_ws[d]);
    g_ws[d] = nullptr;
    g_ws_bytes[d] = 0;
  }
  size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
  if (!check_hip(hipMalloc(&g_ws[d], want), "workspace")) return nullptr;
  g_ws_bytes[d] = want;
  return g_ws[d];
}

int* device_flag() {
  std::lock_guard<std::mutex> lock(g_mu);
  int d = current_device();
  if (!g_flag[d]) {
    if (!check_hip(hipMalloc(&g_flag[d], sizeof(int)), "device_flag")) return nullptr;
    (void)hipMemsetAsync(g_flag[d], 0, sizeof(int), g_stream);
  }
  return g_flag[d];
}

int take_device_flag(const char* who) {
  int* f = device_flag();
  if (!f) return 0;
  int h = 0;
  if (!check_hip(hipMemcpyAsync(&h, f, sizeof(int), hipMemcpyDeviceToHost, g_stream), who)) return 0;
  if (!check_hip(hipStreamSynchronize(g_stream), who)) return 0;
  if (h) {
    (void)hipMemsetAsync(f, 0, sizeof(int), g_stream);
    set_error(kErrIndex, "%s: an index-list point or one of its stencil neighbours lies outside an array's "
              "ghost box (such points were skipped)", who);
  }
  return h;
}

}  // namespace lsmhip

using namespace lsmhip;

// ---------------------------------------------------------------------------
// Library control (no counterpart in the reference; declared in include/lsmhip.h)
// ---------------------------------------------------------------------------
LSMHIP_EXPORT int lsmhip_get_last_error(void) { return g_err; }
LSMHIP_EXPORT const char* lsmhip_get_last_error_string(void) { return g_msg; }
LSMHIP_EXPORT void lsmhip_clear_error(void) {
  g_err = kOk;
  g_msg[0] = 0;
}
LSMHIP_EXPORT void lsmhip_set_stream(void* s, int synchronous) {
  g_stream = (hipStream_t)s;
  g_sync = synchronous != 0;
}
LSMHIP_EXPORT void* lsmhip_get_stream(void) { return (void*)g_stream; }
LSMHIP_EXPORT int lsmhip_real_size(void) { return (int)sizeof(real_t); }
LSMHIP_EXPORT int lsmhip_check_device_errors(void) { return take_device_flag("lsmhip_check_device_errors"); }
LSMHIP_EXPORT int lsmhip_synchronize(void) {
  return check_hip(hipStreamSynchronize(g_stream), "lsmhip_synchronize") ? 0 : -1;
}
