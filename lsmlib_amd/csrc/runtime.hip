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

// n contiguous scalars from a (naturally aligned) set to v: 16-byte stores over
// the 16-byte-aligned body, one thread per vector; scalar stores for the at
// most W - 1 elements before the body and after it.
constexpr int kFillThreads = 256;
constexpr int kFillW = 16 / (int)sizeof(real_t);
typedef real_t fill_vec_t __attribute__((ext_vector_type(kFillW)));

__global__ void __launch_bounds__(kFillThreads) fill_contiguous_kernel(real_t* a, long long head, long long nvec,
                                                                      long long n, real_t v) {
  const long long t = blockIdx.x * (long long)kFillThreads + threadIdx.x;
  if (t < nvec) reinterpret_cast<fill_vec_t*>(a + head)[t] = fill_vec_t(v);
  const long long tail0 = head + nvec * kFillW;
  if (t < head) a[t] = v;
  if (t < n - tail0) a[tail0 + t] = v;
}
}  // namespace

bool fill_contiguous(real_t* a, long long n, real_t v, const char* who) {
  if (n <= 0) return true;
  const long long mis = (long long)(((uintptr_t)a) & 15) / (long long)sizeof(real_t);
  long long head = mis ? kFillW - mis : 0;
  if (head > n) head = n;
  const long long nvec = (n - head) / kFillW;
  const long long blocks = (nvec + kFillThreads - 1) / kFillThreads;
  hipLaunchKernelGGL(fill_contiguous_kernel, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(kFillThreads), 0,
                     g_stream, a, head, nvec, n, v);
  return check_hip(hipGetLastError(), who);
}

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

bool report_null_array(const char* who, const char* names, int position) {
  // the position-th name of "a, b, c"
  const char* s = names;
  for (int q = 0; q < position && s; q++) {
    s = strchr(s, ',');
    if (s) s++;
  }
  char name[64] = "?";
  if (s) {
    while (*s == ' ') s++;
    size_t n = strcspn(s, ", ");
    if (n >= sizeof(name)) n = sizeof(name) - 1;
    memcpy(name, s, n);
    name[n] = 0;
  }
  set_error(kErrArg, "%s: array argument %s is NULL (nothing was done)", who, name);
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
