// lsmhip_core.hpp -- shared types and runtime plumbing of the HIP level-set library.
//
// Arrays follow the reference's conventions: column-major, each array
// addressed from the first element of its own ghost box (i fastest), boxes as
// inclusive (ilo, ihi, jlo, jhi, klo, khi).  Array arguments of the exported
// entry points are device pointers; scalars and boxes stay host pointers as in
// the reference signatures.
//
// The library is compiled once per precision: LSMHIP_REAL=double gives
// liblsmhip.so (the FP64 drop-in, operation order of the reference, no FMA
// contraction), LSMHIP_REAL=float gives liblsmhip_f32.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#ifndef LSMHIP_REAL
#define LSMHIP_REAL double
#endif

typedef LSMHIP_REAL real_t;

// The public headers (include/) declare every exported entry point; including
// them here makes the compiler check each definition against its declaration.
#define LSMLIB_REAL LSMHIP_REAL
#include "lsmhip_all.h"

namespace lsmhip {

constexpr bool kDouble = sizeof(real_t) == 8;
// Template constants of the reference build (src/CMakeLists.txt:10-20):
// zero tolerance 1e-11 in double precision and 1e-5 in single precision,
// tiny nonzero number 1e-36 in both.
constexpr real_t kZeroTol = kDouble ? real_t(1.e-11) : real_t(1.e-5);
constexpr real_t kTiny = real_t(1.e-36);
constexpr real_t kBig = real_t(1.e10);

struct Box {
  int ilo, ihi, jlo, jhi, klo, khi;
  __host__ __device__ int nx() const { return ihi - ilo + 1; }
  __host__ __device__ int ny() const { return jhi - jlo + 1; }
  __host__ __device__ int nz() const { return khi - klo + 1; }
  __host__ __device__ bool empty() const { return ihi < ilo || jhi < jlo || khi < klo; }
  __host__ __device__ bool has(int i, int j, int k) const {
    return i >= ilo && i <= ihi && j >= jlo && j <= jhi && k >= klo && k <= khi;
  }
  bool contains(const Box& b) const {
    return b.empty() || (b.ilo >= ilo && b.ihi <= ihi && b.jlo >= jlo && b.jhi <= jhi && b.klo >= klo && b.khi <= khi);
  }
  Box widen(int wx, int wy, int wz) const {
    return Box{ilo - wx, ihi + wx, jlo - wy, jhi + wy, klo - wz, khi + wz};
  }
  __host__ __device__ long long volume() const { return empty() ? 0 : (long long)nx() * ny() * nz(); }
};

// Device view of one array: element (i,j,k) of an array whose ghost box
// starts at (ilo,jlo,klo) with row stride sy and plane stride sz.
template <class T>
struct View {
  T* p;
  int ilo, jlo, klo, sy, sz;
  __device__ __forceinline__ T& operator()(int i, int j, int k) const {
    return p[(i - ilo) + sy * (j - jlo) + sz * (k - klo)];
  }
  __device__ __forceinline__ int offset(int i, int j, int k) const {
    return (i - ilo) + sy * (j - jlo) + sz * (k - klo);
  }
};

template <class T>
inline View<T> view(T* p, const Box& b) {
  return View<T>{p, b.ilo, b.jlo, b.klo, b.nx(), b.nx() * b.ny()};
}

inline Box box_of(const int* ilo, const int* ihi, const int* jlo, const int* jhi, const int* klo, const int* khi) {
  return Box{*ilo, *ihi, *jlo, *jhi, *klo, *khi};
}

// ----- error side channel and stream state (runtime.hip) ---------------------
enum ErrorCode {
  kOk = 0,
  kErrBox = 1,    // a fill/include box is not inside an array's ghost box
  kErrHip = 2,    // a HIP runtime call failed
  kErrArg = 3,    // invalid scalar argument (null pointer, bad index, ...)
  kErrIndex = 4,  // an index-list entry (or its stencil) left an array's ghost box
  kErrSize = 5,   // an array exceeds the 2^31-element addressing limit
};

void set_error(int code, const char* fmt, ...);
hipStream_t stream();
bool synchronous();
// Ends an entry point: records a launch error and, in synchronous mode (the
// default, matching the reference's completion semantics), waits for the stream.
bool finish(const char* who);  // launch status, and completion in synchronous mode
bool check_hip(hipError_t e, const char* who);
// Sets n contiguous scalars from a to v on the library's stream (a kernel, so a
// caller may capture it in a graph; launch status only).
bool fill_contiguous(real_t* a, long long n, real_t v, const char* who);
// Device scratch owned by the library, grown on demand (never shrinks).
void* workspace(size_t bytes);
// Per-device int flag that index-list kernels raise when an index or a stencil
// neighbour lies outside an array's ghost box (the point is then skipped).
int* device_flag();
// Reads and clears the device flag (synchronises the stream).
int take_device_flag(const char* who);

bool check_box(const Box& outer, const Box& inner, const char* who, const char* what);
bool check_size(const Box& b, const char* who);

// Device-array arguments.  The reference's routines dereference every array
// they are given; here a NULL array would be dereferenced by a kernel, and a
// kernel that faults can end the whole device, not just the caller.  Every
// exported entry point therefore tests its device-array arguments on the host
// before it launches anything: a NULL one sets kErrArg, naming the argument,
// and the call does nothing.  names is the comma-separated list of the
// arguments' names (the macro's stringised argument list).
bool report_null_array(const char* who, const char* names, int position);
template <class... P>
inline bool arrays_present(const char* who, const char* names, P... p) {
  const void* const a[] = {static_cast<const void*>(p)...};
  for (int q = 0; q < (int)sizeof...(P); q++)
    if (a[q] == nullptr) return report_null_array(who, names, q);
  return true;
}
#define LSMHIP_REQUIRE_ARRAYS(...) \
  do {                             \
    if (!::lsmhip::arrays_present(__func__, #__VA_ARGS__, __VA_ARGS__)) return; \
  } while (0)
#define LSMHIP_REQUIRE_ARRAYS_OR(rv, ...) \
  do {                                    \
    if (!::lsmhip::arrays_present(__func__, #__VA_ARGS__, __VA_ARGS__)) return rv; \
  } while (0)

// Launch shape for kernels that map one thread to one cell of a box: kBX x kBY
// tiles of (i, j) in one plane k.  Workgroups are dealt to the kXcds XCDs
// round robin in flattened-id order, and each XCD has its own L2.  The grid's x
// extent is kXcds x (tiles along i), so block (bx, by, bz) runs on XCD bx % kXcds
// whatever by and bz are; the macro gives XCD x the tiles of rows
// [x * gridDim.y, (x + 1) * gridDim.y) of tiles along j, in every plane.  Tiles
// that are neighbours along i, along j (inside a slab) and along k then share
// one L2, so the cache lines at tile edges (rows start inside a line: fill
// boxes sit three cells into the ghost box) and the stencils' neighbour rows
// and planes are fetched once per XCD instead of once per neighbouring tile.
// Tiles past the box (the last slab is partly empty) return at once.
constexpr int kBX = 64, kBY = 4, kXcds = 8;
inline dim3 grid_for(const Box& b) {
  const unsigned ntx = (unsigned)((b.nx() + kBX - 1) / kBX), nty = (unsigned)((b.ny() + kBY - 1) / kBY);
  return dim3(kXcds * ntx, (nty + kXcds - 1) / kXcds, (unsigned)b.nz());
}
inline dim3 block2d() { return dim3(kBX, kBY, 1); }

// NK cells along k per thread, (i, j, k0) .. (i, j, k0 + NK - 1) (nk of them
// inside the box), on the same tiles and XCD slabs as LSMHIP_CELL: for
// kernels that issue the loads of all their cells before any store, so each
// wave keeps NK times the bytes in flight.  The stencils take NK = 2; the
// pointwise kernels take kPointNK, two cells in FP32 and one in FP64 (where
// one cell already keeps enough bytes in flight, and two measured slower).
constexpr int kPointNK = sizeof(real_t) == 4 ? 2 : 1;
inline dim3 grid_for_cells(const Box& b, int NK) {
  const dim3 g = grid_for(b);
  return dim3(g.x, g.y, (unsigned)((b.nz() + NK - 1) / NK));
}
inline dim3 grid_for_pair(const Box& b) { return grid_for_cells(b, 2); }
#define LSMHIP_CELLS(b, NK)                                                                       \
  const int i = (b).ilo + (int)(blockIdx.x / kXcds) * kBX + (int)threadIdx.x;                    \
  const int j = (b).jlo + (int)((blockIdx.x % kXcds) * gridDim.y + blockIdx.y) * kBY + (int)threadIdx.y; \
  const int k0 = (b).klo + (NK) * (int)blockIdx.z;                                                \
  const int nk = (b).khi - k0 + 1 < (NK) ? (b).khi - k0 + 1 : (NK);                               \
  if (i > (b).ihi || j > (b).jhi) return
#define LSMHIP_CELL_PAIR(b) LSMHIP_CELLS(b, 2)

// The cells of LSMHIP_CELLS (nk <= 2): load(k) for each (every load issued
// before any store), then store(k, loaded) for each.
template <class L, class S>
__device__ __forceinline__ void pair_cells(int k0, int nk, const L& load, const S& store) {
  const auto a = load(k0);
  if (nk > 1) {
    const auto b = load(k0 + 1);
    store(k0, a);
    store(k0 + 1, b);
  } else {
    store(k0, a);
  }
}

#define LSMHIP_CELL(b)                                                                            \
  const int i = (b).ilo + (int)(blockIdx.x / kXcds) * kBX + (int)threadIdx.x;                    \
  const int j = (b).jlo + (int)((blockIdx.x % kXcds) * gridDim.y + blockIdx.y) * kBY + (int)threadIdx.y; \
  const int k = (b).klo + (int)blockIdx.z;                                                        \
  if (i > (b).ihi || j > (b).jhi) return

// Exported-signature helpers: six box bounds passed by pointer.
#define LSMHIP_BOXP(n) \
  const int *ilo_##n, const int *ihi_##n, const int *jlo_##n, const int *jhi_##n, const int *klo_##n, const int *khi_##n
#define LSMHIP_BOX(n) ::lsmhip::box_of(ilo_##n, ihi_##n, jlo_##n, jhi_##n, klo_##n, khi_##n)

template <class T>
__device__ __forceinline__ T sq(T x) {
  return x * x;
}
template <class T>
__device__ __forceinline__ T maxr(T a, T b) {
  return a > b ? a : b;
}
template <class T>
__device__ __forceinline__ T minr(T a, T b) {
  return a < b ? a : b;
}

}  // namespace lsmhip

#define LSMHIP_EXPORT extern "C" __attribute__((visibility("default")))
