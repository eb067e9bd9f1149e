// local.hpp -- machinery shared by the 3D and 2D narrow-band ("local") kernels.
//
// One thread per index-list entry l: (i,j,k) = (index_x, index_y, index_z)(l),
// where an index array declared index_x(nlo:nhi) in the reference is a
// pointer to its element nlo; 2D lists have no index_z (k = 0, and 2D boxes
// span k = 0..0).  Points whose narrow_band mark (a signed byte) exceeds the
// fill-box mark are skipped, as in the reference.  Each index and stencil
// neighbour is checked against its array's ghost box; a point that would leave
// it is skipped and raises the device flag (lsmhip_get_last_error then
// reports kErrIndex) where the reference would read or write out of bounds.
#pragma once

#include <type_traits>

#include <cstring>

#include "lsmhip_core.hpp"
#include "reduce.hpp"

namespace lsmhip {
namespace {

constexpr int kT = 256;

struct IndexList {
  const int *ix, *iy, *iz;  // iz null: a 2D list (k = 0)
  int lo, hi, base;
  __device__ __forceinline__ void at(int l, int& i, int& j, int& k) const {
    i = ix[l - base];
    j = iy[l - base];
    k = iz ? iz[l - base] : 0;
  }
};

struct Mask {
  const unsigned char* p;
  Box b;
  int sy, sz;
  signed char mark;
  int filter;  // 0: no mark filter
  __device__ __forceinline__ signed char at(int i, int j, int k) const {
    return (signed char)p[(i - b.ilo) + sy * (j - b.jlo) + sz * (k - b.klo)];
  }
};

Mask make_mask(const unsigned char* nb, const Box& b, const unsigned char* mark, int filter = 1) {
  return Mask{nb, b, b.nx(), b.nx() * b.ny(), mark ? (signed char)*mark : (signed char)0, filter};
}
Mask no_mask() { return Mask{nullptr, Box{0, -1, 0, -1, 0, -1}, 0, 0, 0, 0}; }

// Checked array view: element access plus ghost-box membership tests.
template <class T>
struct CV {
  View<T> v;
  Box b;
  __device__ __forceinline__ T& operator()(int i, int j, int k) const { return v(i, j, k); }
  __device__ __forceinline__ bool has(int i, int j, int k) const { return b.has(i, j, k); }
  // (i,j,k) with its +-w neighbours along each axis (the corners of the star's box)
  __device__ __forceinline__ bool star(int i, int j, int k, int w) const {
    return b.has(i - w, j - w, k - w) && b.has(i + w, j + w, k + w);
  }
  // the same in the (i, j) plane of a 2D array
  __device__ __forceinline__ bool star2(int i, int j, int k, int w) const {
    return b.has(i - w, j - w, k) && b.has(i + w, j + w, k);
  }
};
template <class T>
CV<T> cv(T* p, const Box& b) {
  return CV<T>{view(p, b), b};
}

// Grids over list entries dealt to the 8 XCDs in contiguous ranges: the
// hardware sends block b to XCD b % 8, so block b takes the entries of virtual
// block (b % 8) * (grid / 8) + b / 8 and each XCD's L2 sees one contiguous
// stretch of the list (neighbouring points of a band are neighbours in the
// list), instead of every eighth block.  Grids are a multiple of 8 blocks.
// Used by the stencil-heavy band kernels and the reductions, where it cuts the
// HBM fetches of the 19-point cross about 2.4x; the one-pass per-point
// kernels keep the plain order, which measured faster for them.
inline unsigned xcd_grid(long long entries) {
  const long long nb = (entries + kT - 1) / kT;
  return (unsigned)(8 * ((nb + 7) / 8));
}
__device__ __forceinline__ long long xcd_block() {
  return (long long)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
}
__device__ __forceinline__ long long xcd_entry() { return xcd_block() * kT + threadIdx.x; }

// Ops return false when an access would leave an array's ghost box.
template <class Op>
__global__ void __launch_bounds__(kT) points_kernel(IndexList L, Mask M, Op op, int* flag) {
  const long long t = (long long)blockIdx.x * kT + threadIdx.x;
  if (t > (long long)L.hi - L.lo) return;
  const int l = L.lo + (int)t;
  int i, j, k;
  L.at(l, i, j, k);
  if (M.filter) {
    if (!M.b.has(i, j, k)) {
      *flag = 1;
      return;
    }
    if (M.at(i, j, k) > M.mark) return;
  }
  if (!op(i, j, k)) *flag = 1;
}

template <class Op>
void launch_points(const char* who, const IndexList& L, const Mask& M, const Op& op) {
  if (L.hi < L.lo) return;
  int* flag = device_flag();
  if (!flag) return;
  const long long n = (long long)L.hi - L.lo + 1;
  hipLaunchKernelGGL(points_kernel<Op>, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, stream(), L, M, op, flag);
  check_hip(hipGetLastError(), who);
}

void finish_local(const char* who) {
  finish(who);
  if (synchronous()) take_device_flag(who);
}

IndexList list(const int* ix, const int* iy, const int* iz, int lo, int hi, int base) {
  return IndexList{ix, iy, iz, lo, hi, base};
}

// ----- undivided differences over index levels (lsm3dComputeDnLOCAL, :43-138) ------------------
struct DnOp {
  CV<real_t> dn;
  CV<const real_t> dm;
  Mask m;
  signed char markD;
  int di, dj, dk;
  real_t sgn;
  __device__ bool operator()(int i, int j, int k) const {
    if (!dn.has(i, j, k) || !m.b.has(i, j, k)) return false;
    if (m.at(i, j, k) <= markD) {
      if (!dm.has(i, j, k) || !dm.has(i - di, j - dj, k - dk)) return false;
      dn(i, j, k) = sgn * (dm(i, j, k) - dm(i - di, j - dj, k - dk));
    } else {
      dn(i, j, k) = kBig;
    }
    return true;
  }
};

// n odd: D(i) = Dm(i) - Dm(i-1); n even: D(i) = -(Dm(i) - Dm(i+1)) along axis d.
void dn_local(const char* who, real_t* dn, const Box& dnb, const real_t* dm, const Box& dmb, int n, int d,
              const IndexList& L, const Mask& nbm, signed char markD) {
  const int off = n % 2 == 1 ? 1 : -1;
  DnOp op{cv(dn, dnb), cv(dm, dmb), nbm, markD, d == 0 ? off : 0, d == 1 ? off : 0, d == 2 ? off : 0,
          n % 2 == 1 ? real_t(1) : real_t(-1)};
  launch_points(who, L, no_mask(), op);
}

// ----- reductions over index lists ----------------------------------------------------------------
// An op of the list maxima is called as op(t, i, j, k, r) when it takes the
// entry's position t in the list (0 for entry lo), else as op(i, j, k, r).
template <class Op>
__device__ __forceinline__ auto call_max_op(const Op& op, long long t, int i, int j, int k, double& r, int)
    -> decltype(op(t, i, j, k, r)) {
  return op(t, i, j, k, r);
}
template <class Op>
__device__ __forceinline__ bool call_max_op(const Op& op, long long, int i, int j, int k, double& r, long) {
  return op(i, j, k, r);
}

// An op with marked(t, i, j, k, M, r) reads the point's mark itself, with its
// own loads, and returns true without a value where the mark exceeds M.mark.
template <class Op, class = void>
struct reads_own_mark : std::false_type {};
template <class Op>
struct reads_own_mark<Op, std::void_t<decltype(&Op::marked)>> : std::true_type {};

// Max over admissible points of non-negative values (-1 stands for "no
// value": a point that is not admissible, or an op that declines it).
template <class Op>
__global__ void __launch_bounds__(kT) points_max_kernel(IndexList L, Mask M, Op op, double* part, int* flag) {
  // one thread per list entry, one partial per block (no atomics: a single
  // accumulator would serialise its updates)
  const long long t = xcd_entry();
  double m = -1.0;  // below every admissible value
  if (t <= (long long)L.hi - L.lo) {
    int i, j, k;
    L.at(L.lo + (int)t, i, j, k);
    if (!M.b.has(i, j, k)) {
      *flag = 1;
    } else if constexpr (reads_own_mark<Op>::value) {
      if (!op.marked(t, i, j, k, M, m)) *flag = 1;
    } else {
      if (M.at(i, j, k) <= M.mark && !call_max_op(op, t, i, j, k, m, 0)) *flag = 1;
    }
  }
  m = reduce::wave_max(m);
  __shared__ double w[kT / 64];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kT / 64; q++) m = w[q] > m ? w[q] : m;
    part[xcd_block()] = m;
  }
}

constexpr int kFinish = 1024;  // threads of the one-block finishing reduction
__global__ void __launch_bounds__(kFinish) max_finish_kernel(const double* part, long long n, double* out) {
  double m = -1.0;
  for (long long t = threadIdx.x; t < n; t += kFinish) m = part[t] > m ? part[t] : m;
  m = reduce::wave_max(m);
  __shared__ double w[kFinish / 64];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kFinish / 64; q++) m = w[q] > m ? w[q] : m;
    *out = m;
  }
}

// Max of op's values over admissible points: *any is false when no point is
// admissible (or every value was below zero).
template <class Op>
bool points_max(const char* who, const IndexList& L, const Mask& M, const Op& op, double* result, bool* any) {
  *any = false;
  if (L.hi < L.lo) return true;
  const long long n = (long long)L.hi - L.lo + 1;
  const long long nb = xcd_grid(n);
  double* part = (double*)workspace(sizeof(double) * (size_t)(nb + 1));
  int* flag = device_flag();
  if (!part || !flag) return false;
  hipLaunchKernelGGL(points_max_kernel<Op>, dim3((unsigned)nb), dim3(kT), 0, stream(), L, M, op, part, flag);
  hipLaunchKernelGGL(max_finish_kernel, dim3(1), dim3(kFinish), 0, stream(), part, nb, part + nb);
  if (!check_hip(hipGetLastError(), who)) return false;
  double h = -1.0;
  if (!check_hip(hipMemcpyAsync(&h, part + nb, sizeof(h), hipMemcpyDeviceToHost, stream()), who)) return false;
  if (!check_hip(hipStreamSynchronize(stream()), who)) return false;
  take_device_flag(who);
  if (!(h >= 0.0)) return true;
  *any = true;
  *result = h;
  return true;
}

// Compensated FP64 sum and count over admissible points (one partial per block,
// combined in a fixed order by reduce::sum_finish_kernel).
template <class Op>
__global__ void __launch_bounds__(kT) points_sum_kernel(IndexList L, Mask M, Op op, double* partials, int* counts,
                                                        int* flag) {
  const long long tt = xcd_entry();
  const int l = L.lo + (int)tt;
  reduce::Pair p{0.0, 0.0};
  int c = 0;
  if (tt <= (long long)L.hi - L.lo) {
    int i, j, k;
    L.at(l, i, j, k);
    double v = 0.0;
    if (!M.b.has(i, j, k)) *flag = 1;
    else if (M.at(i, j, k) <= M.mark) {
      if (op(i, j, k, v)) {
        p = reduce::add(p, v);
        c = 1;
      } else {
        *flag = 1;
      }
    }
  }
  p = reduce::wave_sum(p);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  __shared__ double ps[kT / 64], pc[kT / 64];
  __shared__ int cs[kT / 64];
  if ((threadIdx.x & 63) == 0) {
    ps[threadIdx.x >> 6] = p.s;
    pc[threadIdx.x >> 6] = p.c;
    cs[threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    reduce::Pair q{ps[0], pc[0]};
    int cc = cs[0];
    for (int w = 1; w < kT / 64; w++) {
      q = reduce::add(q, reduce::Pair{ps[w], pc[w]});
      cc += cs[w];
    }
    const long long vb = xcd_block();
    partials[2 * vb] = q.s;
    partials[2 * vb + 1] = q.c;
    counts[vb] = cc;
  }
}

__global__ void count_finish_kernel(const int* counts, int n, long long* out) {
  long long c = 0;
  for (int t = threadIdx.x; t < n; t += 64) c += counts[t];
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if (threadIdx.x == 0) *out = c;
}

// Mean of op's values over admissible points (0 when there are none): an FP64
// compensated sum in a fixed order and a count, one partial per block.
template <class Op>
void points_mean(const char* who, const IndexList& L, const Mask& M, const Op& op, real_t* ave) {
  *ave = real_t(0);
  if (L.hi < L.lo) return;
  const long long n = (long long)L.hi - L.lo + 1;
  const int nb_ = (int)xcd_grid(n);
  char* ws = (char*)workspace(sizeof(double) * (2 * (size_t)nb_ + 2) + sizeof(int) * (size_t)nb_ + 16);
  int* flag = device_flag();
  if (!ws || !flag) return;
  double* partials = (double*)ws;
  double* total = partials + 2 * nb_;
  long long* count = (long long*)(total + 1);
  int* counts = (int*)(count + 1);
  hipLaunchKernelGGL(points_sum_kernel<Op>, dim3(nb_), dim3(kT), 0, stream(), L, M, op, partials, counts, flag);
  hipLaunchKernelGGL(reduce::sum_finish_kernel, dim3(1), dim3(reduce::kThreads), 0, stream(), partials, nb_, total);
  hipLaunchKernelGGL(count_finish_kernel, dim3(1), dim3(64), 0, stream(), counts, nb_, count);
  if (!check_hip(hipGetLastError(), who)) return;
  double s = 0;
  long long c = 0;
  if (!check_hip(hipMemcpyAsync(&s, total, sizeof(s), hipMemcpyDeviceToHost, stream()), who)) return;
  if (!check_hip(hipMemcpyAsync(&c, count, sizeof(c), hipMemcpyDeviceToHost, stream()), who)) return;
  if (!check_hip(hipStreamSynchronize(stream()), who)) return;
  take_device_flag(who);
  *ave = c > 0 ? real_t(s / (double)c) : real_t(0);
}

}  // namespace
}  // namespace lsmhip
