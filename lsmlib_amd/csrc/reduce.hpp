// reduce.hpp -- grid-wide reductions over a box for the CFL and integral
// entry points.  Wave-level reduction with cross-lane shuffles, then LDS across
// the block's waves, then either an atomic max on the IEEE bit pattern (all
// reduced values are >= 0, for which bit order equals numeric order) or
// per-block partial sums combined by one block in a fixed order with
// Neumaier compensation, so sums are deterministic run to run.
#pragma once

#include <cstring>
#include <type_traits>

#include "lsmhip_core.hpp"

namespace lsmhip {
namespace reduce {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 4096;

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// Neumaier pair (s, c): the value is s + c.
struct Pair {
  double s, c;
};
__device__ __forceinline__ Pair add(Pair a, double x) {
  const double t = a.s + x;
  if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
  else a.c += (x - t) + a.s;
  a.s = t;
  return a;
}
__device__ __forceinline__ Pair add(Pair a, Pair b) {
  a = add(a, b.s);
  a.c += b.c;
  return a;
}
__device__ __forceinline__ Pair wave_sum(Pair v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    Pair w{__shfl_xor(v.s, o, 64), __shfl_xor(v.c, o, 64)};
    v = add(v, w);
  }
  return v;
}

// Cells are visited one row (fixed j, k) per wave at a time: the lanes stride
// along i, so loads are coalesced and the row index is split into (j, k) once
// per row rather than once per cell; each lane issues kLoads loads before it
// accumulates, into kUnroll independent accumulators (no serial dependence
// between them), which the caller combines in a fixed order.  Out-of-row
// slots contribute the identity (0 for the maxima, whose terms are >= 0; 0 for
// the sums, which leaves a Neumaier pair as is).
constexpr int kUnroll = 4;
// Slots loaded per lane and pass: twice kUnroll for 4-byte reals, so that a
// wave has as many bytes in flight as with doubles (slot u accumulates into
// acc[u % kUnroll]).
constexpr int kLoads = sizeof(real_t) == 4 ? 2 * kUnroll : kUnroll;

//
// An op with `In load(i, j, k)` and `double eval(const In&, i, j, k)` has its
// unconditional loads split from its arithmetic: the kLoads loads of a lane
// go out together (at clamped, in-box indices), and only then are the terms
// evaluated.  Called whole, an op's branches would keep the compiler from
// issuing the next slot's loads before this slot's term, one memory round
// trip per slot.
template <class Op, class = void>
struct split_op : std::false_type {};
template <class Op>
struct split_op<Op, std::void_t<typename Op::In>> : std::true_type {};

template <class Op, class Acc, class F>
__device__ __forceinline__ void for_rows(const Box& b, const Op& op, Acc (&acc)[kUnroll], F combine,
                                         double identity = 0.0) {
  const int nx = b.nx(), ny = b.ny();
  const long long rows = (long long)ny * b.nz();
  const int lane = threadIdx.x & 63;
  const long long nwaves = (long long)gridDim.x * (kThreads / 64);
  for (long long r = blockIdx.x * (long long)(kThreads / 64) + (threadIdx.x >> 6); r < rows; r += nwaves) {
    const int j = b.jlo + (int)(r % ny), k = b.klo + (int)(r / ny);
    for (int i0 = lane; i0 < nx; i0 += 64 * kLoads) {
      double v[kLoads];
      if constexpr (split_op<Op>::value) {
        typename Op::In in[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; u++) in[u] = op.load(b.ilo + min(i0 + 64 * u, nx - 1), j, k);
#pragma unroll
        for (int u = 0; u < kLoads; u++) {
          const int i = i0 + 64 * u;
          v[u] = i < nx ? op.eval(in[u], b.ilo + i, j, k) : identity;
        }
      } else {
#pragma unroll
        for (int u = 0; u < kLoads; u++) {
          const int i = i0 + 64 * u;
          v[u] = i < nx ? op(b.ilo + i, j, k) : identity;
        }
      }
#pragma unroll
      for (int u = 0; u < kLoads; u++) acc[u % kUnroll] = combine(acc[u % kUnroll], v[u]);
    }
  }
}

// Maximum of op over the box as the IEEE bits of a non-negative double.
// PRESENT: op returns a negative value for a cell that does not take part;
// the accumulator then holds 0 when no cell took part and the bits of the
// maximum plus one otherwise (still in numeric order).
template <class Op, bool PRESENT = false>
__global__ void __launch_bounds__(kThreads) max_kernel(Box b, Op op, unsigned long long* acc) {
  constexpr double kNone = PRESENT ? -1.0 : 0.0;
  double ms[kUnroll] = {kNone, kNone, kNone, kNone};
  static_assert(kUnroll == 4, "initialiser above");
  for_rows(b, op, ms, [](double a, double v) { return v > a ? v : a; }, kNone);
  double m = ms[0];
#pragma unroll
  for (int u = 1; u < kUnroll; u++) m = ms[u] > m ? ms[u] : m;
  m = wave_max(m);
  __shared__ double part[kThreads / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; w++) m = part[w] > m ? part[w] : m;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
    if (!PRESENT) atomicMax(acc, bits);
    else if (m >= 0.0) atomicMax(acc, bits + 1ull);
  }
}

template <class Op>
__global__ void __launch_bounds__(kThreads) sum_kernel(Box b, Op op, double* partials) {
  Pair ps_[kUnroll] = {};
  for_rows(b, op, ps_, [](Pair a, double v) { return add(a, v); });
  Pair p = ps_[0];
#pragma unroll
  for (int u = 1; u < kUnroll; u++) p = add(p, ps_[u]);
  p = wave_sum(p);
  __shared__ double ps[kThreads / 64], pc[kThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    ps[threadIdx.x >> 6] = p.s;
    pc[threadIdx.x >> 6] = p.c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Pair q{ps[0], pc[0]};
    for (int w = 1; w < kThreads / 64; w++) q = add(q, Pair{ps[w], pc[w]});
    partials[2 * blockIdx.x] = q.s;
    partials[2 * blockIdx.x + 1] = q.c;
  }
}

// One block: combines nblocks partial pairs in a fixed order.
static __global__ void __launch_bounds__(kThreads) sum_finish_kernel(const double* partials, int nblocks, double* out) {
  Pair p{0.0, 0.0};
  for (int t = threadIdx.x; t < nblocks; t += kThreads) p = add(p, Pair{partials[2 * t], partials[2 * t + 1]});
  p = wave_sum(p);
  __shared__ double ps[kThreads / 64], pc[kThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    ps[threadIdx.x >> 6] = p.s;
    pc[threadIdx.x >> 6] = p.c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Pair q{ps[0], pc[0]};
    for (int w = 1; w < kThreads / 64; w++) q = add(q, Pair{ps[w], pc[w]});
    out[0] = q.s + q.c;
  }
}

// One wave per row until the block cap.
inline int blocks_for(const Box& b) {
  long long n = ((long long)b.ny() * b.nz() + kThreads / 64 - 1) / (kThreads / 64);
  return (int)(n < kMaxBlocks ? (n < 1 ? 1 : n) : kMaxBlocks);
}

// Blocking: the result is needed on the host (the reference returns it through a pointer).
// PRESENT: cells for which op is negative do not take part, and *result is
// -1 when no cell did.
template <class Op, bool PRESENT = false>
bool box_max(const char* who, const Box& b, const Op& op, double* result) {
  unsigned long long* acc = (unsigned long long*)workspace(sizeof(unsigned long long));
  if (!acc) return false;
  if (!check_hip(hipMemsetAsync(acc, 0, sizeof(unsigned long long), stream()), who)) return false;
  hipLaunchKernelGGL((max_kernel<Op, PRESENT>), dim3(blocks_for(b)), dim3(kThreads), 0, stream(), b, op, acc);
  if (!check_hip(hipGetLastError(), who)) return false;
  unsigned long long h = 0;
  if (!check_hip(hipMemcpyAsync(&h, acc, sizeof(h), hipMemcpyDeviceToHost, stream()), who)) return false;
  if (!check_hip(hipStreamSynchronize(stream()), who)) return false;
  if (PRESENT) {
    if (h == 0) {
      *result = -1.0;
      return true;
    }
    h -= 1;
  }
  memcpy(result, &h, sizeof(double));
  return true;
}

// buf: caller-provided device memory of at least 2 * kMaxBlocks + 2 doubles
// for the partials and the result, or null for the library workspace.
template <class Op>
bool box_sum(const char* who, const Box& b, const Op& op, double* result, double* buf = nullptr) {
  const int nb = blocks_for(b);
  double* ws = buf ? buf : (double*)workspace(sizeof(double) * (2 * nb + 2));
  if (!ws) return false;
  hipLaunchKernelGGL(sum_kernel<Op>, dim3(nb), dim3(kThreads), 0, stream(), b, op, ws);
  hipLaunchKernelGGL(sum_finish_kernel, dim3(1), dim3(kThreads), 0, stream(), ws, nb, ws + 2 * nb);
  if (!check_hip(hipGetLastError(), who)) return false;
  double h = 0;
  if (!check_hip(hipMemcpyAsync(&h, ws + 2 * nb, sizeof(h), hipMemcpyDeviceToHost, stream()), who)) return false;
  if (!check_hip(hipStreamSynchronize(stream()), who)) return false;
  *result = h;
  return true;
}

}  // namespace reduce
}  // namespace lsmhip
