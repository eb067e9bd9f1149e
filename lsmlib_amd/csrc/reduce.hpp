// reduce.hpp -- grid-wide reductions over a box for the CFL and integral
// entry points.  Wave-level reduction with cross-lane shuffles, then LDS across
// the block's waves, then either an atomic max on the IEEE bit pattern (all
// reduced values are >= 0, for which bit order equals numeric order) or
// per-block partial sums combined by one block in a fixed order with
// Neumaier compensation, so sums are deterministic run to run.
#pragma once

#include <cstring>

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

template <class Op>
__global__ void __launch_bounds__(kThreads) max_kernel(Box b, Op op, unsigned long long* acc) {
  const long long n = b.volume();
  const int nx = b.nx(), ny = b.ny();
  double m = 0.0;
  for (long long t = blockIdx.x * (long long)kThreads + threadIdx.x; t < n; t += (long long)gridDim.x * kThreads) {
    const int i = b.ilo + (int)(t % nx);
    const int j = b.jlo + (int)((t / nx) % ny);
    const int k = b.klo + (int)(t / ((long long)nx * ny));
    const double v = op(i, j, k);
    m = v > m ? v : m;
  }
  m = wave_max(m);
  __shared__ double part[kThreads / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; w++) m = part[w] > m ? part[w] : m;
    atomicMax(acc, (unsigned long long)__double_as_longlong(m));
  }
}

template <class Op>
__global__ void __launch_bounds__(kThreads) sum_kernel(Box b, Op op, double* partials) {
  const long long n = b.volume();
  const int nx = b.nx(), ny = b.ny();
  Pair p{0.0, 0.0};
  for (long long t = blockIdx.x * (long long)kThreads + threadIdx.x; t < n; t += (long long)gridDim.x * kThreads) {
    const int i = b.ilo + (int)(t % nx);
    const int j = b.jlo + (int)((t / nx) % ny);
    const int k = b.klo + (int)(t / ((long long)nx * ny));
    p = add(p, op(i, j, k));
  }
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

inline int blocks_for(const Box& b) {
  long long n = (b.volume() + kThreads - 1) / kThreads;
  return (int)(n < kMaxBlocks ? (n < 1 ? 1 : n) : kMaxBlocks);
}

// Blocking: the result is needed on the host (the reference returns it through a pointer).
template <class Op>
bool box_max(const char* who, const Box& b, const Op& op, double* result) {
  unsigned long long* acc = (unsigned long long*)workspace(sizeof(unsigned long long));
  if (!acc) return false;
  if (!check_hip(hipMemsetAsync(acc, 0, sizeof(unsigned long long), stream()), who)) return false;
  hipLaunchKernelGGL(max_kernel<Op>, dim3(blocks_for(b)), dim3(kThreads), 0, stream(), b, op, acc);
  if (!check_hip(hipGetLastError(), who)) return false;
  unsigned long long h = 0;
  if (!check_hip(hipMemcpyAsync(&h, acc, sizeof(h), hipMemcpyDeviceToHost, stream()), who)) return false;
  if (!check_hip(hipStreamSynchronize(stream()), who)) return false;
  memcpy(result, &h, sizeof(double));
  return true;
}

template <class Op>
bool box_sum(const char* who, const Box& b, const Op& op, double* result) {
  const int nb = blocks_for(b);
  double* ws = (double*)workspace(sizeof(double) * (2 * nb + 2));
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
