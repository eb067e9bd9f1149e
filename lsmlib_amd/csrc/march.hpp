// march.hpp -- plane-marching launch for the whole-grid kernels whose stencil
// is the 7-point star (one neighbour along each axis).
//
// A block owns a kMX x kMY tile of the (i, j) plane and marches along k
// through a chunk of kChunk planes.  Each thread keeps its own column at k-1,
// k, k+1 in registers and has the next plane's values (own column and halo) in
// flight while the current plane is worked on; the current plane with a
// one-cell x/y halo is staged in LDS,
// so every value of phi is read from memory once per block (plus the halo),
// instead of once per neighbour that uses it.  Tiles are dealt to the XCDs in
// contiguous ranges (block b runs on XCD b % 8), so neighbouring tiles share
// an L2 for their halos.
//
// The operation receives the centre value and its six neighbours; it does the
// same arithmetic as the one-cell-per-thread form, so results are bitwise
// unchanged.
#pragma once

#include "lsmhip_core.hpp"

namespace lsmhip {
namespace march {

constexpr int kMX = 64, kMY = 8, kChunk = 16;

struct Tiles {
  int nti, ntj, ntk, kc;
};

inline Tiles tiles_for(const Box& fb) {
  return Tiles{(fb.nx() + kMX - 1) / kMX, (fb.ny() + kMY - 1) / kMY, (fb.nz() + kChunk - 1) / kChunk, kChunk};
}
inline dim3 grid_for(const Tiles& t) {
  const long long n = (long long)t.nti * t.ntj * t.ntk;
  return dim3((unsigned)(((n + 7) / 8) * 8));
}
inline dim3 block() { return dim3(kMX, kMY, 1); }

// Position (px, py) in the (kMX + 2) x (kMY + 2) plane tile of halo element h:
// the full row above, the full row below, then the left and right cells of
// the kMY interior rows.
__device__ __forceinline__ void halo_pos(int h, int& px, int& py) {
  constexpr int PX = kMX + 2;
  if (h < PX) {
    py = 0;
    px = h;
  } else if (h < 2 * PX) {
    py = kMY + 1;
    px = h - PX;
  } else {
    const int t = h - 2 * PX;
    py = 1 + t / 2;
    px = (t & 1) ? kMX + 1 : 0;
  }
}

// op(i, j, k, c, xm, xp, ym, yp, zm, zp) at every cell of fb; phi must cover
// fb widened by one cell on every side.  (A deeper prefetch, 64 x 4 tiles and
// chunks of 32 or 64 planes measured no faster on gfx950.)
template <class Op>
__global__ void __launch_bounds__(kMX* kMY) star_march_kernel(Box fb, View<const real_t> phi, Tiles tl, Op op) {
  constexpr int PX = kMX + 2, PLANE = PX * (kMY + 2), HALO = PLANE - kMX * kMY, DEPTH = 1;
  __shared__ real_t plane[PLANE];
  static_assert(HALO <= kMX * kMY, "one halo element per thread at most");
  const int tid = threadIdx.y * kMX + threadIdx.x;
  const int nt = tl.nti * tl.ntj * tl.ntk;
  const int per = (nt + 7) / 8;
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= nt) return;
  const int ti = t % tl.nti, tj = (t / tl.nti) % tl.ntj, tk = t / (tl.nti * tl.ntj);
  const int i0 = fb.ilo + ti * kMX, j0 = fb.jlo + tj * kMY;
  const int kb = fb.klo + tk * tl.kc;
  const int ke = min(kb + tl.kc - 1, fb.khi);
  const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
  const bool active = i <= fb.ihi && j <= fb.jhi;
  // the own column is a neighbour of the tile's cells when it lies within one
  // cell of fb (inside phi's box); columns further out are never read
  const bool own_in = i <= fb.ihi + 1 && j <= fb.jhi + 1;
  const int own = own_in ? phi.offset(i, j, kb) : 0;
  int hoff = -1, hslot = 0;
  if (tid < HALO) {
    int px, py;
    halo_pos(tid, px, py);
    const int gi = i0 - 1 + px, gj = j0 - 1 + py;
    hslot = py * PX + px;
    if (gi <= fb.ihi + 1 && gj <= fb.jhi + 1) hoff = phi.offset(gi, gj, kb);
  }
  const int sz = phi.sz;
  const int slot = (threadIdx.y + 1) * PX + threadIdx.x + 1;
  // z[s] = phi(i, j, k - 1 + s), h[s] = the halo element at plane k + s
  real_t z[3 + DEPTH], h[1 + DEPTH];
#pragma unroll
  for (int s = 0; s < 2 + DEPTH; s++) z[s] = own_in && kb + s - 1 <= fb.khi + 1 ? phi.p[own + (s - 1) * sz] : real_t(0);
#pragma unroll
  for (int s = 0; s < DEPTH; s++) h[s] = hoff >= 0 && kb + s <= ke ? phi.p[hoff + s * sz] : real_t(0);
  for (int k = kb; k <= ke; k++) {
    const int dk = (k - kb) * sz;
    // DEPTH planes ahead, in flight while this plane is worked on
    z[2 + DEPTH] = own_in && k + 1 + DEPTH <= fb.khi + 1 ? phi.p[own + dk + (1 + DEPTH) * sz] : real_t(0);
    h[DEPTH] = hoff >= 0 && k + DEPTH <= ke ? phi.p[hoff + dk + DEPTH * sz] : real_t(0);
    __syncthreads();  // the previous plane's LDS reads are done
    plane[slot] = z[1];
    if (tid < HALO) plane[hslot] = h[0];
    __syncthreads();
    if (active) op(i, j, k, z[1], plane[slot - 1], plane[slot + 1], plane[slot - PX], plane[slot + PX], z[0], z[2]);
#pragma unroll
    for (int s = 0; s < 2 + DEPTH; s++) z[s] = z[s + 1];
#pragma unroll
    for (int s = 0; s < DEPTH; s++) h[s] = h[s + 1];
  }
}

}  // namespace march
}  // namespace lsmhip
