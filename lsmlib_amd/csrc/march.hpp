// march.hpp -- plane-marching launch for the whole-grid kernels whose stencil
// is the 7-point star (one neighbour along each axis), the 19-point cross
// (the star and the twelve edge neighbours) or the lines of R cells either
// side along each axis (ENO / WENO, R <= 3).
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
// The operation receives the centre value and its neighbours and does the
// same arithmetic, in the same order, as a kernel that loads them itself, so
// results are bitwise those of the one-cell-per-thread form.
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

// Position (px, py) in the (kMX + 2R) x (kMY + 2R) plane tile of halo element
// h: R full rows above, R full rows below, then the R cells left and right of
// each of the kMY interior rows.
template <int R>
__device__ __forceinline__ void halo_pos(int h, int& px, int& py) {
  constexpr int PX = kMX + 2 * R;
  if (h < R * PX) {
    py = h / PX;
    px = h % PX;
  } else if (h < 2 * R * PX) {
    py = R + kMY + (h - R * PX) / PX;
    px = (h - R * PX) % PX;
  } else {
    const int t = h - 2 * R * PX, c = t % (2 * R);
    py = R + t / (2 * R);
    px = c < R ? c : kMX + c;
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
    halo_pos<1>(tid, px, py);
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

// star_march_kernel over tiles twice as wide: each thread owns the columns
// i0 + x and i0 + kMX + x of a 2 kMX x kMY tile.  Per cache-line request a
// plane row then carries 512 rather than 256 bytes of FP32 (1024 / 512 of
// FP64), and the two one-cell halo columns serve twice the cells.
inline Tiles tiles_for_x2(const Box& fb) {
  return Tiles{(fb.nx() + 2 * kMX - 1) / (2 * kMX), (fb.ny() + kMY - 1) / kMY, (fb.nz() + kChunk - 1) / kChunk, kChunk};
}
template <class Op>
__global__ void __launch_bounds__(kMX* kMY) star_march_x2_kernel(Box fb, View<const real_t> phi, Tiles tl, Op op) {
  constexpr int TW = 2 * kMX, PX = TW + 2, PLANE = PX * (kMY + 2), HALO = PLANE - TW * kMY;
  __shared__ real_t plane[PLANE];
  static_assert(HALO <= kMX * kMY, "one halo element per thread at most");
  const int tid = threadIdx.y * kMX + threadIdx.x;
  const int nt = tl.nti * tl.ntj * tl.ntk;
  const int per = (nt + 7) / 8;
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= nt) return;
  const int ti = t % tl.nti, tj = (t / tl.nti) % tl.ntj, tk = t / (tl.nti * tl.ntj);
  const int i0 = fb.ilo + ti * TW, j0 = fb.jlo + tj * kMY;
  const int kb = fb.klo + tk * tl.kc;
  const int ke = min(kb + tl.kc - 1, fb.khi);
  const int ia = i0 + (int)threadIdx.x, ib = ia + kMX, j = j0 + (int)threadIdx.y;
  const bool act_a = ia <= fb.ihi && j <= fb.jhi, act_b = ib <= fb.ihi && j <= fb.jhi;
  const bool in_a = ia <= fb.ihi + 1 && j <= fb.jhi + 1, in_b = ib <= fb.ihi + 1 && j <= fb.jhi + 1;
  const int own_a = in_a ? phi.offset(ia, j, kb) : 0, own_b = in_b ? phi.offset(ib, j, kb) : 0;
  // halo: the row above and the row below (PX cells each), then the cells left
  // and right of each interior row
  int hoff = -1, hslot = 0;
  if (tid < HALO) {
    int px, py;
    if (tid < 2 * PX) {
      py = tid < PX ? 0 : kMY + 1;
      px = tid % PX;
    } else {
      const int h = tid - 2 * PX;
      py = 1 + h / 2;
      px = h % 2 ? TW + 1 : 0;
    }
    const int gi = i0 - 1 + px, gj = j0 - 1 + py;
    hslot = py * PX + px;
    if (gi <= fb.ihi + 1 && gj <= fb.jhi + 1) hoff = phi.offset(gi, gj, kb);
  }
  const int sz = phi.sz;
  const int sa = (threadIdx.y + 1) * PX + threadIdx.x + 1, sb = sa + kMX;
  // za[s] = phi(ia, j, k - 1 + s), zb likewise, h[s] = the halo element at plane k + s
  real_t za[4], zb[4], h[2];
#pragma unroll
  for (int s = 0; s < 3; s++) {
    za[s] = in_a && kb + s - 1 <= fb.khi + 1 ? phi.p[own_a + (s - 1) * sz] : real_t(0);
    zb[s] = in_b && kb + s - 1 <= fb.khi + 1 ? phi.p[own_b + (s - 1) * sz] : real_t(0);
  }
  h[0] = hoff >= 0 ? phi.p[hoff] : real_t(0);
  for (int k = kb; k <= ke; k++) {
    const int dk = (k - kb) * sz;
    // the next plane, in flight while this plane is worked on
    za[3] = in_a && k + 2 <= fb.khi + 1 ? phi.p[own_a + dk + 2 * sz] : real_t(0);
    zb[3] = in_b && k + 2 <= fb.khi + 1 ? phi.p[own_b + dk + 2 * sz] : real_t(0);
    h[1] = hoff >= 0 && k + 1 <= ke ? phi.p[hoff + dk + sz] : real_t(0);
    __syncthreads();  // the previous plane's LDS reads are done
    plane[sa] = za[1];
    plane[sb] = zb[1];
    if (tid < HALO) plane[hslot] = h[0];
    __syncthreads();
    if (act_a) op(ia, j, k, za[1], plane[sa - 1], plane[sa + 1], plane[sa - PX], plane[sa + PX], za[0], za[2]);
    if (act_b) op(ib, j, k, zb[1], plane[sb - 1], plane[sb + 1], plane[sb - PX], plane[sb + PX], zb[0], zb[2]);
#pragma unroll
    for (int s = 0; s < 3; s++) {
      za[s] = za[s + 1];
      zb[s] = zb[s + 1];
    }
    h[0] = h[1];
  }
}

// The 19 values of the compact second-order cross around (i, j, k): the cell,
// its six face neighbours and the twelve edge neighbours, the edges of each
// plane pair ordered (+,+), (+,-), (-,+), (-,-).
struct Cross19 {
  real_t c, xp, xm, yp, ym, zp, zm;
  real_t xy[4], xz[4], yz[4];
};

// op.load(i, j, k) -> Op::In (pointwise inputs, issued before the plane's
// barriers) and op.store(i, j, k, const Cross19&, const Op::In&) at every cell
// of fb; phi must cover fb widened by one cell on every side.  The plane k+1
// (with a one-cell halo, corners included) is staged in LDS while cell k is
// computed: its face neighbours along x and y serve as the k+1 edges of cell
// k, and carried one and two iterations serve as the plane-k values and the
// k-1 edges.
template <class Op>
__global__ void __launch_bounds__(kMX* kMY) cross19_march_kernel(Box fb, View<const real_t> phi, Tiles tl, Op op) {
  constexpr int PX = kMX + 2, PLANE = PX * (kMY + 2), HALO = PLANE - kMX * kMY;
  __shared__ real_t plane[PLANE];
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
  const bool own_in = i <= fb.ihi + 1 && j <= fb.jhi + 1;
  const int own = own_in ? phi.offset(i, j, kb) : 0;
  int hoff = -1, hslot = 0;
  if (tid < HALO) {
    int px, py;
    halo_pos<1>(tid, px, py);
    const int gi = i0 - 1 + px, gj = j0 - 1 + py;
    hslot = py * PX + px;
    if (gi <= fb.ihi + 1 && gj <= fb.jhi + 1) hoff = phi.offset(gi, gj, kb);
  }
  const int sz = phi.sz;
  const int slot = (threadIdx.y + 1) * PX + threadIdx.x + 1;
  // in-plane neighbours of this thread's cell in the staged plane
  struct Ring {
    real_t xp, xm, yp, ym, pp, pm, mp, mm;  // faces and the (+,+), (+,-), (-,+), (-,-) corners
  };
  auto own_at = [&](int d) { return own_in ? phi.p[own + d * sz] : real_t(0); };
  auto halo_at = [&](int d) { return hoff >= 0 ? phi.p[hoff + d * sz] : real_t(0); };
  auto put = [&](real_t own_v, real_t halo_v) {
    __syncthreads();  // the previous plane's LDS reads are done
    plane[slot] = own_v;
    if (tid < HALO) plane[hslot] = halo_v;
    __syncthreads();
    return Ring{plane[slot + 1],      plane[slot - 1],      plane[slot + PX],     plane[slot - PX],
                plane[slot + PX + 1], plane[slot - PX + 1], plane[slot + PX - 1], plane[slot - PX - 1]};
  };
  // planes kb - 1 and kb, and plane kb + 1 in flight (all inside phi's box)
  real_t zm = own_at(-1), zc = own_at(0), on = own_at(1);
  const real_t hm = halo_at(-1), h0 = halo_at(0);
  real_t hn = halo_at(1);
  Ring rm = put(zm, hm), r0 = put(zc, h0);
  for (int k = kb; k <= ke; k++) {
    typename Op::In in{};
    if (active) in = op.load(i, j, k);
    // plane k + 2, in flight while plane k + 1 is staged and cell k computed
    const int d2 = k - kb + 2;
    const real_t on2 = k + 2 <= ke + 1 ? own_at(d2) : real_t(0);
    const real_t hn2 = k + 2 <= ke + 1 ? halo_at(d2) : real_t(0);
    const Ring rp = put(on, hn);  // plane k + 1
    const real_t zp = on;
    if (active) {
      Cross19 x;
      x.c = zc;
      x.xp = r0.xp;
      x.xm = r0.xm;
      x.yp = r0.yp;
      x.ym = r0.ym;
      x.zp = zp;
      x.zm = zm;
      // (i+1, j+1), (i+1, j-1), (i-1, j+1), (i-1, j-1) of plane k
      x.xy[0] = r0.pp;
      x.xy[1] = r0.pm;
      x.xy[2] = r0.mp;
      x.xy[3] = r0.mm;
      // (i+1, k+1), (i+1, k-1), (i-1, k+1), (i-1, k-1)
      x.xz[0] = rp.xp;
      x.xz[1] = rm.xp;
      x.xz[2] = rp.xm;
      x.xz[3] = rm.xm;
      // (j+1, k+1), (j+1, k-1), (j-1, k+1), (j-1, k-1)
      x.yz[0] = rp.yp;
      x.yz[1] = rm.yp;
      x.yz[2] = rp.ym;
      x.yz[3] = rm.ym;
      op.store(i, j, k, x, in);
    }
    rm = r0;
    r0 = rp;
    zm = zc;
    zc = zp;
    on = on2;
    hn = hn2;
  }
}

// op.load(i, j, k) -> Op::In (pointwise inputs, issued before the plane's
// barriers) and op.store(i, j, k, fx, fy, fz, in) at every cell of fb, where
// f?[s + 3] = phi at offset s (|s| <= R) along x, y, z; phi must cover fb
// widened by R cells on every side.  The plane with an R-cell x/y halo goes
// through LDS; the thread's own column from k - R to k + R is carried in
// registers, with plane k + R + 1 in flight.
template <int R, class Op>
__global__ void __launch_bounds__(kMX* kMY) line_march_kernel(Box fb, View<const real_t> phi, Tiles tl, Op op) {
  static_assert(R >= 1 && R <= 3, "f? holds offsets -3 .. 3");
  constexpr int PX = kMX + 2 * R, PLANE = PX * (kMY + 2 * R), HALO = PLANE - kMX * kMY, W = 2 * R + 1;
  static_assert(HALO <= kMX * kMY, "one halo element per thread at most");
  __shared__ real_t plane[PLANE];
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
  // the own column is a neighbour of the tile's cells when it lies within R
  // cells of fb (inside phi's box); columns further out are never read
  const bool own_in = i <= fb.ihi + R && j <= fb.jhi + R;
  const int own = own_in ? phi.offset(i, j, kb) : 0;
  int hoff = -1, hslot = 0;
  if (tid < HALO) {
    int px, py;
    halo_pos<R>(tid, px, py);
    const int gi = i0 - R + px, gj = j0 - R + py;
    hslot = py * PX + px;
    if (gi <= fb.ihi + R && gj <= fb.jhi + R) hoff = phi.offset(gi, gj, kb);
  }
  const int sz = phi.sz;
  const int slot = (threadIdx.y + R) * PX + threadIdx.x + R;
  // z[s] = phi(i, j, k - R + s); z[W] = plane k + R + 1 in flight
  real_t z[W + 1];
#pragma unroll
  for (int s = 0; s < W; s++) z[s] = own_in ? phi.p[own + (s - R) * sz] : real_t(0);
  real_t h0 = hoff >= 0 ? phi.p[hoff] : real_t(0);
  for (int k = kb; k <= ke; k++) {
    const int dk = (k - kb) * sz;
    typename Op::In in{};
    if (active) in = op.load(i, j, k);
    // the next cell's window and halo, needed only while k + 1 <= ke
    z[W] = own_in && k < ke ? phi.p[own + dk + (R + 1) * sz] : real_t(0);
    const real_t h1 = hoff >= 0 && k < ke ? phi.p[hoff + dk + sz] : real_t(0);
    __syncthreads();  // the previous plane's LDS reads are done
    plane[slot] = z[R];
    if (tid < HALO) plane[hslot] = h0;
    __syncthreads();
    if (active) {
      real_t fx[7] = {}, fy[7] = {}, fz[7] = {};
#pragma unroll
      for (int s = -R; s <= R; s++) {
        fx[s + 3] = plane[slot + s];
        fy[s + 3] = plane[slot + s * PX];
        fz[s + 3] = z[s + R];
      }
      op.store(i, j, k, fx, fy, fz, in);
    }
#pragma unroll
    for (int s = 0; s < W; s++) z[s] = z[s + 1];
    h0 = h1;
  }
}

}  // namespace march
}  // namespace lsmhip
