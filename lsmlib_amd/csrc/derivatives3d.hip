// derivatives3d.hip -- whole-grid HJ-ENO1/2/3 and HJ-WENO5 derivatives
// (plain and velocity-upwinded), central gradients, Laplacian, averaged and
// magnitude of the gradient.
//
// The reference (src/toolbox/lsm_spatial_derivatives3d.f.in) materialises the
// undivided differences D1/D2/D3 in caller-owned scratch arrays, one full-grid
// pass per order and axis.  Here each thread recomputes the differences it
// needs from phi in registers with exactly the same floating-point operations
// (D1 = phi(i) - phi(i-1), D2 = -(D1(i) - D1(i+1)), D3 = D2(i) - D2(i-1)), so
// the results are bitwise those of the reference while the scratch arrays are
// neither read nor written.  Every difference a fill-box cell uses lies inside
// the range the reference computes, so its 1e10 fill values never enter.
#include "lsmhip_core.hpp"
#include "march.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

struct Axes3 {
  View<real_t> o[3];
};

// HJ-ENO1/2/3 and HJ-WENO5 plus and minus derivatives, and their velocity-
// upwinded forms, on the plane march: the lines of phi come from LDS (x, y)
// and the thread's carried column (z); the arithmetic is stencil.hpp's, in the
// reference's operation order.
template <int ORDER>
struct PlusMinusLine {
  Axes3 plus, minus;
  real_t inv[3];
  struct In {};
  __device__ In load(int, int, int) const { return In{}; }
  __device__ void store(int i, int j, int k, const real_t (&fx)[7], const real_t (&fy)[7], const real_t (&fz)[7],
                        const In&) const {
    const real_t* f[3] = {fx, fy, fz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      real_t p, m;
      stencil::plus_minus<ORDER>(f[d], inv[d], p, m);
      plus.o[d](i, j, k) = p;
      minus.o[d](i, j, k) = m;
    }
  }
};
template <int ORDER>
struct UpwindLine {
  Axes3 out;
  View<const real_t> vx, vy, vz;
  real_t inv[3];
  struct In {
    real_t v[3];
  };
  __device__ In load(int i, int j, int k) const { return In{{vx(i, j, k), vy(i, j, k), vz(i, j, k)}}; }
  __device__ void store(int i, int j, int k, const real_t (&fx)[7], const real_t (&fy)[7], const real_t (&fz)[7],
                        const In& in) const {
    const real_t* f[3] = {fx, fy, fz};
#pragma unroll
    for (int d = 0; d < 3; d++)
      out.o[d](i, j, k) =
          fabs(in.v[d]) < kZeroTol ? real_t(0) : stencil::upwind<ORDER>(f[d], inv[d], in.v[d] > real_t(0));
  }
};

// Second-order central gradient and Laplacian on the plane march
// (src/toolbox/lsm_spatial_derivatives3d.f.in, lsm3dCentralGradOrder2 and
// lsm3dLaplacianOrder2), in the reference's operation order.
struct Central2Star {
  Axes3 out;
  real_t fx, fy, fz;
  __device__ void operator()(int i, int j, int k, real_t, real_t xm, real_t xp, real_t ym, real_t yp, real_t zm,
                             real_t zp) const {
    out.o[0](i, j, k) = (xp - xm) * fx;
    out.o[1](i, j, k) = (yp - ym) * fy;
    out.o[2](i, j, k) = (zp - zm) * fz;
  }
};
struct Laplacian2Star {
  View<real_t> out;
  real_t ax, ay, az;
  __device__ void operator()(int i, int j, int k, real_t c, real_t xm, real_t xp, real_t ym, real_t yp, real_t zm,
                             real_t zp) const {
    out(i, j, k) = ax * (xp - real_t(2) * c + xm) + ay * (yp - real_t(2) * c + ym) + az * (zp - real_t(2) * c + zm);
  }
};

// Fourth-order central gradient on the plane march (lsm3dCentralGradOrder4),
// in the reference's operation order.
struct Central4Line {
  Axes3 out;
  real_t r[3];
  struct In {};
  __device__ In load(int, int, int) const { return In{}; }
  __device__ void store(int i, int j, int k, const real_t (&fx)[7], const real_t (&fy)[7], const real_t (&fz)[7],
                        const In&) const {
    const real_t e = real_t(8);
    const real_t* f[3] = {fx, fy, fz};
#pragma unroll
    for (int d = 0; d < 3; d++) out.o[d](i, j, k) = (-f[d][5] + e * f[d][4] + f[d][1] - e * f[d][2]) * r[d];
  }
};

// lsm3dCentralHessian (src/toolbox/lsm_spatial_derivatives3d.f.in:3658-3715):
// compact second differences over dx*dx etc., mixed ones as 0.25 (f(+,+) -
// f(-,+) - f(+,-) + f(-,-)) / (dx*dy), in that operation order, on the plane
// march over the 19-point cross.
struct Hessian6 {
  View<real_t> xx, xy, xz, yy, yz, zz;
};
struct HessianMarch {
  Hessian6 out;
  real_t dxx, dxy, dxz, dyy, dyz, dzz;
  struct In {};
  __device__ In load(int, int, int) const { return In{}; }
  __device__ void store(int i, int j, int k, const march::Cross19& c, const In&) const {
    const real_t q = real_t(0.25);
    out.xx(i, j, k) = (c.xp - real_t(2) * c.c + c.xm) / dxx;
    out.xy(i, j, k) = q * (c.xy[0] - c.xy[2] - c.xy[1] + c.xy[3]) / dxy;
    out.xz(i, j, k) = q * (c.xz[0] - c.xz[2] - c.xz[1] + c.xz[3]) / dxz;
    out.yy(i, j, k) = (c.yp - real_t(2) * c.c + c.ym) / dyy;
    out.yz(i, j, k) = q * (c.yz[0] - c.yz[2] - c.yz[1] + c.yz[3]) / dyz;
    out.zz(i, j, k) = (c.zp - real_t(2) * c.c + c.zm) / dzz;
  }
};

// One axis of lsm3dPhiUpwindGradF (src/toolbox/lsm_spatial_derivatives3d.f.in:
// 3270-3458): the one-sided derivative from the neighbour whose phi is nearer
// zero (smaller for phi > 0, larger for phi < 0), the plus one where the minus
// neighbour does not qualify, 0 where neither does.
__device__ __forceinline__ real_t phi_upwind_pick(real_t c, real_t nm, real_t np, real_t fp, real_t fm) {
  if (c > real_t(0)) {
    if (nm <= c) return np < nm ? fp : fm;
    return np <= c ? fp : real_t(0);
  }
  if (nm >= c) return np > nm ? fp : fm;
  return np >= c ? fp : real_t(0);
}
__global__ void __launch_bounds__(256) phi_upwind_grad_kernel(Box fb, Axes3 out, Axes3 plus, Axes3 minus,
                                                              View<const real_t> phi) {
  LSMHIP_CELL(fb);
  const real_t c = phi(i, j, k);
  if (!(c > real_t(0)) && !(c < real_t(0))) return;  // phi == 0 (or NaN): the outputs are left as they are
  const real_t n[6] = {phi(i - 1, j, k), phi(i + 1, j, k), phi(i, j - 1, k),
                       phi(i, j + 1, k), phi(i, j, k - 1), phi(i, j, k + 1)};
  real_t fp[3], fm[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    fp[d] = plus.o[d](i, j, k);
    fm[d] = minus.o[d](i, j, k);
  }
#pragma unroll
  for (int d = 0; d < 3; d++) out.o[d](i, j, k) = phi_upwind_pick(c, n[2 * d], n[2 * d + 1], fp[d], fm[d]);
}

__global__ void __launch_bounds__(256) average_grad_kernel(Box fb, Axes3 out, Axes3 plus, Axes3 minus) {
  LSMHIP_CELL(fb);
  real_t p[3], m[3];  // loads first: a load after a store to a possibly aliasing output would wait for it
#pragma unroll
  for (int d = 0; d < 3; d++) {
    p[d] = plus.o[d](i, j, k);
    m[d] = minus.o[d](i, j, k);
  }
#pragma unroll
  for (int d = 0; d < 3; d++) out.o[d](i, j, k) = real_t(0.5) * (p[d] + m[d]);
}

__global__ void __launch_bounds__(256) grad_mag_kernel(Box fb, Axes3 g, View<real_t> mag) {
  LSMHIP_CELLS(fb, kPointNK);
  struct V {
    real_t x, y, z;
  };
  pair_cells(
      k0, nk, [&](int k) { return V{g.o[0](i, j, k), g.o[1](i, j, k), g.o[2](i, j, k)}; },
      [&](int k, const V& a) { mag(i, j, k) = stencil::grad_magnitude(a.x, a.y, a.z, real_t(1.e-8)); });
}

Axes3 axes(real_t* x, real_t* y, real_t* z, const Box& b) { return Axes3{{view(x, b), view(y, b), view(z, b)}}; }

// FP32 HJ-WENO5 plus and minus derivatives on the plane march, each WENO5
// window formed once.  The minus derivative at a cell and the plus derivative
// at the cell before it along an axis come from one window (the cell's own),
// so each cell forms its window once per axis, takes both derivatives from it
// (stencil::fastf::fwd_rev), and passes the plus derivative to the cell before
// it: through LDS along x and y, carried in a register along z.  One wave forms
// the plus derivatives of the cells one column past the tile; the last row
// forms those of the row past it.  F is float: a template parameter so that the
// FP64 build, which never launches this kernel, does not compile it.
template <class F>
__global__ void __launch_bounds__(march::kMX* march::kMY)
    weno5_pm_own_window_kernel(Box fb, View<const F> phi, march::Tiles tl, Axes3 plus, Axes3 minus, F ix, F iy, F iz) {
  using namespace stencil::fastf;
  constexpr int R = 3, MX = march::kMX, MY = march::kMY, PX = MX + 2 * R, PLANE = PX * (MY + 2 * R);
  constexpr int HALO = PLANE - MX * MY;
  static_assert(HALO <= MX * MY, "one halo element per thread at most");
  __shared__ F plane[PLANE];
  __shared__ F yplus[MY][MX];
  __shared__ F xplus[MY][MX + 1];  // column MX: the cells one past the tile
  const int tid = threadIdx.y * MX + threadIdx.x;
  const int nt = tl.nti * tl.ntj * tl.ntk;
  const int per = (nt + 7) / 8;
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= nt) return;
  const int ti = t % tl.nti, tj = (t / tl.nti) % tl.ntj, tk = t / (tl.nti * tl.ntj);
  const int i0 = fb.ilo + ti * MX, j0 = fb.jlo + tj * MY;
  const int kb = fb.klo + tk * tl.kc;
  const int ke = min(kb + tl.kc - 1, fb.khi);
  const int i = i0 + (int)threadIdx.x, j = j0 + (int)threadIdx.y;
  const bool active = i <= fb.ihi && j <= fb.jhi;
  // columns within R cells of fb lie inside phi's box; further out they are never read
  const bool own_in = i <= fb.ihi + R && j <= fb.jhi + R;
  const int own = own_in ? phi.offset(i, j, kb) : 0;
  int hoff = -1, hslot = 0;
  if (tid < HALO) {
    int px, py;
    march::halo_pos<R>(tid, px, py);
    const int gi = i0 - R + px, gj = j0 - R + py;
    hslot = py * PX + px;
    if (gi <= fb.ihi + R && gj <= fb.jhi + R) hoff = phi.offset(gi, gj, kb);
  }
  const int sz = phi.sz;
  const int slot = (threadIdx.y + R) * PX + threadIdx.x + R;
  // z[s] = phi(i, j, k - 3 + s); z[7] = plane k + 4 in flight
  F z[8];
#pragma unroll
  for (int s = 0; s < 7; s++) z[s] = own_in ? phi.p[own + (s - R) * sz] : F(0);
  F h0 = hoff >= 0 ? phi.p[hoff] : F(0);
  // minus derivative along z at the first plane, from its window D(k-2 .. k+2)
  F zm, unused;
  fwd_rev(win5(z[1] - z[0], z[2] - z[1], z[3] - z[2], z[4] - z[3], z[5] - z[4]), z[2] - z[1], z[3] - z[2],
          z[4] - z[3], zm, unused);
  for (int k = kb; k <= ke; k++) {
    const int dk = (k - kb) * sz;
    z[7] = own_in && k < ke ? phi.p[own + dk + (R + 1) * sz] : F(0);
    const F h1 = hoff >= 0 && k < ke ? phi.p[hoff + dk + sz] : F(0);
    __syncthreads();  // the previous plane's LDS reads are done
    plane[slot] = z[R];
    if (tid < HALO) plane[hslot] = h0;
    __syncthreads();
    F dx[6], dy[6];  // D(i-2 .. i+3), D(j-2 .. j+3)
#pragma unroll
    for (int q = -2; q <= 3; q++) {
      dx[q + 2] = plane[slot + q] - plane[slot + q - 1];
      dy[q + 2] = plane[slot + q * PX] - plane[slot + (q - 1) * PX];
    }
    F xm, xpp, ym, ypp, zmn, zp;
    fwd_rev(win5(dx[0], dx[1], dx[2], dx[3], dx[4]), dx[1], dx[2], dx[3], xm, xpp);
    xplus[threadIdx.y][threadIdx.x] = xpp * ix;
    if (threadIdx.y == 0 && threadIdx.x < MY) {  // the column one past the tile, one row per lane
      const int row = ((int)threadIdx.x + R) * PX + MX + R;
      F e[5], ef, er;
#pragma unroll
      for (int q = -2; q <= 2; q++) e[q + 2] = plane[row + q] - plane[row + q - 1];
      fwd_rev(win5(e[0], e[1], e[2], e[3], e[4]), e[1], e[2], e[3], ef, er);
      xplus[threadIdx.x][MX] = er * ix;
    }
    fwd_rev(win5(dy[0], dy[1], dy[2], dy[3], dy[4]), dy[1], dy[2], dy[3], ym, ypp);
    yplus[threadIdx.y][threadIdx.x] = ypp * iy;
    // z: the window D(k-1 .. k+3) gives the plus derivative here and the minus one at k + 1
    fwd_rev(win5(z[2] - z[1], z[3] - z[2], z[4] - z[3], z[5] - z[4], z[6] - z[5]), z[3] - z[2], z[4] - z[3],
            z[5] - z[4], zmn, zp);
    __syncthreads();
    F yp;
    if (threadIdx.y == MY - 1) {  // the row past the tile
      F nf, nr;
      fwd_rev(win5(dy[1], dy[2], dy[3], dy[4], dy[5]), dy[2], dy[3], dy[4], nf, nr);
      yp = nr * iy;
    } else {
      yp = yplus[threadIdx.y + 1][threadIdx.x];
    }
    const F xp = xplus[threadIdx.y][threadIdx.x + 1];
    if (active) {
      plus.o[0](i, j, k) = xp;
      plus.o[1](i, j, k) = yp;
      plus.o[2](i, j, k) = zp * iz;
      minus.o[0](i, j, k) = xm * ix;
      minus.o[1](i, j, k) = ym * iy;
      minus.o[2](i, j, k) = zm * iz;
    }
    zm = zmn;
#pragma unroll
    for (int s = 0; s < 7; s++) z[s] = z[s + 1];
    h0 = h1;
  }
}

template <int ORDER>
void hj_plus_minus(const char* who, real_t* pxp, real_t* pyp, real_t* pzp, const Box& gp, real_t* pxm, real_t* pym,
                   real_t* pzm, const Box& gm, const real_t* phi, const Box& pg, const Box& fb, real_t dx, real_t dy,
                   real_t dz) {
  if (fb.empty()) return;
  constexpr int w = ORDER == 1 ? 1 : (ORDER == 2 ? 2 : 3);
  if (!check_box(gp, fb, who, "grad_phi_plus") || !check_box(gm, fb, who, "grad_phi_minus") ||
      !check_box(pg, fb.widen(w, w, w), who, "phi") || !check_size(pg, who) || !check_size(gp, who) ||
      !check_size(gm, who))
    return;
  const march::Tiles tl = march::tiles_for(fb);
  if constexpr (ORDER == 5 && sizeof(real_t) == 4) {
    hipLaunchKernelGGL(weno5_pm_own_window_kernel<real_t>, march::grid_for(tl), march::block(), 0, stream(), fb,
                       view(phi, pg), tl, axes(pxp, pyp, pzp, gp), axes(pxm, pym, pzm, gm), real_t(1.0) / dx,
                       real_t(1.0) / dy, real_t(1.0) / dz);
  } else {
    hipLaunchKernelGGL((march::line_march_kernel<w, PlusMinusLine<ORDER>>), march::grid_for(tl), march::block(), 0,
                       stream(), fb, view(phi, pg), tl,
                       PlusMinusLine<ORDER>{axes(pxp, pyp, pzp, gp), axes(pxm, pym, pzm, gm),
                                            {real_t(1.0) / dx, real_t(1.0) / dy, real_t(1.0) / dz}});
  }
  finish(who);
}

template <int ORDER>
void hj_upwind(const char* who, real_t* px, real_t* py, real_t* pz, const Box& g, const real_t* phi, const Box& pg,
               const real_t* vx, const real_t* vy, const real_t* vz, const Box& vg, const Box& fb, real_t dx,
               real_t dy, real_t dz) {
  if (fb.empty()) return;
  constexpr int w = ORDER == 1 ? 1 : (ORDER == 2 ? 2 : 3);
  if (!check_box(g, fb, who, "grad_phi") || !check_box(vg, fb, who, "vel") ||
      !check_box(pg, fb.widen(w, w, w), who, "phi") || !check_size(pg, who) || !check_size(g, who) ||
      !check_size(vg, who))
    return;
  const march::Tiles tl = march::tiles_for(fb);
  hipLaunchKernelGGL((march::line_march_kernel<w, UpwindLine<ORDER>>), march::grid_for(tl), march::block(), 0,
                     stream(), fb, view(phi, pg), tl,
                     UpwindLine<ORDER>{axes(px, py, pz, g), view(vx, vg), view(vy, vg), view(vz, vg),
                                       {real_t(1.0) / dx, real_t(1.0) / dy, real_t(1.0) / dz}});
  finish(who);
}

}  // namespace

// ---------------------------------------------------------------------------
// Drop-in entry points (src/toolbox/lsm_spatial_derivatives3d.h)
// ---------------------------------------------------------------------------
#define PM_ARGS                                                                                            \
  real_t *pxp, real_t *pyp, real_t *pzp, LSMHIP_BOXP(gp), real_t *pxm, real_t *pym, real_t *pzm,          \
      LSMHIP_BOXP(gm), const real_t *phi, LSMHIP_BOXP(pg)
#define PM_PASS pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm), phi, LSMHIP_BOX(pg)

LSMHIP_EXPORT void lsm3dhjeno1_(PM_ARGS, real_t*, LSMHIP_BOXP(d1), LSMHIP_BOXP(fb), const real_t* dx,
                                const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, phi);
  hj_plus_minus<1>("lsm3dhjeno1_", PM_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}
LSMHIP_EXPORT void lsm3dhjeno2_(PM_ARGS, real_t*, LSMHIP_BOXP(d1), real_t*, LSMHIP_BOXP(d2), LSMHIP_BOXP(fb),
                                const real_t* dx, const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, phi);
  hj_plus_minus<2>("lsm3dhjeno2_", PM_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}
LSMHIP_EXPORT void lsm3dhjeno3_(PM_ARGS, real_t*, LSMHIP_BOXP(d1), real_t*, LSMHIP_BOXP(d2), real_t*,
                                LSMHIP_BOXP(d3), LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, phi);
  hj_plus_minus<3>("lsm3dhjeno3_", PM_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}
LSMHIP_EXPORT void lsm3dhjweno5_(PM_ARGS, real_t*, LSMHIP_BOXP(d1), LSMHIP_BOXP(fb), const real_t* dx,
                                 const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, phi);
  hj_plus_minus<5>("lsm3dhjweno5_", PM_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}

#define UP_ARGS                                                                                         \
  real_t *px, real_t *py, real_t *pz, LSMHIP_BOXP(g), const real_t *phi, LSMHIP_BOXP(pg), const real_t *vx, \
      const real_t *vy, const real_t *vz, LSMHIP_BOXP(vg)
#define UP_PASS px, py, pz, LSMHIP_BOX(g), phi, LSMHIP_BOX(pg), vx, vy, vz, LSMHIP_BOX(vg)

LSMHIP_EXPORT void lsm3dupwindhjeno1_(UP_ARGS, real_t*, LSMHIP_BOXP(d1), LSMHIP_BOXP(fb), const real_t* dx,
                                      const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi, vx, vy, vz);
  hj_upwind<1>("lsm3dupwindhjeno1_", UP_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}
LSMHIP_EXPORT void lsm3dupwindhjeno2_(UP_ARGS, real_t*, LSMHIP_BOXP(d1), real_t*, LSMHIP_BOXP(d2), LSMHIP_BOXP(fb),
                                      const real_t* dx, const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi, vx, vy, vz);
  hj_upwind<2>("lsm3dupwindhjeno2_", UP_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}
LSMHIP_EXPORT void lsm3dupwindhjeno3_(UP_ARGS, real_t*, LSMHIP_BOXP(d1), real_t*, LSMHIP_BOXP(d2), real_t*,
                                      LSMHIP_BOXP(d3), LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                      const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi, vx, vy, vz);
  hj_upwind<3>("lsm3dupwindhjeno3_", UP_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}
LSMHIP_EXPORT void lsm3dupwindhjweno5_(UP_ARGS, real_t*, LSMHIP_BOXP(d1), LSMHIP_BOXP(fb), const real_t* dx,
                                       const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi, vx, vy, vz);
  hj_upwind<5>("lsm3dupwindhjweno5_", UP_PASS, LSMHIP_BOX(fb), *dx, *dy, *dz);
}

LSMHIP_EXPORT void lsm3dcentralgradorder2_(real_t* px, real_t* py, real_t* pz, LSMHIP_BOXP(g), const real_t* phi,
                                           LSMHIP_BOXP(pg), LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                           const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi);
  const char* who = "lsm3dcentralgradorder2_";
  Box fb = LSMHIP_BOX(fb), g = LSMHIP_BOX(g), pg = LSMHIP_BOX(pg);
  if (fb.empty()) return;
  if (!check_box(g, fb, who, "grad_phi") || !check_box(pg, fb.widen(1, 1, 1), who, "phi") || !check_size(g, who) ||
      !check_size(pg, who))
    return;
  const real_t fx = real_t(0.5) / *dx, fy = real_t(0.5) / *dy, fz = real_t(0.5) / *dz;
  const march::Tiles tl = march::tiles_for(fb);
  hipLaunchKernelGGL(march::star_march_kernel<Central2Star>, march::grid_for(tl), march::block(), 0, stream(), fb,
                     view(phi, pg), tl, Central2Star{axes(px, py, pz, g), fx, fy, fz});
  finish(who);
}

LSMHIP_EXPORT void lsm3dcentralgradorder4_(real_t* px, real_t* py, real_t* pz, LSMHIP_BOXP(g), const real_t* phi,
                                           LSMHIP_BOXP(pg), LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                           const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi);
  const char* who = "lsm3dcentralgradorder4_";
  Box fb = LSMHIP_BOX(fb), g = LSMHIP_BOX(g), pg = LSMHIP_BOX(pg);
  if (fb.empty()) return;
  if (!check_box(g, fb, who, "grad_phi") || !check_box(pg, fb.widen(2, 2, 2), who, "phi") || !check_size(g, who) ||
      !check_size(pg, who))
    return;
  const real_t c = real_t(0.0833333333333333333333);
  const march::Tiles tl = march::tiles_for(fb);
  hipLaunchKernelGGL((march::line_march_kernel<2, Central4Line>), march::grid_for(tl), march::block(), 0, stream(),
                     fb, view(phi, pg), tl, Central4Line{axes(px, py, pz, g), {c / *dx, c / *dy, c / *dz}});
  finish(who);
}

LSMHIP_EXPORT void lsm3dlaplacianorder2_(real_t* lap, LSMHIP_BOXP(lg), const real_t* phi, LSMHIP_BOXP(pg),
                                         LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(lap, phi);
  const char* who = "lsm3dlaplacianorder2_";
  Box fb = LSMHIP_BOX(fb), lg = LSMHIP_BOX(lg), pg = LSMHIP_BOX(pg);
  if (fb.empty()) return;
  if (!check_box(lg, fb, who, "laplacian_phi") || !check_box(pg, fb.widen(1, 1, 1), who, "phi") ||
      !check_size(lg, who) || !check_size(pg, who))
    return;
  const real_t ax = real_t(1.0) / *dx / *dx, ay = real_t(1.0) / *dy / *dy, az = real_t(1.0) / *dz / *dz;
  // 128-wide tiles: 315 -> 253 us FP32 and unchanged FP64 at 512^3 (the central
  // gradient, with three outputs, gained nothing from them and keeps 64-wide tiles)
  const march::Tiles tl = march::tiles_for_x2(fb);
  hipLaunchKernelGGL(march::star_march_x2_kernel<Laplacian2Star>, march::grid_for(tl), march::block(), 0, stream(), fb,
                     view(phi, pg), tl, Laplacian2Star{view(lap, lg), ax, ay, az});
  finish(who);
}

LSMHIP_EXPORT void lsm3dcentralhessian_(real_t* pxx, real_t* pxy, real_t* pxz, real_t* pyy, real_t* pyz, real_t* pzz,
                                        LSMHIP_BOXP(hg), const real_t* phi, LSMHIP_BOXP(pg), LSMHIP_BOXP(fb),
                                        const real_t* dx, const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(pxx, pxy, pxz, pyy, pyz, pzz, phi);
  const char* who = "lsm3dcentralhessian_";
  const Box fb = LSMHIP_BOX(fb), hg = LSMHIP_BOX(hg), pg = LSMHIP_BOX(pg);
  if (fb.empty()) return;
  if (!check_box(hg, fb, who, "hessian") || !check_box(pg, fb.widen(1, 1, 1), who, "phi") || !check_size(hg, who) ||
      !check_size(pg, who))
    return;
  const march::Tiles tl = march::tiles_for(fb);
  hipLaunchKernelGGL(march::cross19_march_kernel<HessianMarch>, march::grid_for(tl), march::block(), 0, stream(), fb,
                     view(phi, pg), tl,
                     HessianMarch{Hessian6{view(pxx, hg), view(pxy, hg), view(pxz, hg), view(pyy, hg), view(pyz, hg),
                                           view(pzz, hg)},
                                  *dx * *dx, *dx * *dy, *dx * *dz, *dy * *dy, *dy * *dz, *dz * *dz});
  finish(who);
}

LSMHIP_EXPORT void lsm3dphiupwindgradf_(real_t* fx, real_t* fy, real_t* fz, LSMHIP_BOXP(g), real_t* fxp, real_t* fyp,
                                        real_t* fzp, LSMHIP_BOXP(gp), real_t* fxm, real_t* fym, real_t* fzm,
                                        LSMHIP_BOXP(gm), const real_t* phi, LSMHIP_BOXP(pg), LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(fx, fy, fz, fxp, fyp, fzp, fxm, fym, fzm, phi);
  const char* who = "lsm3dphiupwindgradf_";
  const Box fb = LSMHIP_BOX(fb), g = LSMHIP_BOX(g), gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm), pg = LSMHIP_BOX(pg);
  if (fb.empty()) return;
  if (!check_box(g, fb, who, "grad_F") || !check_box(gp, fb, who, "grad_F_plus") ||
      !check_box(gm, fb, who, "grad_F_minus") || !check_box(pg, fb.widen(1, 1, 1), who, "phi") ||
      !check_size(g, who) || !check_size(gp, who) || !check_size(gm, who) || !check_size(pg, who))
    return;
  hipLaunchKernelGGL(phi_upwind_grad_kernel, grid_for(fb), block2d(), 0, stream(), fb, axes(fx, fy, fz, g),
                     axes(fxp, fyp, fzp, gp), axes(fxm, fym, fzm, gm), view(phi, pg));
  finish(who);
}

LSMHIP_EXPORT void lsm3daveragegradphi_(real_t* px, real_t* py, real_t* pz, LSMHIP_BOXP(g), real_t* pxp,
                                        real_t* pyp, real_t* pzp, LSMHIP_BOXP(gp), real_t* pxm, real_t* pym,
                                        real_t* pzm, LSMHIP_BOXP(gm), LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, pxp, pyp, pzp, pxm, pym, pzm);
  const char* who = "lsm3daveragegradphi_";
  Box fb = LSMHIP_BOX(fb), g = LSMHIP_BOX(g), gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm);
  if (fb.empty()) return;
  if (!check_box(g, fb, who, "grad_phi") || !check_box(gp, fb, who, "grad_phi_plus") ||
      !check_box(gm, fb, who, "grad_phi_minus") || !check_size(g, who) || !check_size(gp, who) ||
      !check_size(gm, who))
    return;
  hipLaunchKernelGGL(average_grad_kernel, grid_for(fb), block2d(), 0, stream(), fb, axes(px, py, pz, g),
                     axes(pxp, pyp, pzp, gp), axes(pxm, pym, pzm, gm));
  finish(who);
}

LSMHIP_EXPORT void lsm3dgradientmagnitude_(const real_t* px, const real_t* py, const real_t* pz, real_t* mag,
                                           LSMHIP_BOXP(g), LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, mag);
  const char* who = "lsm3dgradientmagnitude_";
  Box fb = LSMHIP_BOX(fb), g = LSMHIP_BOX(g);
  if (fb.empty()) return;
  if (!check_box(g, fb, who, "grad_phi") || !check_size(g, who)) return;
  hipLaunchKernelGGL(grad_mag_kernel, grid_for_cells(fb, kPointNK), block2d(), 0, stream(), fb,
                     axes((real_t*)px, (real_t*)py, (real_t*)pz, g), view(mag, g));
  finish(who);
}
