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
#include "stencil.hpp"

using namespace lsmhip;

namespace {

struct Axes3 {
  View<real_t> o[3];
};

// Loads phi at offsets -R..R along axis d (f[s+3] holds offset s).
template <int R>
__device__ __forceinline__ void load_line(const View<const real_t>& phi, int i, int j, int k, int d, real_t f[7]) {
  const int di = d == 0, dj = d == 1, dk = d == 2;
#pragma unroll
  for (int s = -R; s <= R; s++) f[s + 3] = phi(i + s * di, j + s * dj, k + s * dk);
}

template <int ORDER>
__global__ void __launch_bounds__(256) hj_plus_minus_kernel(Box fb, Axes3 plus, Axes3 minus,
                                                            View<const real_t> phi, real_t ix, real_t iy,
                                                            real_t iz) {
  LSMHIP_CELL(fb);
  const real_t inv[3] = {ix, iy, iz};
  // All three lines are loaded before the first store: an output may alias phi
  // as far as the compiler knows, so a load after a store would wait for it.
  // (One cell per thread: a pair of cells along k measured slower here, the
  // registers of two full crosses costing more occupancy than they gain.)
  real_t f[3][7];
#pragma unroll
  for (int d = 0; d < 3; d++) load_line<(ORDER == 1 ? 1 : (ORDER == 2 ? 2 : 3))>(phi, i, j, k, d, f[d]);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    real_t p, m;
    stencil::plus_minus<ORDER>(f[d], inv[d], p, m);
    plus.o[d](i, j, k) = p;
    minus.o[d](i, j, k) = m;
  }
}

template <int ORDER>
__global__ void __launch_bounds__(256) hj_upwind_kernel(Box fb, Axes3 out, View<const real_t> phi,
                                                        View<const real_t> vx, View<const real_t> vy,
                                                        View<const real_t> vz, real_t ix, real_t iy, real_t iz) {
  LSMHIP_CELL_PAIR(fb);
  const real_t inv[3] = {ix, iy, iz};
  // Every line of both cells is loaded (inside phi's checked box) before the
  // first store, also where the speed below zero_tol leaves it unused.
  real_t vel[2][3], f[2][3][7];
#pragma unroll
  for (int c = 0; c < 2; c++)
    if (c < nk) {
      vel[c][0] = vx(i, j, k0 + c);
      vel[c][1] = vy(i, j, k0 + c);
      vel[c][2] = vz(i, j, k0 + c);
#pragma unroll
      for (int d = 0; d < 3; d++) load_line<(ORDER == 1 ? 1 : (ORDER == 2 ? 2 : 3))>(phi, i, j, k0 + c, d, f[c][d]);
    }
#pragma unroll
  for (int c = 0; c < 2; c++)
    if (c < nk)
#pragma unroll
      for (int d = 0; d < 3; d++)
        out.o[d](i, j, k0 + c) = fabs(vel[c][d]) < kZeroTol
                                     ? real_t(0)
                                     : stencil::upwind<ORDER>(f[c][d], inv[d], vel[c][d] > real_t(0));
}

__global__ void __launch_bounds__(256) central2_kernel(Box fb, Axes3 out, View<const real_t> phi, real_t fx,
                                                       real_t fy, real_t fz) {
  LSMHIP_CELL_PAIR(fb);
  // loads of both cells first: a load after a store to a possibly aliasing
  // output would wait for it
  real_t xp[2], xm[2], yp[2], ym[2], zp[2], zm[2];
#pragma unroll
  for (int c = 0; c < 2; c++)
    if (c < nk) {
      const int k = k0 + c;
      xp[c] = phi(i + 1, j, k);
      xm[c] = phi(i - 1, j, k);
      yp[c] = phi(i, j + 1, k);
      ym[c] = phi(i, j - 1, k);
      zp[c] = phi(i, j, k + 1);
      zm[c] = phi(i, j, k - 1);
    }
#pragma unroll
  for (int c = 0; c < 2; c++)
    if (c < nk) {
      const int k = k0 + c;
      out.o[0](i, j, k) = (xp[c] - xm[c]) * fx;
      out.o[1](i, j, k) = (yp[c] - ym[c]) * fy;
      out.o[2](i, j, k) = (zp[c] - zm[c]) * fz;
    }
}

__global__ void __launch_bounds__(256) central4_kernel(Box fb, Axes3 out, View<const real_t> phi, real_t fx,
                                                       real_t fy, real_t fz) {
  LSMHIP_CELL(fb);
  const real_t e = real_t(8);
  real_t f[3][7];  // loads first, as in central2_kernel
#pragma unroll
  for (int d = 0; d < 3; d++) load_line<2>(phi, i, j, k, d, f[d]);
  const real_t r[3] = {fx, fy, fz};
#pragma unroll
  for (int d = 0; d < 3; d++) out.o[d](i, j, k) = (-f[d][5] + e * f[d][4] + f[d][1] - e * f[d][2]) * r[d];
}

__global__ void __launch_bounds__(256) laplacian2_kernel(Box fb, View<real_t> out, View<const real_t> phi,
                                                         real_t ax, real_t ay, real_t az) {
  LSMHIP_CELL_PAIR(fb);
  real_t v[2][7];  // centre, x+, x-, y+, y-, z+, z- of both cells, loaded first
#pragma unroll
  for (int c = 0; c < 2; c++)
    if (c < nk) {
      const int k = k0 + c;
      v[c][0] = phi(i, j, k);
      v[c][1] = phi(i + 1, j, k);
      v[c][2] = phi(i - 1, j, k);
      v[c][3] = phi(i, j + 1, k);
      v[c][4] = phi(i, j - 1, k);
      v[c][5] = phi(i, j, k + 1);
      v[c][6] = phi(i, j, k - 1);
    }
#pragma unroll
  for (int c = 0; c < 2; c++)
    if (c < nk)
      out(i, j, k0 + c) = ax * (v[c][1] - real_t(2) * v[c][0] + v[c][2]) +
                          ay * (v[c][3] - real_t(2) * v[c][0] + v[c][4]) +
                          az * (v[c][5] - real_t(2) * v[c][0] + v[c][6]);
}

__global__ void __launch_bounds__(256) average_grad_kernel(Box fb, Axes3 out, Axes3 plus, Axes3 minus) {
  LSMHIP_CELL(fb);
  real_t p[3], m[3];  // loads first, as in central2_kernel
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

template <int ORDER>
void hj_plus_minus(const char* who, real_t* pxp, real_t* pyp, real_t* pzp, const Box& gp, real_t* pxm, real_t* pym,
                   real_t* pzm, const Box& gm, const real_t* phi, const Box& pg, const Box& fb, real_t dx, real_t dy,
                   real_t dz) {
  if (fb.empty()) return;
  const int w = ORDER == 1 ? 1 : (ORDER == 2 ? 2 : 3);
  if (!check_box(gp, fb, who, "grad_phi_plus") || !check_box(gm, fb, who, "grad_phi_minus") ||
      !check_box(pg, fb.widen(w, w, w), who, "phi") || !check_size(pg, who) || !check_size(gp, who) ||
      !check_size(gm, who))
    return;
  hipLaunchKernelGGL(hj_plus_minus_kernel<ORDER>, grid_for(fb), block2d(), 0, stream(), fb,
                     axes(pxp, pyp, pzp, gp), axes(pxm, pym, pzm, gm), view(phi, pg), real_t(1.0) / dx,
                     real_t(1.0) / dy, real_t(1.0) / dz);
  finish(who);
}

template <int ORDER>
void hj_upwind(const char* who, real_t* px, real_t* py, real_t* pz, const Box& g, const real_t* phi, const Box& pg,
               const real_t* vx, const real_t* vy, const real_t* vz, const Box& vg, const Box& fb, real_t dx,
               real_t dy, real_t dz) {
  if (fb.empty()) return;
  const int w = ORDER == 1 ? 1 : (ORDER == 2 ? 2 : 3);
  if (!check_box(g, fb, who, "grad_phi") || !check_box(vg, fb, who, "vel") ||
      !check_box(pg, fb.widen(w, w, w), who, "phi") || !check_size(pg, who) || !check_size(g, who) ||
      !check_size(vg, who))
    return;
  hipLaunchKernelGGL(hj_upwind_kernel<ORDER>, grid_for_pair(fb), block2d(), 0, stream(), fb, axes(px, py, pz, g),
                     view(phi, pg), view(vx, vg), view(vy, vg), view(vz, vg), real_t(1.0) / dx, real_t(1.0) / dy,
                     real_t(1.0) / dz);
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
  hipLaunchKernelGGL(central2_kernel, grid_for_pair(fb), block2d(), 0, stream(), fb, axes(px, py, pz, g), view(phi, pg),
                     real_t(0.5) / *dx, real_t(0.5) / *dy, real_t(0.5) / *dz);
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
  hipLaunchKernelGGL(central4_kernel, grid_for(fb), block2d(), 0, stream(), fb, axes(px, py, pz, g), view(phi, pg),
                     c / *dx, c / *dy, c / *dz);
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
  hipLaunchKernelGGL(laplacian2_kernel, grid_for_pair(fb), block2d(), 0, stream(), fb, view(lap, lg), view(phi, pg),
                     real_t(1.0) / *dx / *dx, real_t(1.0) / *dy / *dy, real_t(1.0) / *dz / *dz);
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
