// toolbox2d.hip -- the 2D twins of the whole-grid path: HJ-ENO1/2/3 and
// HJ-WENO5 derivatives (plain and velocity-upwinded), central gradients,
// Laplacian, averaged and magnitude of the gradient, level-set RHS terms,
// TVD Runge-Kutta stages, reinitialization and orthogonalization RHS, mean
// curvature (orders 2 and 4, and from a given Hessian), unit normals,
// field-extension RHS, CFL steps, max-norm difference, smoothed area /
// perimeter / volume and surface integrals, and ghost-cell fills.
// References: src/toolbox/lsm_spatial_derivatives2d.f.in,
// src/toolbox/lsm_level_set_evolution2d.f.in, src/toolbox/lsm_tvd_runge_kutta2d.f,
// src/reinitialization/lsm_reinitialization2d.f.in, src/geometry/lsm_curvature2d.f.in,
// src/geometry/lsm_geometry2d.f.in, src/toolbox/lsm_math_utils2d.f.in,
// src/field_extension/lsm_field_extension2d.f,
// src/boundary_conditions/lsm_boundary_conditions2d.f.
//
// A 2D array over (i, j) is addressed as a 3D array of one plane (k = 0), so
// the 64 x 4 one-thread-per-cell launch shape and the per-axis stencil forms of
// the 3D drop-ins carry over; only the formulas that differ in 2D are written
// here (two-term sums, the 2D zero-tolerance comparisons, the 2D smoothed
// signs).  FP64 results are bitwise the reference's, and the integrals are
// FP64 compensated sums as in 3D.
#include "lsmhip_core.hpp"
#include "reduce.hpp"
#include "stencil.hpp"

using namespace lsmhip;

#define LSMHIP_BOXP2(n) const int *ilo_##n, const int *ihi_##n, const int *jlo_##n, const int *jhi_##n
#define LSMHIP_BOX2(n) ::lsmhip::Box{*ilo_##n, *ihi_##n, *jlo_##n, *jhi_##n, 0, 0}

namespace {

using CV = View<const real_t>;
using MV = View<real_t>;

bool inside(const char* who, const Box& fb, std::initializer_list<std::pair<Box, const char*>> arrays) {
  for (auto& a : arrays)
    if (!check_box(a.first, fb, who, a.second) || !check_size(a.first, who)) return false;
  return true;
}

// phi read w cells away along i and along j (a cross-shaped stencil).
bool cross_inside(const char* who, const Box& fb, const Box& g, int w, const char* what) {
  return check_box(g, fb.widen(w, 0, 0), who, what) && check_box(g, fb.widen(0, w, 0), who, what) &&
         check_size(g, who);
}

// Loads phi at offsets -R..R along axis d (f[s+3] holds offset s).
template <int R>
__device__ __forceinline__ void line2(const CV& phi, int i, int j, int d, real_t f[7]) {
#pragma unroll
  for (int s = -R; s <= R; s++) f[s + 3] = phi(i + (d == 0 ? s : 0), j + (d == 1 ? s : 0), 0);
}
constexpr int reach(int order) { return order == 1 ? 1 : (order == 2 ? 2 : 3); }

// ---------------------------------------------------------------------------
// Spatial derivatives (lsm_spatial_derivatives2d.f.in)
// ---------------------------------------------------------------------------
template <int ORDER>
__global__ void __launch_bounds__(256) hj_pm2_kernel(Box fb, MV xp, MV yp, MV xm, MV ym, CV phi, real_t ix,
                                                     real_t iy) {
  LSMHIP_CELL(fb);
  // Both lines are loaded before the first store: an output may alias phi as
  // far as the compiler knows, so a load after a store would wait for it.
  real_t f[7], g[7], p, m;
  line2<reach(ORDER)>(phi, i, j, 0, f);
  line2<reach(ORDER)>(phi, i, j, 1, g);
  stencil::plus_minus<ORDER>(f, ix, p, m);
  xp(i, j, k) = p;
  xm(i, j, k) = m;
  stencil::plus_minus<ORDER>(g, iy, p, m);
  yp(i, j, k) = p;
  ym(i, j, k) = m;
}

template <int ORDER>
__global__ void __launch_bounds__(256) hj_up2_kernel(Box fb, MV ox, MV oy, CV phi, CV vx, CV vy, real_t ix,
                                                     real_t iy) {
  LSMHIP_CELL(fb);
  const real_t vel[2] = {vx(i, j, k), vy(i, j, k)}, inv[2] = {ix, iy};
  // Both lines are loaded (inside phi's checked box) before the first store,
  // also where the speed below zero_tol leaves one unused.
  real_t f[2][7];
#pragma unroll
  for (int d = 0; d < 2; d++) line2<reach(ORDER)>(phi, i, j, d, f[d]);
#pragma unroll
  for (int d = 0; d < 2; d++)
    (d == 0 ? ox : oy)(i, j, k) =
        fabs(vel[d]) < kZeroTol ? real_t(0) : stencil::upwind<ORDER>(f[d], inv[d], vel[d] > real_t(0));
}

__global__ void __launch_bounds__(256) central2_2d_kernel(Box fb, MV ox, MV oy, CV phi, real_t fx, real_t fy) {
  LSMHIP_CELL(fb);
  const real_t xp = phi(i + 1, j, k), xm = phi(i - 1, j, k), yp = phi(i, j + 1, k), ym = phi(i, j - 1, k);
  ox(i, j, k) = (xp - xm) * fx;
  oy(i, j, k) = (yp - ym) * fy;
}

__global__ void __launch_bounds__(256) central4_2d_kernel(Box fb, MV ox, MV oy, CV phi, real_t fx, real_t fy) {
  LSMHIP_CELL(fb);
  const real_t e = real_t(8);
  const real_t a2 = phi(i + 2, j, k), a1 = phi(i + 1, j, k), b2 = phi(i - 2, j, k), b1 = phi(i - 1, j, k);
  const real_t c2 = phi(i, j + 2, k), c1 = phi(i, j + 1, k), d2 = phi(i, j - 2, k), d1 = phi(i, j - 1, k);
  ox(i, j, k) = (-a2 + e * a1 + b2 - e * b1) * fx;
  oy(i, j, k) = (-c2 + e * c1 + d2 - e * d1) * fy;
}

__global__ void __launch_bounds__(256) laplacian2_2d_kernel(Box fb, MV out, CV phi, real_t ax, real_t ay) {
  LSMHIP_CELL(fb);
  const real_t c = phi(i, j, k);
  out(i, j, k) = ax * (phi(i + 1, j, k) - real_t(2) * c + phi(i - 1, j, k)) +
                 ay * (phi(i, j + 1, k) - real_t(2) * c + phi(i, j - 1, k));
}

__global__ void __launch_bounds__(256) average_grad_2d_kernel(Box fb, MV ox, MV oy, CV xp, CV yp, CV xm, CV ym) {
  LSMHIP_CELL(fb);
  const real_t a = xp(i, j, k), b = xm(i, j, k), c = yp(i, j, k), d = ym(i, j, k);
  ox(i, j, k) = real_t(0.5) * (a + b);
  oy(i, j, k) = real_t(0.5) * (c + d);
}

// lsm_spatial_derivatives2d.f.in:2395-2448: both cut-offs are zero_tol in 2D.
__global__ void __launch_bounds__(256) grad_mag_2d_kernel(Box fb, CV x, CV y, MV mag) {
  LSMHIP_CELL(fb);
  real_t sx = x(i, j, k) * x(i, j, k), sy = y(i, j, k) * y(i, j, k);
  if (sx < kZeroTol) sx = real_t(0);
  if (sy < kZeroTol) sy = real_t(0);
  const real_t t = sqrt(sx + sy);
  mag(i, j, k) = t < kZeroTol ? real_t(0) : t;
}

// ---------------------------------------------------------------------------
// Level-set RHS terms (lsm_level_set_evolution2d.f.in)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) advection_2d_kernel(Box fb, MV rhs, CV px, CV py, CV vx, CV vy) {
  LSMHIP_CELL(fb);
  rhs(i, j, k) = rhs(i, j, k) - (vx(i, j, k) * px(i, j, k) + vy(i, j, k) * py(i, j, k));
}

// Godunov |grad phi|^2, :157-166 and :240-249.
__device__ __forceinline__ real_t godunov2(bool vpos, real_t xp, real_t yp, real_t xm, real_t ym) {
  const real_t z = real_t(0);
  if (vpos) return maxr(sq(maxr(xm, z)), sq(minr(xp, z))) + maxr(sq(maxr(ym, z)), sq(minr(yp, z)));
  return maxr(sq(minr(xm, z)), sq(maxr(xp, z))) + maxr(sq(minr(ym, z)), sq(maxr(yp, z)));
}

// vn.p null: constant speed vconst (the caller skips |vconst| < zero_tol).
__global__ void __launch_bounds__(256) normal_2d_kernel(Box fb, MV rhs, CV xp, CV yp, CV xm, CV ym, CV vn,
                                                        real_t vconst) {
  LSMHIP_CELL(fb);
  const real_t v = vn.p ? vn(i, j, k) : vconst;
  if (fabs(v) >= kZeroTol) {
    const real_t s = godunov2(v > real_t(0), xp(i, j, k), yp(i, j, k), xm(i, j, k), ym(i, j, k));
    rhs(i, j, k) = rhs(i, j, k) - v * sqrt(s);
  }
}

__global__ void __launch_bounds__(256) const_curv_2d_kernel(Box fb, MV rhs, CV px, CV py, CV pxx, CV pxy, CV pyy,
                                                            real_t b) {
  LSMHIP_CELL(fb);
  const real_t x = px(i, j, k), y = py(i, j, k);
  const real_t g = x * x + y * y;
  real_t curv = real_t(0);
  if (!(g < kZeroTol)) {
    curv = pxx(i, j, k) * y * y + pyy(i, j, k) * x * x - real_t(2) * pxy(i, j, k) * x * y;
    curv = curv / g;
  }
  rhs(i, j, k) = rhs(i, j, k) + b * curv;
}

__global__ void __launch_bounds__(256) precomputed_curv_2d_kernel(Box lb, MV rhs, CV kappa, CV gmag, real_t b) {
  LSMHIP_CELL(lb);
  rhs(i, j, k) = rhs(i, j, k) + b * kappa(i, j, k) * gmag(i, j, k);
}

// :399-522: the advection and normal parts are added only where |V| >= zero_tol.
__global__ void __launch_bounds__(256) ext_normal_2d_kernel(Box fb, MV rhs, CV xp, CV yp, CV xm, CV ym, CV vn, CV vx,
                                                            CV vy) {
  LSMHIP_CELL(fb);
  const real_t n = vn(i, j, k), u = vx(i, j, k), v = vy(i, j, k);
  const real_t px = stencil::ext_normal_pick(u, n, xp(i, j, k), xm(i, j, k));
  const real_t py = stencil::ext_normal_pick(v, n, yp(i, j, k), ym(i, j, k));
  const real_t s = px * px + py * py;
  if (fabs(n) >= kZeroTol) rhs(i, j, k) = rhs(i, j, k) - u * px - v * py - n * sqrt(s);
}

// RK kinds as in 3D: 1 = u + dt L; 2 = 0.5 (u + u1 + dt L); 32 = 0.75 u + 0.25 (u1 + dt L);
// 33 = 1/3 u + 2/3 (u2 + dt L)  (lsm_tvd_runge_kutta2d.f:68, :202, :340, :412).
template <int KIND>
__global__ void __launch_bounds__(256) rk_2d_kernel(Box fb, MV out, CV us, CV uc, CV rhs, real_t dt) {
  LSMHIP_CELL(fb);
  const real_t r = rhs(i, j, k);
  if (KIND == 1) out(i, j, k) = uc(i, j, k) + dt * r;
  if (KIND == 2) out(i, j, k) = real_t(0.5) * (uc(i, j, k) + us(i, j, k) + dt * r);
  if (KIND == 32) out(i, j, k) = real_t(0.75) * uc(i, j, k) + real_t(0.25) * (us(i, j, k) + dt * r);
  if (KIND == 33) out(i, j, k) = real_t(1.0 / 3.0) * uc(i, j, k) + real_t(2.0 / 3.0) * (us(i, j, k) + dt * r);
}

// ---------------------------------------------------------------------------
// Reinitialization and orthogonalization (lsm_reinitialization2d.f.in)
// ---------------------------------------------------------------------------
// :48-227 -- zero where |sign source| <= zero_tol (strict, unlike the 3D >=).
__global__ void __launch_bounds__(256) reinit_2d_kernel(Box fb, MV out, CV phi, CV phi0, CV xp, CV yp, CV xm, CV ym,
                                                        real_t dx_sq, int use_phi0) {
  LSMHIP_CELL(fb);
  const real_t s = use_phi0 ? phi0(i, j, k) : phi(i, j, k);
  const real_t gp[2] = {xp(i, j, k), yp(i, j, k)}, gm[2] = {xm(i, j, k), ym(i, j, k)};
  real_t star[2];
#pragma unroll
  for (int d = 0; d < 2; d++) {
    real_t p, m;
    if (s > real_t(0)) {
      p = maxr(-gp[d], real_t(0));
      m = maxr(gm[d], real_t(0));
    } else {
      p = maxr(gp[d], real_t(0));
      m = maxr(-gm[d], real_t(0));
    }
    star[d] = maxr(p, m);
  }
  real_t r = real_t(0);
  if (fabs(s) > kZeroTol) {
    const real_t n2 = star[0] * star[0] + star[1] * star[1];
    const real_t sgn = use_phi0 ? s / sqrt(s * s + dx_sq) : s / sqrt(s * s + n2 * dx_sq);
    r = sgn * (real_t(1) - sqrt(n2));
  }
  out(i, j, k) = r;
}

// :255-386 -- the smoothed sign of psi is psi / sqrt(psi^2 + max(dx,dy)^2).
__global__ void __launch_bounds__(256) ortho_2d_kernel(Box fb, MV rhs, CV xp, CV yp, CV xm, CV ym, CV psi, CV qxp,
                                                       CV qyp, CV qxm, CV qym, real_t dx_sq) {
  LSMHIP_CELL(fb);
  const real_t g1 = real_t(0.5) * (qxp(i, j, k) + qxm(i, j, k));
  const real_t g2 = real_t(0.5) * (qyp(i, j, k) + qym(i, j, k));
  const real_t ng = sqrt(g1 * g1 + g2 * g2);
  const real_t s = psi(i, j, k);
  real_t out = real_t(0);
  if (fabs(s) > kZeroTol && ng > kZeroTol) {
    const real_t sgn = s / sqrt(s * s + dx_sq);
    const real_t a1 = sgn * g1, a2 = sgn * g2, z = real_t(0);
    out = real_t(-1) / ng *
          (maxr(a1, z) * xm(i, j, k) + minr(a1, z) * xp(i, j, k) + maxr(a2, z) * ym(i, j, k) +
           minr(a2, z) * yp(i, j, k));
  }
  rhs(i, j, k) = out;
}

// ---------------------------------------------------------------------------
// Curvature and normals (lsm_curvature2d.f.in, lsm_geometry2d.f.in)
// ---------------------------------------------------------------------------
// Order 2 (curvature2d :15-92) and order 4 (:252-337, phi_xy = d/dy phi_x).
template <int ORDER>
__global__ void __launch_bounds__(256) mean_curv_2d_kernel(Box fb, MV kappa, CV F, CV px, CV py, CV gmag, real_t fxx,
                                                           real_t fyy, real_t fxy) {
  LSMHIP_CELL(fb);
  const real_t c = F(i, j, k);
  real_t xx, yy, xy;
  if (ORDER == 2) {
    xx = (F(i + 1, j, k) - real_t(2) * c + F(i - 1, j, k)) * fxx;
    yy = (F(i, j + 1, k) - real_t(2) * c + F(i, j - 1, k)) * fyy;
    xy = (F(i + 1, j + 1, k) - F(i + 1, j - 1, k) - F(i - 1, j + 1, k) + F(i - 1, j - 1, k)) * fxy;
  } else {
    xx = (-F(i + 2, j, k) + real_t(16) * F(i + 1, j, k) - real_t(30) * c - F(i - 2, j, k) +
          real_t(16) * F(i - 1, j, k)) * fxx;
    yy = (-F(i, j + 2, k) + real_t(16) * F(i, j + 1, k) - real_t(30) * c - F(i, j - 2, k) +
          real_t(16) * F(i, j - 1, k)) * fyy;
    xy = (-px(i, j + 2, k) + real_t(8) * px(i, j + 1, k) + px(i, j - 2, k) - real_t(8) * px(i, j - 1, k)) * fxy;
  }
  const real_t g = gmag(i, j, k), den = g * g * g;
  real_t out = real_t(0);
  if (!(den < kZeroTol)) {
    const real_t x = px(i, j, k), y = py(i, j, k);
    out = (xx * y * y + yy * x * x - real_t(2) * xy * x * y) / den;
  }
  kappa(i, j, k) = out;
}

// geometry2d :218-286, over (|grad phi|^2)^1.5
__global__ void __launch_bounds__(256) mean_curv_hessian_2d_kernel(Box fb, MV kappa, CV px, CV py, CV pxx, CV pxy,
                                                                   CV pyy) {
  LSMHIP_CELL(fb);
  const real_t x = px(i, j, k), y = py(i, j, k), ngs = x * x + y * y;
  real_t out = real_t(0);
  if (!(ngs < kZeroTol))
    out = (pxx(i, j, k) * y * y - real_t(2) * pxy(i, j, k) * x * y + pyy(i, j, k) * x * x) /
          stencil::pow_three_halves(ngs);
  kappa(i, j, k) = out;
}

// geometry2d :22-83 (SIGNED = 0) and :103-195 (SIGNED = 1)
template <int SIGNED>
__global__ void __launch_bounds__(256) normal_2d_kernel_n(Box fb, MV nx, MV ny, CV px, CV py, CV phi, real_t dx_sq) {
  LSMHIP_CELL(fb);
  const real_t x = px(i, j, k), y = py(i, j, k);
  real_t ox = real_t(1), oy = real_t(0);
  if (SIGNED) {
    const real_t c = phi(i, j, k);
    if (fabs(c) > kZeroTol) {
      const real_t ngs = x * x + y * y;
      if (ngs >= kZeroTol) {
        const real_t sgn = c / sqrt(c * c + ngs * dx_sq), inv = real_t(1) / sqrt(ngs);
        ox = sgn * x * inv;
        oy = sgn * y * inv;
      }
    } else {
      ox = real_t(0);
    }
  } else {
    const real_t norm = sqrt(x * x + y * y);
    if (norm >= kZeroTol) {
      const real_t inv = real_t(1) / norm;
      ox = x * inv;
      oy = y * inv;
    }
  }
  nx(i, j, k) = ox;
  ny(i, j, k) = oy;
}

// lsm_field_extension2d.f:31-125
__global__ void __launch_bounds__(256) field_ext_2d_kernel(Box fb, MV rhs, CV P, CV sx, CV sy, CV nx, CV ny,
                                                           real_t cutoff) {
  LSMHIP_CELL(fb);
  const real_t c = P(i, j, k);
  const bool away = fabs(c) > cutoff || (c * P(i - 1, j, k) > real_t(0) && c * P(i + 1, j, k) > real_t(0) &&
                                         c * P(i, j - 1, k) > real_t(0) && c * P(i, j + 1, k) > real_t(0));
  rhs(i, j, k) = away ? -(nx(i, j, k) * sx(i, j, k) + ny(i, j, k) * sy(i, j, k)) : real_t(0);
}

// ---------------------------------------------------------------------------
// Reductions (lsm_math_utils2d.f.in, lsm_geometry2d.f.in)
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;
constexpr double kOneOverPi = 0.31830988618379;  // the reference's truncated constant

struct MaxNormDiff2 {
  CV a, b;
  __device__ double operator()(int i, int j, int k) const { return (double)fabs(a(i, j, k) - b(i, j, k)); }
};
struct AdvectionCfl2 {
  CV u, v;
  real_t ix, iy;
  __device__ double operator()(int i, int j, int k) const {
    return (double)(fabs(u(i, j, k)) * ix + fabs(v(i, j, k)) * iy);
  }
};
struct NormalCfl2 {
  CV vn, xp, yp, xm, ym;
  real_t vconst, mdx2, ix, iy;
  __device__ double operator()(int i, int j, int k) const {
    const real_t av = vn.p ? fabs(vn(i, j, k)) : vconst;
    const real_t x = maxr(fabs(xp(i, j, k)), fabs(xm(i, j, k))), y = maxr(fabs(yp(i, j, k)), fabs(ym(i, j, k)));
    return (double)(av / sqrt(x * x + y * y + mdx2) * (x * ix + y * iy));
  }
};
// sign = -1: smoothed measure of {phi < 0}; +1: of {phi > 0}.  F.p null means F = 1.
struct Region2 {
  CV phi, F;
  real_t eps, dA;
  int sign;
  __device__ double operator()(int i, int j, int k) const {
    // q = phi / eps only inside the smoothing band, where it is used
    const real_t p = phi(i, j, k);
    const real_t f = F.p ? F(i, j, k) : real_t(1);
    if (sign < 0) {
      if (p < -eps) return (double)(F.p ? f * dA : dA);
      if (p < eps) {
        const real_t q = p / eps;
        const real_t h = real_t(0.5) * (real_t(1) - q - real_t(kOneOverPi) * sin(real_t(kPi) * q));
        return (double)(F.p ? h * f * dA : h * dA);
      }
    } else {
      if (p > eps) return (double)(F.p ? f * dA : dA);
      if (p > -eps) {
        const real_t q = p / eps;
        const real_t h = real_t(0.5) * (real_t(1) + q + real_t(kOneOverPi) * sin(real_t(kPi) * q));
        return (double)(F.p ? h * f * dA : h * dA);
      }
    }
    return 0.0;
  }
};
struct Surface2 {
  CV phi, F, x, y;
  real_t eps, ie, dA;
  __device__ double operator()(int i, int j, int k) const {
    const real_t p = phi(i, j, k);
    if (!(fabs(p) < eps)) return 0.0;
    const real_t delta = real_t(0.5) * ie * (real_t(1) + cos(real_t(kPi) * p * ie));
    const real_t gx = x(i, j, k), gy = y(i, j, k);
    const real_t n = sqrt(gx * gx + gy * gy);
    return (double)(F.p ? delta * n * F(i, j, k) * dA : delta * n * dA);
  }
};

real_t cfl_dt(double m, real_t cfl) { return cfl / (real_t(m) + kTiny); }

// ---------------------------------------------------------------------------
// Ghost fills (lsm_boundary_conditions2d.f): face idx 0..3, one thread per
// ghost column; each face sweeps the other axis' full ghost range.
// ---------------------------------------------------------------------------
enum { kLinear = 0, kSigned = 1, kCopy = 2 };

__global__ void __launch_bounds__(256) face_2d_kernel(MV P, Box gb, Box fb, int idx, int kind) {
  const int d = idx / 2, hi = idx % 2, a = 1 - d;
  const int lo_g[2] = {gb.ilo, gb.jlo}, hi_g[2] = {gb.ihi, gb.jhi};
  const int lo_f[2] = {fb.ilo, fb.jlo}, hi_f[2] = {fb.ihi, fb.jhi};
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > hi_g[a] - lo_g[a]) return;
  const int ta = lo_g[a] + t;
  const int edge = hi ? hi_f[d] : lo_f[d], inner = hi ? hi_f[d] - 1 : lo_f[d] + 1;
  const int g0 = hi ? hi_f[d] + 1 : lo_g[d], g1 = hi ? hi_g[d] : lo_f[d] - 1;
  int e[2], n[2];
  e[d] = edge;
  e[a] = ta;
  n[d] = inner;
  n[a] = ta;
  const real_t pb = P(e[0], e[1], 0);
  real_t slope = real_t(0);
  if (kind == kLinear) slope = pb - P(n[0], n[1], 0);
  if (kind == kSigned) slope = copysign(real_t(1), pb) * fabs(pb - P(n[0], n[1], 0));
  for (int g = g0; g <= g1; g++) {
    int q[2];
    q[d] = g;
    q[a] = ta;
    P(q[0], q[1], 0) = kind == kCopy ? pb : pb + slope * real_t(hi ? g - edge : edge - g);
  }
}

void face2(const char* who, real_t* phi, const Box& gb, const Box& fb, int idx, int kind) {
  if (idx < 0 || idx > 3) return;  // the reference ignores other indices
  if (gb.empty() || fb.empty()) return;
  if (!check_box(gb, fb, who, "phi") || !check_size(gb, who)) return;
  const int d = idx / 2;
  const int lo_f[2] = {fb.ilo, fb.jlo}, hi_f[2] = {fb.ihi, fb.jhi};
  const int lo_g[2] = {gb.ilo, gb.jlo}, hi_g[2] = {gb.ihi, gb.jhi};
  const int depth = idx % 2 ? hi_g[d] - hi_f[d] : lo_f[d] - lo_g[d];
  if (depth <= 0) return;
  if (kind != kCopy) {
    const int inner = idx % 2 ? hi_f[d] - 1 : lo_f[d] + 1;  // the slope reads this layer
    if (inner < lo_g[d] || inner > hi_g[d]) {
      set_error(kErrBox, "%s: the layer next to face %d lies outside the ghost box", who, idx);
      return;
    }
  }
  const int lines = hi_g[1 - d] - lo_g[1 - d] + 1;
  hipLaunchKernelGGL(face_2d_kernel, dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, stream(), view(phi, gb), gb,
                     fb, idx, kind);
  finish(who);
}

template <int KIND>
void rk2d(const char* who, real_t* out, const Box& og, const real_t* us, const Box& sg, const real_t* uc,
          const Box& cg, const real_t* rhs, const Box& rg, const Box& fb, real_t dt) {
  if (fb.empty() || !inside(who, fb, {{og, "u_next"}, {cg, "u_cur"}, {rg, "rhs"}})) return;
  if (KIND != 1 && !inside(who, fb, {{sg, "u_stage"}})) return;
  hipLaunchKernelGGL(rk_2d_kernel<KIND>, grid_for(fb), block2d(), 0, stream(), fb, view(out, og), view(us, sg),
                     view(uc, cg), view(rhs, rg), dt);
  finish(who);
}

template <int ORDER>
void hj_pm2(const char* who, real_t* xp, real_t* yp, const Box& gp, real_t* xm, real_t* ym, const Box& gm,
            const real_t* phi, const Box& pg, const Box& fb, real_t dx, real_t dy) {
  if (fb.empty() || !inside(who, fb, {{gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}}) ||
      !cross_inside(who, fb, pg, reach(ORDER), "phi"))
    return;
  hipLaunchKernelGGL(hj_pm2_kernel<ORDER>, grid_for(fb), block2d(), 0, stream(), fb, view(xp, gp), view(yp, gp),
                     view(xm, gm), view(ym, gm), view(phi, pg), real_t(1.0) / dx, real_t(1.0) / dy);
  finish(who);
}

template <int ORDER>
void hj_up2(const char* who, real_t* ox, real_t* oy, const Box& g, const real_t* phi, const Box& pg,
            const real_t* vx, const real_t* vy, const Box& vg, const Box& fb, real_t dx, real_t dy) {
  if (fb.empty() || !inside(who, fb, {{g, "grad_phi"}, {vg, "vel"}}) ||
      !cross_inside(who, fb, pg, reach(ORDER), "phi"))
    return;
  hipLaunchKernelGGL(hj_up2_kernel<ORDER>, grid_for(fb), block2d(), 0, stream(), fb, view(ox, g), view(oy, g),
                     view(phi, pg), view(vx, vg), view(vy, vg), real_t(1.0) / dx, real_t(1.0) / dy);
  finish(who);
}

template <class Op>
bool box_max2(const char* who, const Box& ib, const Op& op, double* m) {
  return reduce::box_max(who, ib, op, m);
}

}  // namespace

// ===========================================================================
// Entry points: spatial derivatives (src/toolbox/lsm_spatial_derivatives2d.h)
// ===========================================================================
LSMHIP_EXPORT void lsm2dhjeno1_(real_t* pxp, real_t* pyp, LSMHIP_BOXP2(gp), real_t* pxm, real_t* pym, LSMHIP_BOXP2(gm),
                                const real_t* phi, LSMHIP_BOXP2(pg), real_t* D1, LSMHIP_BOXP2(d1), LSMHIP_BOXP2(fb),
                                const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1);
  (void)D1;
  hj_pm2<1>("lsm2dhjeno1_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
            LSMHIP_BOX2(fb), *dx, *dy);
}
LSMHIP_EXPORT void lsm2dhjeno2_(real_t* pxp, real_t* pyp, LSMHIP_BOXP2(gp), real_t* pxm, real_t* pym, LSMHIP_BOXP2(gm),
                                const real_t* phi, LSMHIP_BOXP2(pg), real_t* D1, LSMHIP_BOXP2(d1), real_t* D2,
                                LSMHIP_BOXP2(d2), LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1, D2);
  (void)D1;
  (void)D2;
  hj_pm2<2>("lsm2dhjeno2_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
            LSMHIP_BOX2(fb), *dx, *dy);
}
LSMHIP_EXPORT void lsm2dhjeno3_(real_t* pxp, real_t* pyp, LSMHIP_BOXP2(gp), real_t* pxm, real_t* pym, LSMHIP_BOXP2(gm),
                                const real_t* phi, LSMHIP_BOXP2(pg), real_t* D1, LSMHIP_BOXP2(d1), real_t* D2,
                                LSMHIP_BOXP2(d2), real_t* D3, LSMHIP_BOXP2(d3), LSMHIP_BOXP2(fb), const real_t* dx,
                                const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1, D2, D3);
  (void)D1;
  (void)D2;
  (void)D3;
  hj_pm2<3>("lsm2dhjeno3_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
            LSMHIP_BOX2(fb), *dx, *dy);
}
LSMHIP_EXPORT void lsm2dhjweno5_(real_t* pxp, real_t* pyp, LSMHIP_BOXP2(gp), real_t* pxm, real_t* pym,
                                 LSMHIP_BOXP2(gm), const real_t* phi, LSMHIP_BOXP2(pg), real_t* D1, LSMHIP_BOXP2(d1),
                                 LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1);
  (void)D1;
  hj_pm2<5>("lsm2dhjweno5_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
            LSMHIP_BOX2(fb), *dx, *dy);
}
LSMHIP_EXPORT void lsm2dupwindhjeno1_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi, LSMHIP_BOXP2(pg),
                                      const real_t* vx, const real_t* vy, LSMHIP_BOXP2(vg), real_t* D1,
                                      LSMHIP_BOXP2(d1), LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi, vx, vy, D1);
  (void)D1;
  hj_up2<1>("lsm2dupwindhjeno1_", px, py, LSMHIP_BOX2(g), phi, LSMHIP_BOX2(pg), vx, vy, LSMHIP_BOX2(vg),
            LSMHIP_BOX2(fb), *dx, *dy);
}
LSMHIP_EXPORT void lsm2dupwindhjeno2_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi, LSMHIP_BOXP2(pg),
                                      const real_t* vx, const real_t* vy, LSMHIP_BOXP2(vg), real_t* D1,
                                      LSMHIP_BOXP2(d1), real_t* D2, LSMHIP_BOXP2(d2), LSMHIP_BOXP2(fb),
                                      const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi, vx, vy, D1, D2);
  (void)D1;
  (void)D2;
  hj_up2<2>("lsm2dupwindhjeno2_", px, py, LSMHIP_BOX2(g), phi, LSMHIP_BOX2(pg), vx, vy, LSMHIP_BOX2(vg),
            LSMHIP_BOX2(fb), *dx, *dy);
}
LSMHIP_EXPORT void lsm2dupwindhjeno3_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi, LSMHIP_BOXP2(pg),
                                      const real_t* vx, const real_t* vy, LSMHIP_BOXP2(vg), real_t* D1,
                                      LSMHIP_BOXP2(d1), real_t* D2, LSMHIP_BOXP2(d2), real_t* D3, LSMHIP_BOXP2(d3),
                                      LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi, vx, vy, D1, D2, D3);
  (void)D1;
  (void)D2;
  (void)D3;
  hj_up2<3>("lsm2dupwindhjeno3_", px, py, LSMHIP_BOX2(g), phi, LSMHIP_BOX2(pg), vx, vy, LSMHIP_BOX2(vg),
            LSMHIP_BOX2(fb), *dx, *dy);
}
LSMHIP_EXPORT void lsm2dupwindhjweno5_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi, LSMHIP_BOXP2(pg),
                                       const real_t* vx, const real_t* vy, LSMHIP_BOXP2(vg), real_t* D1,
                                       LSMHIP_BOXP2(d1), LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi, vx, vy, D1);
  (void)D1;
  hj_up2<5>("lsm2dupwindhjweno5_", px, py, LSMHIP_BOX2(g), phi, LSMHIP_BOX2(pg), vx, vy, LSMHIP_BOX2(vg),
            LSMHIP_BOX2(fb), *dx, *dy);
}

LSMHIP_EXPORT void lsm2dcentralgradorder2_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi,
                                           LSMHIP_BOXP2(pg), LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi);
  const char* who = "lsm2dcentralgradorder2_";
  Box fb = LSMHIP_BOX2(fb), g = LSMHIP_BOX2(g), pg = LSMHIP_BOX2(pg);
  if (fb.empty() || !inside(who, fb, {{g, "grad_phi"}}) || !cross_inside(who, fb, pg, 1, "phi")) return;
  hipLaunchKernelGGL(central2_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(px, g), view(py, g),
                     view(phi, pg), real_t(0.5) / *dx, real_t(0.5) / *dy);
  finish(who);
}
LSMHIP_EXPORT void lsm2dcentralgradorder4_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi,
                                           LSMHIP_BOXP2(pg), LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi);
  const char* who = "lsm2dcentralgradorder4_";
  Box fb = LSMHIP_BOX2(fb), g = LSMHIP_BOX2(g), pg = LSMHIP_BOX2(pg);
  if (fb.empty() || !inside(who, fb, {{g, "grad_phi"}}) || !cross_inside(who, fb, pg, 2, "phi")) return;
  const real_t c = real_t(0.0833333333333333333333);
  hipLaunchKernelGGL(central4_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(px, g), view(py, g),
                     view(phi, pg), c / *dx, c / *dy);
  finish(who);
}
LSMHIP_EXPORT void lsm2dlaplacianorder2_(real_t* lap, LSMHIP_BOXP2(lg), const real_t* phi, LSMHIP_BOXP2(pg),
                                         LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(lap, phi);
  const char* who = "lsm2dlaplacianorder2_";
  Box fb = LSMHIP_BOX2(fb), lg = LSMHIP_BOX2(lg), pg = LSMHIP_BOX2(pg);
  if (fb.empty() || !inside(who, fb, {{lg, "laplacian_phi"}}) || !cross_inside(who, fb, pg, 1, "phi")) return;
  hipLaunchKernelGGL(laplacian2_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(lap, lg), view(phi, pg),
                     real_t(1.0) / *dx / *dx, real_t(1.0) / *dy / *dy);
  finish(who);
}
LSMHIP_EXPORT void lsm2daveragegradphi_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* pxp,
                                        const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm, const real_t* pym,
                                        LSMHIP_BOXP2(gm), LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pxp, pyp, pxm, pym);
  const char* who = "lsm2daveragegradphi_";
  Box fb = LSMHIP_BOX2(fb), g = LSMHIP_BOX2(g), gp = LSMHIP_BOX2(gp), gm = LSMHIP_BOX2(gm);
  if (fb.empty() || !inside(who, fb, {{g, "grad_phi"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}})) return;
  hipLaunchKernelGGL(average_grad_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(px, g), view(py, g),
                     view(pxp, gp), view(pyp, gp), view(pxm, gm), view(pym, gm));
  finish(who);
}
LSMHIP_EXPORT void lsm2dgradientmagnitude_(const real_t* px, const real_t* py, real_t* mag, LSMHIP_BOXP2(g),
                                           LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(px, py, mag);
  const char* who = "lsm2dgradientmagnitude_";
  Box fb = LSMHIP_BOX2(fb), g = LSMHIP_BOX2(g);
  if (fb.empty() || !inside(who, fb, {{g, "grad_phi"}})) return;
  hipLaunchKernelGGL(grad_mag_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(px, g), view(py, g),
                     view(mag, g));
  finish(who);
}

// ===========================================================================
// Level-set RHS (src/toolbox/lsm_level_set_evolution2d.h)
// ===========================================================================
LSMHIP_EXPORT void lsm2dzerooutlevelseteqnrhs_(real_t* rhs, LSMHIP_BOXP2(rg)) {
  LSMHIP_REQUIRE_ARRAYS(rhs);
  const char* who = "lsm2dzerooutlevelseteqnrhs_";
  Box rg = LSMHIP_BOX2(rg);
  if (rg.empty() || !check_size(rg, who)) return;
  // the whole ghost box, as the reference zeroes it: rg.volume() contiguous scalars from rhs
  if (fill_contiguous(rhs, rg.volume(), real_t(0), who)) finish(who);
}
LSMHIP_EXPORT void lsm2daddadvectiontermtolserhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* px, const real_t* py,
                                                  LSMHIP_BOXP2(g), const real_t* vx, const real_t* vy,
                                                  LSMHIP_BOXP2(vg), LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, vx, vy);
  const char* who = "lsm2daddadvectiontermtolserhs_";
  Box fb = LSMHIP_BOX2(fb), rg = LSMHIP_BOX2(rg), g = LSMHIP_BOX2(g), vg = LSMHIP_BOX2(vg);
  if (fb.empty() || !inside(who, fb, {{rg, "lse_rhs"}, {g, "grad_phi"}, {vg, "vel"}})) return;
  hipLaunchKernelGGL(advection_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(px, g),
                     view(py, g), view(vx, vg), view(vy, vg));
  finish(who);
}
static void normal2(const char* who, real_t* rhs, const Box& rg, const real_t* pxp, const real_t* pyp, const Box& gp,
                    const real_t* pxm, const real_t* pym, const Box& gm, const real_t* vn, const Box& vg,
                    real_t vconst, const Box& fb) {
  if (fb.empty() || !inside(who, fb, {{rg, "lse_rhs"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}})) return;
  if (vn && !inside(who, fb, {{vg, "vel_n"}})) return;
  hipLaunchKernelGGL(normal_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(pxp, gp),
                     view(pyp, gp), view(pxm, gm), view(pym, gm),
                     vn ? view(vn, vg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0}, vconst);
  finish(who);
}
LSMHIP_EXPORT void lsm2daddnormalveltermtolserhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp, const real_t* pyp,
                                                  LSMHIP_BOXP2(gp), const real_t* pxm, const real_t* pym,
                                                  LSMHIP_BOXP2(gm), const real_t* vn, LSMHIP_BOXP2(vg),
                                                  LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym, vn);
  normal2("lsm2daddnormalveltermtolserhs_", rhs, LSMHIP_BOX2(rg), pxp, pyp, LSMHIP_BOX2(gp), pxm, pym,
          LSMHIP_BOX2(gm), vn, LSMHIP_BOX2(vg), real_t(0), LSMHIP_BOX2(fb));
}
LSMHIP_EXPORT void lsm2daddconstnormalveltermtolserhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp,
                                                       const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                       const real_t* pym, LSMHIP_BOXP2(gm), const real_t* vn,
                                                       LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym);
  if (fabs(*vn) < kZeroTol) return;  // lsm_level_set_evolution2d.f.in:231
  const Box none{0, -1, 0, -1, 0, -1};
  normal2("lsm2daddconstnormalveltermtolserhs_", rhs, LSMHIP_BOX2(rg), pxp, pyp, LSMHIP_BOX2(gp), pxm, pym,
          LSMHIP_BOX2(gm), nullptr, none, *vn, LSMHIP_BOX2(fb));
}
LSMHIP_EXPORT void lsm2daddconstcurvtermtolserhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* px, const real_t* py,
                                                  LSMHIP_BOXP2(g), const real_t* pxx, const real_t* pxy,
                                                  const real_t* pyy, LSMHIP_BOXP2(g2), const real_t* b,
                                                  LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, pxx, pxy, pyy);
  const char* who = "lsm2daddconstcurvtermtolserhs_";
  Box fb = LSMHIP_BOX2(fb), rg = LSMHIP_BOX2(rg), g = LSMHIP_BOX2(g), g2 = LSMHIP_BOX2(g2);
  if (fb.empty() || !inside(who, fb, {{rg, "lse_rhs"}, {g, "grad_phi"}, {g2, "grad2_phi"}})) return;
  hipLaunchKernelGGL(const_curv_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(px, g),
                     view(py, g), view(pxx, g2), view(pxy, g2), view(pyy, g2), *b);
  finish(who);
}
// The reference loops j = jlo_rhs_fb..jhi_rhs_gb, i = ilo_rhs_gb..ihi_rhs_gb
// (lsm_level_set_evolution2d.f.in:383-384); kept.
LSMHIP_EXPORT void lsm2daddconstprecomputedcurvtermtolserhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* kappa,
                                                             LSMHIP_BOXP2(kg), const real_t* gmag, LSMHIP_BOXP2(gg),
                                                             const real_t* b, LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, kappa, gmag);
  const char* who = "lsm2daddconstprecomputedcurvtermtolserhs_";
  Box rg = LSMHIP_BOX2(rg), kg = LSMHIP_BOX2(kg), gg = LSMHIP_BOX2(gg);
  Box lb{rg.ilo, rg.ihi, *jlo_fb, rg.jhi, 0, 0};
  if (lb.empty() || !inside(who, lb, {{rg, "lse_rhs"}, {kg, "kappa"}, {gg, "grad_phi_mag"}})) return;
  hipLaunchKernelGGL(precomputed_curv_2d_kernel, grid_for(lb), block2d(), 0, stream(), lb, view(rhs, rg),
                     view(kappa, kg), view(gmag, gg), *b);
  finish(who);
}
LSMHIP_EXPORT void lsm2daddexternalandnormalveltermtolserhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp,
                                                             const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                             const real_t* pym, LSMHIP_BOXP2(gm), const real_t* vn,
                                                             const real_t* vx, const real_t* vy, LSMHIP_BOXP2(vg),
                                                             LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym, vn, vx, vy);
  const char* who = "lsm2daddexternalandnormalveltermtolserhs_";
  Box fb = LSMHIP_BOX2(fb), rg = LSMHIP_BOX2(rg), gp = LSMHIP_BOX2(gp), gm = LSMHIP_BOX2(gm), vg = LSMHIP_BOX2(vg);
  if (fb.empty() || !inside(who, fb, {{rg, "lse_rhs"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}, {vg, "vel"}}))
    return;
  hipLaunchKernelGGL(ext_normal_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(pxp, gp),
                     view(pyp, gp), view(pxm, gm), view(pym, gm), view(vn, vg), view(vx, vg), view(vy, vg));
  finish(who);
}

// ===========================================================================
// TVD Runge-Kutta (src/toolbox/lsm_tvd_runge_kutta2d.h)
// ===========================================================================
LSMHIP_EXPORT void lsm2drk1step_(real_t* un, LSMHIP_BOXP2(ng), const real_t* uc, LSMHIP_BOXP2(cg), const real_t* rhs,
                                 LSMHIP_BOXP2(rg), LSMHIP_BOXP2(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs);
  rk2d<1>("lsm2drk1step_", un, LSMHIP_BOX2(ng), nullptr, LSMHIP_BOX2(ng), uc, LSMHIP_BOX2(cg), rhs, LSMHIP_BOX2(rg),
          LSMHIP_BOX2(fb), *dt);
}
LSMHIP_EXPORT void lsm2dtvdrk2stage1_(real_t* un, LSMHIP_BOXP2(ng), const real_t* uc, LSMHIP_BOXP2(cg),
                                      const real_t* rhs, LSMHIP_BOXP2(rg), LSMHIP_BOXP2(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs);
  rk2d<1>("lsm2dtvdrk2stage1_", un, LSMHIP_BOX2(ng), nullptr, LSMHIP_BOX2(ng), uc, LSMHIP_BOX2(cg), rhs,
          LSMHIP_BOX2(rg), LSMHIP_BOX2(fb), *dt);
}
LSMHIP_EXPORT void lsm2dtvdrk2stage2_(real_t* un, LSMHIP_BOXP2(ng), const real_t* u1, LSMHIP_BOXP2(g1),
                                      const real_t* uc, LSMHIP_BOXP2(cg), const real_t* rhs, LSMHIP_BOXP2(rg),
                                      LSMHIP_BOXP2(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, u1, uc, rhs);
  rk2d<2>("lsm2dtvdrk2stage2_", un, LSMHIP_BOX2(ng), u1, LSMHIP_BOX2(g1), uc, LSMHIP_BOX2(cg), rhs, LSMHIP_BOX2(rg),
          LSMHIP_BOX2(fb), *dt);
}
LSMHIP_EXPORT void lsm2dtvdrk3stage1_(real_t* un, LSMHIP_BOXP2(ng), const real_t* uc, LSMHIP_BOXP2(cg),
                                      const real_t* rhs, LSMHIP_BOXP2(rg), LSMHIP_BOXP2(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs);
  rk2d<1>("lsm2dtvdrk3stage1_", un, LSMHIP_BOX2(ng), nullptr, LSMHIP_BOX2(ng), uc, LSMHIP_BOX2(cg), rhs,
          LSMHIP_BOX2(rg), LSMHIP_BOX2(fb), *dt);
}
LSMHIP_EXPORT void lsm2dtvdrk3stage2_(real_t* u2, LSMHIP_BOXP2(g2), const real_t* u1, LSMHIP_BOXP2(g1),
                                      const real_t* uc, LSMHIP_BOXP2(cg), const real_t* rhs, LSMHIP_BOXP2(rg),
                                      LSMHIP_BOXP2(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(u2, u1, uc, rhs);
  rk2d<32>("lsm2dtvdrk3stage2_", u2, LSMHIP_BOX2(g2), u1, LSMHIP_BOX2(g1), uc, LSMHIP_BOX2(cg), rhs, LSMHIP_BOX2(rg),
           LSMHIP_BOX2(fb), *dt);
}
LSMHIP_EXPORT void lsm2dtvdrk3stage3_(real_t* un, LSMHIP_BOXP2(ng), const real_t* u2, LSMHIP_BOXP2(g2),
                                      const real_t* uc, LSMHIP_BOXP2(cg), const real_t* rhs, LSMHIP_BOXP2(rg),
                                      LSMHIP_BOXP2(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, u2, uc, rhs);
  rk2d<33>("lsm2dtvdrk3stage3_", un, LSMHIP_BOX2(ng), u2, LSMHIP_BOX2(g2), uc, LSMHIP_BOX2(cg), rhs, LSMHIP_BOX2(rg),
           LSMHIP_BOX2(fb), *dt);
}

// ===========================================================================
// Reinitialization / orthogonalization (src/reinitialization/lsm_reinitialization2d.h)
// ===========================================================================
LSMHIP_EXPORT void lsm2dcomputereinitializationeqnrhs_(real_t* out, LSMHIP_BOXP2(rg), const real_t* phi,
                                                       LSMHIP_BOXP2(pg), const real_t* phi0, LSMHIP_BOXP2(p0g),
                                                       const real_t* pxp, const real_t* pyp, LSMHIP_BOXP2(gp),
                                                       const real_t* pxm, const real_t* pym, LSMHIP_BOXP2(gm),
                                                       LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy,
                                                       const int* use_phi0_for_sgn) {
  LSMHIP_REQUIRE_ARRAYS(out, phi, pxp, pyp, pxm, pym);
  if (*use_phi0_for_sgn == 1) LSMHIP_REQUIRE_ARRAYS(phi0);
  const char* who = "lsm2dcomputereinitializationeqnrhs_";
  Box fb = LSMHIP_BOX2(fb), rg = LSMHIP_BOX2(rg), pg = LSMHIP_BOX2(pg), p0g = LSMHIP_BOX2(p0g), gp = LSMHIP_BOX2(gp),
      gm = LSMHIP_BOX2(gm);
  if (fb.empty()) return;
  const int use0 = *use_phi0_for_sgn == 1;
  if (!inside(who, fb, {{rg, "reinit_rhs"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}})) return;
  if (!inside(who, fb, {{use0 ? p0g : pg, use0 ? "phi0" : "phi"}})) return;
  const real_t h = *dx > *dy ? *dx : *dy;
  hipLaunchKernelGGL(reinit_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(out, rg), view(phi, pg),
                     view(phi0, p0g), view(pxp, gp), view(pyp, gp), view(pxm, gm), view(pym, gm), h * h, use0);
  finish(who);
}
LSMHIP_EXPORT void lsm2dcomputeorthogonalizationeqnrhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp,
                                                        const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                        const real_t* pym, LSMHIP_BOXP2(gm), const real_t* psi,
                                                        LSMHIP_BOXP2(sg), const real_t* qxp, const real_t* qyp,
                                                        LSMHIP_BOXP2(qp), const real_t* qxm, const real_t* qym,
                                                        LSMHIP_BOXP2(qm), LSMHIP_BOXP2(fb), const real_t* dx,
                                                        const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym, psi, qxp, qyp, qxm, qym);
  const char* who = "lsm2dcomputeorthogonalizationeqnrhs_";
  Box fb = LSMHIP_BOX2(fb), rg = LSMHIP_BOX2(rg), gp = LSMHIP_BOX2(gp), gm = LSMHIP_BOX2(gm), sg = LSMHIP_BOX2(sg),
      qp = LSMHIP_BOX2(qp), qm = LSMHIP_BOX2(qm);
  if (fb.empty() || !inside(who, fb,
                            {{rg, "ortho_rhs"},
                             {gp, "grad_phi_plus"},
                             {gm, "grad_phi_minus"},
                             {sg, "psi"},
                             {qp, "grad_psi_plus"},
                             {qm, "grad_psi_minus"}}))
    return;
  const real_t h = *dx > *dy ? *dx : *dy;
  hipLaunchKernelGGL(ortho_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(pxp, gp),
                     view(pyp, gp), view(pxm, gm), view(pym, gm), view(psi, sg), view(qxp, qp), view(qyp, qp),
                     view(qxm, qm), view(qym, qm), h * h);
  finish(who);
}

// ===========================================================================
// Curvature, normals, field extension
// ===========================================================================
LSMHIP_EXPORT void lsm2dcomputemeancurvatureorder2_(real_t* kappa, LSMHIP_BOXP2(kg), const real_t* phi,
                                                    LSMHIP_BOXP2(pg), const real_t* px, const real_t* py,
                                                    const real_t* gmag, LSMHIP_BOXP2(gg), LSMHIP_BOXP2(fb),
                                                    const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, gmag);
  const char* who = "lsm2dcomputemeancurvatureorder2_";
  Box fb = LSMHIP_BOX2(fb), kg = LSMHIP_BOX2(kg), pg = LSMHIP_BOX2(pg), gg = LSMHIP_BOX2(gg);
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}, {gg, "grad_phi"}}) ||
      !check_box(pg, fb.widen(1, 1, 0), who, "phi") || !check_size(pg, who))
    return;
  hipLaunchKernelGGL(mean_curv_2d_kernel<2>, grid_for(fb), block2d(), 0, stream(), fb, view(kappa, kg),
                     view(phi, pg), view(px, gg), view(py, gg), view(gmag, gg), real_t(1.0) / *dx / *dx,
                     real_t(1.0) / *dy / *dy, real_t(0.25) / *dx / *dy);
  finish(who);
}
LSMHIP_EXPORT void lsm2dcomputemeancurvatureorder4_(real_t* kappa, LSMHIP_BOXP2(kg), const real_t* phi,
                                                    LSMHIP_BOXP2(pg), const real_t* px, const real_t* py,
                                                    const real_t* gmag, LSMHIP_BOXP2(gg), LSMHIP_BOXP2(fb),
                                                    const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, gmag);
  const char* who = "lsm2dcomputemeancurvatureorder4_";
  Box fb = LSMHIP_BOX2(fb), kg = LSMHIP_BOX2(kg), pg = LSMHIP_BOX2(pg), gg = LSMHIP_BOX2(gg);
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}}) || !check_box(gg, fb.widen(0, 2, 0), who, "grad_phi") ||
      !check_size(gg, who) || !cross_inside(who, fb, pg, 2, "phi"))
    return;
  const real_t c = real_t(0.0833333333333333333333);
  hipLaunchKernelGGL(mean_curv_2d_kernel<4>, grid_for(fb), block2d(), 0, stream(), fb, view(kappa, kg),
                     view(phi, pg), view(px, gg), view(py, gg), view(gmag, gg), c / *dx / *dx, c / *dy / *dy,
                     c / *dy);
  finish(who);
}
LSMHIP_EXPORT void lsm2dcomputemeancurvature_(real_t* kappa, LSMHIP_BOXP2(kg), const real_t* px, const real_t* py,
                                              LSMHIP_BOXP2(gg), const real_t* pxx, const real_t* pxy,
                                              const real_t* pyy, LSMHIP_BOXP2(hg), LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(kappa, px, py, pxx, pxy, pyy);
  const char* who = "lsm2dcomputemeancurvature_";
  Box fb = LSMHIP_BOX2(fb), kg = LSMHIP_BOX2(kg), gg = LSMHIP_BOX2(gg), hg = LSMHIP_BOX2(hg);
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}, {gg, "grad_phi"}, {hg, "hessian_phi"}})) return;
  hipLaunchKernelGGL(mean_curv_hessian_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(kappa, kg),
                     view(px, gg), view(py, gg), view(pxx, hg), view(pxy, hg), view(pyy, hg));
  finish(who);
}
LSMHIP_EXPORT void lsm2dcomputeunitnormal_(real_t* nx, real_t* ny, LSMHIP_BOXP2(ng), const real_t* px,
                                           const real_t* py, LSMHIP_BOXP2(gg), LSMHIP_BOXP2(fb)) {
  LSMHIP_REQUIRE_ARRAYS(nx, ny, px, py);
  const char* who = "lsm2dcomputeunitnormal_";
  Box fb = LSMHIP_BOX2(fb), ng = LSMHIP_BOX2(ng), gg = LSMHIP_BOX2(gg);
  if (fb.empty() || !inside(who, fb, {{ng, "normal"}, {gg, "grad_phi"}})) return;
  hipLaunchKernelGGL(normal_2d_kernel_n<0>, grid_for(fb), block2d(), 0, stream(), fb, view(nx, ng), view(ny, ng),
                     view(px, gg), view(py, gg), View<const real_t>{nullptr, 0, 0, 0, 0, 0}, real_t(0));
  finish(who);
}
LSMHIP_EXPORT void lsm2dcomputesignedunitnormal_(real_t* nx, real_t* ny, LSMHIP_BOXP2(ng), const real_t* px,
                                                 const real_t* py, LSMHIP_BOXP2(gg), const real_t* phi,
                                                 LSMHIP_BOXP2(pg), LSMHIP_BOXP2(fb), const real_t* dx,
                                                 const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(nx, ny, px, py, phi);
  const char* who = "lsm2dcomputesignedunitnormal_";
  Box fb = LSMHIP_BOX2(fb), ng = LSMHIP_BOX2(ng), gg = LSMHIP_BOX2(gg), pg = LSMHIP_BOX2(pg);
  if (fb.empty() || !inside(who, fb, {{ng, "normal"}, {gg, "grad_phi"}, {pg, "phi"}})) return;
  const real_t h = *dx > *dy ? *dx : *dy;
  hipLaunchKernelGGL(normal_2d_kernel_n<1>, grid_for(fb), block2d(), 0, stream(), fb, view(nx, ng), view(ny, ng),
                     view(px, gg), view(py, gg), view(phi, pg), h * h);
  finish(who);
}
LSMHIP_EXPORT void lsm2dcomputefieldextensioneqnrhs_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* S,
                                                     LSMHIP_BOXP2(sg), const real_t* phi, LSMHIP_BOXP2(pg),
                                                     const real_t* sx, const real_t* sy, LSMHIP_BOXP2(ug),
                                                     const real_t* nx, const real_t* ny, LSMHIP_BOXP2(ng),
                                                     LSMHIP_BOXP2(fb), const real_t* dx, const real_t* dy) {
  LSMHIP_REQUIRE_ARRAYS(rhs, S, phi, sx, sy, nx, ny);
  const char* who = "lsm2dcomputefieldextensioneqnrhs_";
  (void)S;  // part of the reference signature; the RHS reads only its upwind gradient
  Box fb = LSMHIP_BOX2(fb), rg = LSMHIP_BOX2(rg), pg = LSMHIP_BOX2(pg), ug = LSMHIP_BOX2(ug), ng = LSMHIP_BOX2(ng);
  if (fb.empty() || !inside(who, fb, {{rg, "rhs"}, {ug, "grad_S_upwind"}, {ng, "signed_normal"}}) ||
      !cross_inside(who, fb, pg, 1, "phi"))
    return;
  const real_t h = *dx > *dy ? *dx : *dy;
  hipLaunchKernelGGL(field_ext_2d_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(phi, pg),
                     view(sx, ug), view(sy, ug), view(nx, ng), view(ny, ng), real_t(3.0) * h);
  finish(who);
}

// ===========================================================================
// Reductions (src/toolbox/lsm_math_utils2d.h, src/geometry/lsm_geometry2d.h)
// ===========================================================================
LSMHIP_EXPORT void lsm2dmaxnormdiff_(real_t* out, const real_t* f1, LSMHIP_BOXP2(g1), const real_t* f2,
                                     LSMHIP_BOXP2(g2), LSMHIP_BOXP2(ib)) {
  LSMHIP_REQUIRE_ARRAYS(f1, f2);
  const char* who = "lsm2dmaxnormdiff_";
  Box ib = LSMHIP_BOX2(ib), g1 = LSMHIP_BOX2(g1), g2 = LSMHIP_BOX2(g2);
  *out = real_t(0);
  if (ib.empty() || !inside(who, ib, {{g1, "field1"}, {g2, "field2"}})) return;
  double m = 0;
  if (box_max2(who, ib, MaxNormDiff2{view(f1, g1), view(f2, g2)}, &m)) *out = real_t(m);
}
LSMHIP_EXPORT void lsm2dcomputestableadvectiondt_(real_t* dt, const real_t* vx, const real_t* vy, LSMHIP_BOXP2(vg),
                                                  LSMHIP_BOXP2(ib), const real_t* dx, const real_t* dy,
                                                  const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(vx, vy);
  const char* who = "lsm2dcomputestableadvectiondt_";
  Box ib = LSMHIP_BOX2(ib), vg = LSMHIP_BOX2(vg);
  double m = -1.0;
  if (!ib.empty()) {
    if (!inside(who, ib, {{vg, "vel"}})) return;
    if (!box_max2(who, ib, AdvectionCfl2{view(vx, vg), view(vy, vg), real_t(1.0) / *dx, real_t(1.0) / *dy}, &m))
      return;
  }
  *dt = cfl_dt(m, *cfl);
}
static void normal_dt2(const char* who, real_t* dt, const real_t* vn, const Box& vg, real_t vconst, const real_t* pxp,
                       const real_t* pyp, const Box& gp, const real_t* pxm, const real_t* pym, const Box& gm,
                       const Box& ib, real_t dx, real_t dy, real_t cfl) {
  double m = -1.0;
  if (!ib.empty()) {
    if (!inside(who, ib, {{gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}})) return;
    if (vn && !inside(who, ib, {{vg, "vel_n"}})) return;
    const real_t h = dx > dy ? dx : dy;
    NormalCfl2 op{vn ? view(vn, vg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0}, view(pxp, gp), view(pyp, gp),
                  view(pxm, gm), view(pym, gm), vconst, h * h, real_t(1.0) / dx, real_t(1.0) / dy};
    if (!box_max2(who, ib, op, &m)) return;
  }
  *dt = cfl_dt(m, cfl);
}
LSMHIP_EXPORT void lsm2dcomputestablenormalveldt_(real_t* dt, const real_t* vn, LSMHIP_BOXP2(vg), const real_t* pxp,
                                                  const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                  const real_t* pym, LSMHIP_BOXP2(gm), LSMHIP_BOXP2(ib),
                                                  const real_t* dx, const real_t* dy, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(vn, pxp, pyp, pxm, pym);
  normal_dt2("lsm2dcomputestablenormalveldt_", dt, vn, LSMHIP_BOX2(vg), real_t(0), pxp, pyp, LSMHIP_BOX2(gp), pxm, pym,
             LSMHIP_BOX2(gm), LSMHIP_BOX2(ib), *dx, *dy, *cfl);
}
LSMHIP_EXPORT void lsm2dcomputestableconstnormalveldt_(real_t* dt, const real_t* vn, const real_t* pxp,
                                                       const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                       const real_t* pym, LSMHIP_BOXP2(gm), LSMHIP_BOXP2(ib),
                                                       const real_t* dx, const real_t* dy, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym);
  const Box none{0, -1, 0, -1, 0, -1};
  normal_dt2("lsm2dcomputestableconstnormalveldt_", dt, nullptr, none, real_t(fabs(*vn)), pxp, pyp, LSMHIP_BOX2(gp),
             pxm, pym, LSMHIP_BOX2(gm), LSMHIP_BOX2(ib), *dx, *dy, *cfl);
}

static void region2(const char* who, real_t* out, const real_t* F, const Box& fg, const real_t* phi, const Box& pg,
                    const Box& ib, real_t dx, real_t dy, real_t eps, int sign) {
  if (ib.empty()) {
    *out = real_t(0);
    return;
  }
  if (!inside(who, ib, {{pg, "phi"}})) return;
  if (F && !inside(who, ib, {{fg, "F"}})) return;
  Region2 op{view(phi, pg), F ? view(F, fg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0}, eps, dx * dy, sign};
  double s = 0;
  if (reduce::box_sum(who, ib, op, &s)) *out = real_t(s);
}
static void surface2(const char* who, real_t* out, const real_t* F, const Box& fg, const real_t* phi, const Box& pg,
                     const real_t* px, const real_t* py, const Box& gg, const Box& ib, real_t dx, real_t dy,
                     real_t eps) {
  if (ib.empty()) {
    *out = real_t(0);
    return;
  }
  if (!inside(who, ib, {{pg, "phi"}, {gg, "grad_phi"}})) return;
  if (F && !inside(who, ib, {{fg, "F"}})) return;
  Surface2 op{view(phi, pg), F ? view(F, fg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0}, view(px, gg),
              view(py, gg), eps, real_t(1.0) / eps, dx * dy};
  double s = 0;
  if (reduce::box_sum(who, ib, op, &s)) *out = real_t(s);
}
LSMHIP_EXPORT void lsm2dvolumeintegralphilessthanzero_(real_t* out, const real_t* F, LSMHIP_BOXP2(fg),
                                                       const real_t* phi, LSMHIP_BOXP2(pg), LSMHIP_BOXP2(ib),
                                                       const real_t* dx, const real_t* dy, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi);
  region2("lsm2dvolumeintegralphilessthanzero_", out, F, LSMHIP_BOX2(fg), phi, LSMHIP_BOX2(pg), LSMHIP_BOX2(ib), *dx,
          *dy, *eps, -1);
}
LSMHIP_EXPORT void lsm2dvolumeintegralphigreaterthanzero_(real_t* out, const real_t* F, LSMHIP_BOXP2(fg),
                                                          const real_t* phi, LSMHIP_BOXP2(pg), LSMHIP_BOXP2(ib),
                                                          const real_t* dx, const real_t* dy, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi);
  region2("lsm2dvolumeintegralphigreaterthanzero_", out, F, LSMHIP_BOX2(fg), phi, LSMHIP_BOX2(pg), LSMHIP_BOX2(ib),
          *dx, *dy, *eps, 1);
}
LSMHIP_EXPORT void lsm2dsurfaceintegral_(real_t* out, const real_t* F, LSMHIP_BOXP2(fg), const real_t* phi,
                                         LSMHIP_BOXP2(pg), const real_t* px, const real_t* py, LSMHIP_BOXP2(gg),
                                         LSMHIP_BOXP2(ib), const real_t* dx, const real_t* dy, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi, px, py);
  surface2("lsm2dsurfaceintegral_", out, F, LSMHIP_BOX2(fg), phi, LSMHIP_BOX2(pg), px, py, LSMHIP_BOX2(gg),
           LSMHIP_BOX2(ib), *dx, *dy, *eps);
}
LSMHIP_EXPORT void lsm2darearegionphilessthanzero_(real_t* out, const real_t* phi, LSMHIP_BOXP2(pg), LSMHIP_BOXP2(ib),
                                                   const real_t* dx, const real_t* dy, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  const Box none{0, -1, 0, -1, 0, -1};
  region2("lsm2darearegionphilessthanzero_", out, nullptr, none, phi, LSMHIP_BOX2(pg), LSMHIP_BOX2(ib), *dx, *dy,
          *eps, -1);
}
LSMHIP_EXPORT void lsm2darearegionphigreaterthanzero_(real_t* out, const real_t* phi, LSMHIP_BOXP2(pg),
                                                      LSMHIP_BOXP2(ib), const real_t* dx, const real_t* dy,
                                                      const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  const Box none{0, -1, 0, -1, 0, -1};
  region2("lsm2darearegionphigreaterthanzero_", out, nullptr, none, phi, LSMHIP_BOX2(pg), LSMHIP_BOX2(ib), *dx, *dy,
          *eps, 1);
}
LSMHIP_EXPORT void lsm2dperimeterzerolevelset_(real_t* out, const real_t* phi, LSMHIP_BOXP2(pg), const real_t* px,
                                               const real_t* py, LSMHIP_BOXP2(gg), LSMHIP_BOXP2(ib), const real_t* dx,
                                               const real_t* dy, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi, px, py);
  const Box none{0, -1, 0, -1, 0, -1};
  surface2("lsm2dperimeterzerolevelset_", out, nullptr, none, phi, LSMHIP_BOX2(pg), px, py, LSMHIP_BOX2(gg),
           LSMHIP_BOX2(ib), *dx, *dy, *eps);
}

// ===========================================================================
// Ghost fills (src/boundary_conditions/lsm_boundary_conditions2d.h)
// ===========================================================================
LSMHIP_EXPORT void lsm2dlinearextrapolation_(real_t* phi, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face2("lsm2dlinearextrapolation_", phi, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *idx, kLinear);
}
LSMHIP_EXPORT void lsm2dsignedlinearextrapolation_(real_t* phi, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face2("lsm2dsignedlinearextrapolation_", phi, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *idx, kSigned);
}
LSMHIP_EXPORT void lsm2dcopyextrapolation_(real_t* phi, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face2("lsm2dcopyextrapolation_", phi, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *idx, kCopy);
}
// The four homogeneous Neumann fills are copy extrapolation (lsm_boundary_conditions2d.f:360-510).
LSMHIP_EXPORT void lsm2dhomogeneousneumanneno1_(real_t* phi, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face2("lsm2dhomogeneousneumanneno1_", phi, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm2dhomogeneousneumanneno2_(real_t* phi, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face2("lsm2dhomogeneousneumanneno2_", phi, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm2dhomogeneousneumanneno3_(real_t* phi, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face2("lsm2dhomogeneousneumanneno3_", phi, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *idx, kCopy);
}
LSMHIP_EXPORT void lsm2dhomogeneousneumannweno5_(real_t* phi, LSMHIP_BOXP2(gb), LSMHIP_BOXP2(fb), const int* idx) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  face2("lsm2dhomogeneousneumannweno5_", phi, LSMHIP_BOX2(gb), LSMHIP_BOX2(fb), *idx, kCopy);
}
