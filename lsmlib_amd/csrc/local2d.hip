// local2d.hip -- the 2D narrow-band ("local") kernels: HJ-ENO1/2/3 and
// HJ-WENO5 derivatives, upwinded HJ-ENO2, central gradients, Laplacian,
// |grad phi| and its band average, RHS terms, TVD Runge-Kutta stages,
// reinitialization and orthogonalization RHS, mean curvature (orders 2 and 4),
// unit normals, CFL steps, max-norm difference, mask/copy helpers and the RHS
// cut-off.
// References: src/toolbox/lsm_spatial_derivatives2d_local.f.in,
// src/toolbox/lsm_level_set_evolution2d_local.f.in, src/toolbox/lsm_tvd_runge_kutta2d_local.f,
// src/toolbox/lsm_math_utils2d_local.f.in, src/reinitialization/lsm_reinitialization2d_local.f.in,
// src/geometry/lsm_curvature2d_local.f.in, src/geometry/lsm_geometry2d_local.f.in,
// src/toolbox/lsm_localization2d.f:621-875.
//
// A 2D array is a 3D array of one plane (k = 0) and a 2D index list has no
// index_z, so the one-thread-per-listed-point launches and box checks of the
// 3D narrow-band kernels carry over.  The derivative routines keep the
// reference's use of the caller's D1/D2/D3 scratch (differences written at the
// listed points of levels 0..L, 1e10 where the mark exceeds mark_D*, then read
// at level 0).  FP64 results are bitwise the reference's; the band average of
// |grad phi| is an FP64 compensated sum.
#include "local.hpp"
#include "stencil.hpp"

using namespace lsmhip;

#define LSMHIP_BOXP2(n) const int *ilo_##n, const int *ihi_##n, const int *jlo_##n, const int *jhi_##n
#define LSMHIP_BOX2(n) ::lsmhip::Box{*ilo_##n, *ihi_##n, *jlo_##n, *jhi_##n, 0, 0}

namespace {

// Level-0 plus/minus derivatives along axis (di, dj) from the difference
// arrays (lsm_spatial_derivatives2d_local.f.in:133-1154).  ORDER 5 is
// HJ-WENO5 with the routine's literal tiny number 1e-35 (:892).
template <int ORDER>
struct Deriv2Op {
  CV<real_t> p, m;
  CV<const real_t> d1, d2, d3;
  int di, dj;
  real_t inv;
  __device__ bool operator()(int i, int j, int k) const {
    if (!p.has(i, j, k) || !m.has(i, j, k)) return false;
    auto D1 = [&](int s) -> real_t { return d1(i + s * di, j + s * dj, k); };
    auto D2 = [&](int s) -> real_t { return d2(i + s * di, j + s * dj, k); };
    auto D3 = [&](int s) -> real_t { return d3(i + s * di, j + s * dj, k); };
    auto in = [&](const CV<const real_t>& a, int s0, int s1) {
      return a.has(i + s0 * di, j + s0 * dj, k) && a.has(i + s1 * di, j + s1 * dj, k);
    };
    const real_t h = real_t(0.5);
    if (ORDER == 1) {
      if (!in(d1, 0, 1)) return false;
      p(i, j, k) = D1(1) * inv;
      m(i, j, k) = D1(0) * inv;
    } else if (ORDER == 2) {
      if (!in(d1, 0, 1) || !in(d2, -1, 1)) return false;
      p(i, j, k) = (fabs(D2(0)) < fabs(D2(1)) ? D1(1) - h * D2(0) : D1(1) - h * D2(1)) * inv;
      m(i, j, k) = (fabs(D2(-1)) < fabs(D2(0)) ? D1(0) + h * D2(-1) : D1(0) + h * D2(0)) * inv;
    } else if (ORDER == 3) {
      if (!in(d1, 0, 1) || !in(d2, -1, 1) || !in(d3, -1, 2)) return false;
      const real_t third = real_t(1.0 / 3.0), sixth = real_t(1.0 / 6.0);
      real_t r = D1(1);
      if (fabs(D2(0)) < fabs(D2(1))) {
        r = r - h * D2(0);
        r = fabs(D3(0)) < fabs(D3(1)) ? r - sixth * D3(0) : r - sixth * D3(1);
      } else {
        r = r - h * D2(1);
        r = fabs(D3(1)) < fabs(D3(2)) ? r + third * D3(1) : r + third * D3(2);
      }
      p(i, j, k) = r * inv;
      r = D1(0);
      if (fabs(D2(-1)) < fabs(D2(0))) {
        r = r + h * D2(-1);
        r = fabs(D3(-1)) < fabs(D3(0)) ? r + third * D3(-1) : r + third * D3(0);
      } else {
        r = r + h * D2(0);
        r = fabs(D3(0)) < fabs(D3(1)) ? r - sixth * D3(0) : r - sixth * D3(1);
      }
      m(i, j, k) = r * inv;
    } else {
      if (!in(d1, -2, 3)) return false;
      const real_t t = real_t(1.e-35);
      p(i, j, k) = stencil::weno5(D1(3) * inv, D1(2) * inv, D1(1) * inv, D1(0) * inv, D1(-1) * inv, t);
      m(i, j, k) = stencil::weno5(D1(-2) * inv, D1(-1) * inv, D1(0) * inv, D1(1) * inv, D1(2) * inv, t);
    }
    return true;
  }
};

// lsm2dUpwindHJENO2LOCAL, :1160-1386: zero where |v| < zero_tol, the minus
// (backward) ENO2 form where v > 0, the plus form otherwise.
struct Upwind2Op {
  CV<real_t> o;
  CV<const real_t> v, d1, d2;
  int di, dj;
  real_t inv;
  __device__ bool operator()(int i, int j, int k) const {
    if (!o.has(i, j, k) || !v.has(i, j, k) || !d1.has(i, j, k) || !d1.has(i + di, j + dj, k) ||
        !d2.has(i - di, j - dj, k) || !d2.has(i + di, j + dj, k))
      return false;
    auto D1 = [&](int s) -> real_t { return d1(i + s * di, j + s * dj, k); };
    auto D2 = [&](int s) -> real_t { return d2(i + s * di, j + s * dj, k); };
    const real_t vel = v(i, j, k), h = real_t(0.5);
    if (fabs(vel) < kZeroTol)
      o(i, j, k) = real_t(0);
    else if (vel > real_t(0))
      o(i, j, k) = (fabs(D2(-1)) < fabs(D2(0)) ? D1(0) + h * D2(-1) : D1(0) + h * D2(0)) * inv;
    else
      o(i, j, k) = (fabs(D2(0)) < fabs(D2(1)) ? D1(1) - h * D2(0) : D1(1) - h * D2(1)) * inv;
    return true;
  }
};

struct Four {
  CV<const real_t> xp, yp, xm, ym;
  __device__ bool has(int i, int j, int k) const { return xp.has(i, j, k) && xm.has(i, j, k); }
};
Four four(const real_t* xp, const real_t* yp, const Box& gp, const real_t* xm, const real_t* ym, const Box& gm) {
  return Four{cv(xp, gp), cv(yp, gp), cv(xm, gm), cv(ym, gm)};
}

struct Central2Op {
  CV<real_t> ox, oy;
  CV<const real_t> F;
  real_t fx, fy;
  __device__ bool operator()(int i, int j, int k) const {
    if (!ox.has(i, j, k) || !F.star2(i, j, k, 1)) return false;
    ox(i, j, k) = (F(i + 1, j, k) - F(i - 1, j, k)) * fx;
    oy(i, j, k) = (F(i, j + 1, k) - F(i, j - 1, k)) * fy;
    return true;
  }
};

struct Central4Op {
  CV<real_t> ox, oy;
  CV<const real_t> F;
  real_t fx, fy;
  __device__ bool operator()(int i, int j, int k) const {
    if (!ox.has(i, j, k) || !F.star2(i, j, k, 2)) return false;
    const real_t e = real_t(8);
    ox(i, j, k) = (-F(i + 2, j, k) + e * F(i + 1, j, k) + F(i - 2, j, k) - e * F(i - 1, j, k)) * fx;
    oy(i, j, k) = (-F(i, j + 2, k) + e * F(i, j + 1, k) + F(i, j - 2, k) - e * F(i, j - 1, k)) * fy;
    return true;
  }
};

struct LaplacianOp {
  CV<real_t> out;
  CV<const real_t> F;
  real_t ax, ay;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.has(i, j, k) || !F.star2(i, j, k, 1)) return false;
    const real_t c = F(i, j, k);
    out(i, j, k) = ax * (F(i + 1, j, k) - real_t(2) * c + F(i - 1, j, k)) +
                   ay * (F(i, j + 1, k) - real_t(2) * c + F(i, j - 1, k));
    return true;
  }
};

// :1689-1763 -- both cut-offs are zero_tol
struct GradMagOp {
  CV<const real_t> x, y;
  CV<real_t> mag;
  __device__ bool operator()(int i, int j, int k) const {
    if (!mag.has(i, j, k)) return false;
    real_t sx = x(i, j, k) * x(i, j, k), sy = y(i, j, k) * y(i, j, k);
    if (sx < kZeroTol) sx = real_t(0);
    if (sy < kZeroTol) sy = real_t(0);
    const real_t t = sqrt(sx + sy);
    mag(i, j, k) = t < kZeroTol ? real_t(0) : t;
    return true;
  }
};

struct GradNormOp {
  CV<const real_t> F;
  real_t fx, fy;
  __device__ bool operator()(int i, int j, int k, double& v) const {
    if (!F.star2(i, j, k, 1)) return false;
    const real_t x = (F(i + 1, j, k) - F(i - 1, j, k)) * fx;
    const real_t y = (F(i, j + 1, k) - F(i, j - 1, k)) * fy;
    v = (double)sqrt(x * x + y * y);
    return true;
  }
};

struct ZeroOp {
  CV<real_t> r;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k)) return false;
    r(i, j, k) = real_t(0);
    return true;
  }
};

struct AdvectOp {
  CV<real_t> r;
  CV<const real_t> px, py, vx, vy;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !px.has(i, j, k) || !vx.has(i, j, k)) return false;
    r(i, j, k) = r(i, j, k) - (vx(i, j, k) * px(i, j, k) + vy(i, j, k) * py(i, j, k));
    return true;
  }
};

// Godunov |grad phi|^2 (lsm_level_set_evolution2d_local.f.in:166-176)
__device__ __forceinline__ real_t godunov2(bool vpos, real_t xp, real_t yp, real_t xm, real_t ym) {
  const real_t z = real_t(0);
  if (vpos) return maxr(sq(maxr(xm, z)), sq(minr(xp, z))) + maxr(sq(maxr(ym, z)), sq(minr(yp, z)));
  return maxr(sq(minr(xm, z)), sq(maxr(xp, z))) + maxr(sq(minr(ym, z)), sq(maxr(yp, z)));
}

// vn.v.p null: constant speed vc (:126-314)
struct NormalOp {
  CV<real_t> r;
  Four g;
  CV<const real_t> vn;
  real_t vc;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !g.has(i, j, k) || (vn.v.p && !vn.has(i, j, k))) return false;
    const real_t v = vn.v.p ? vn(i, j, k) : vc;
    const real_t s = godunov2(v > real_t(0), g.xp(i, j, k), g.yp(i, j, k), g.xm(i, j, k), g.ym(i, j, k));
    if (fabs(v) >= kZeroTol) r(i, j, k) = r(i, j, k) - v * sqrt(s);
    return true;
  }
};

// :320-402
struct ConstCurvOp {
  CV<real_t> r;
  CV<const real_t> px, py, xx, xy, yy;
  real_t b;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !px.has(i, j, k) || !xx.has(i, j, k)) return false;
    const real_t x = px(i, j, k), y = py(i, j, k);
    const real_t g = x * x + y * y;
    real_t curv = real_t(0);
    if (!(g < kZeroTol)) {
      curv = xx(i, j, k) * y * y + yy(i, j, k) * x * x - real_t(2) * xy(i, j, k) * x * y;
      curv = curv / g;
    }
    r(i, j, k) = r(i, j, k) + b * curv;
    return true;
  }
};

struct PrecompCurvOp {
  CV<real_t> r;
  CV<const real_t> kap, gm;
  real_t b;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !kap.has(i, j, k) || !gm.has(i, j, k)) return false;
    r(i, j, k) = r(i, j, k) + b * kap(i, j, k) * gm(i, j, k);
    return true;
  }
};

// :474-615 -- the advection part everywhere on the band, the normal part where
// |V| >= zero_tol and the picked |grad phi|^2 >= zero_tol
struct ExtNormalOp {
  CV<real_t> r;
  Four g;
  CV<const real_t> vn, vx, vy;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !g.has(i, j, k) || !vn.has(i, j, k)) return false;
    const real_t n = vn(i, j, k), u = vx(i, j, k), v = vy(i, j, k);
    const real_t px = stencil::ext_normal_pick(u, n, g.xp(i, j, k), g.xm(i, j, k));
    const real_t py = stencil::ext_normal_pick(v, n, g.yp(i, j, k), g.ym(i, j, k));
    const real_t s = px * px + py * py;
    if (fabs(n) >= kZeroTol && s >= kZeroTol)
      r(i, j, k) = r(i, j, k) - u * px - v * py - n * sqrt(s);
    else
      r(i, j, k) = r(i, j, k) - u * px - v * py;
    return true;
  }
};

// lsm_tvd_runge_kutta2d_local.f:88, :274, :458, :554 (kinds as in 3D)
template <int KIND>
struct RkOp {
  CV<real_t> out;
  CV<const real_t> us, uc, r;
  real_t dt;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.has(i, j, k) || !uc.has(i, j, k) || !r.has(i, j, k) || (KIND != 1 && !us.has(i, j, k))) return false;
    const real_t q = r(i, j, k);
    if (KIND == 1) out(i, j, k) = uc(i, j, k) + dt * q;
    if (KIND == 2) out(i, j, k) = real_t(0.5) * (uc(i, j, k) + us(i, j, k) + dt * q);
    if (KIND == 32) out(i, j, k) = real_t(0.75) * uc(i, j, k) + real_t(0.25) * (us(i, j, k) + dt * q);
    if (KIND == 33)
      out(i, j, k) = real_t(1.0 / 3.0) * uc(i, j, k) + real_t(2.0 / 3.0) * (us(i, j, k) + dt * q);
    return true;
  }
};

// lsm_reinitialization2d_local.f.in:49-251 -- zero where |sign source| < zero_tol
// (the local routine's >=, unlike the 2D whole-grid >)
struct ReinitOp {
  CV<real_t> out;
  CV<const real_t> phi, phi0;
  Four g;
  real_t dx_sq;
  int use0;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.has(i, j, k) || !g.has(i, j, k)) return false;
    if (use0 ? !phi0.has(i, j, k) : !phi.has(i, j, k)) return false;
    const real_t s = use0 ? phi0(i, j, k) : phi(i, j, k);
    const real_t gp[2] = {g.xp(i, j, k), g.yp(i, j, k)}, gm[2] = {g.xm(i, j, k), g.ym(i, j, k)};
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
    if (fabs(s) >= kZeroTol) {
      const real_t n2 = star[0] * star[0] + star[1] * star[1];
      const real_t sgn = use0 ? s / sqrt(s * s + dx_sq) : s / sqrt(s * s + n2 * dx_sq);
      r = sgn * (real_t(1) - sqrt(n2));
    }
    out(i, j, k) = r;
    return true;
  }
};

// :284-432
struct OrthoOp {
  CV<real_t> out;
  Four g;
  CV<const real_t> psi;
  Four q;
  real_t dx_sq;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.has(i, j, k) || !g.has(i, j, k) || !psi.has(i, j, k) || !q.has(i, j, k)) return false;
    const real_t g1 = real_t(0.5) * (q.xp(i, j, k) + q.xm(i, j, k));
    const real_t g2 = real_t(0.5) * (q.yp(i, j, k) + q.ym(i, j, k));
    const real_t ng = sqrt(g1 * g1 + g2 * g2);
    const real_t s = psi(i, j, k);
    real_t r = real_t(0);
    if (fabs(s) > kZeroTol && ng > kZeroTol) {
      const real_t sgn = s / sqrt(s * s + dx_sq);
      const real_t a1 = sgn * g1, a2 = sgn * g2, z = real_t(0);
      r = real_t(-1) / ng *
          (maxr(a1, z) * g.xm(i, j, k) + minr(a1, z) * g.xp(i, j, k) + maxr(a2, z) * g.ym(i, j, k) +
           minr(a2, z) * g.yp(i, j, k));
    }
    out(i, j, k) = r;
    return true;
  }
};

// lsm_curvature2d_local.f.in:15-108 (ORDER 2) and :299-402 (ORDER 4, phi_xy = d/dy phi_x)
template <int ORDER>
struct MeanCurvOp {
  CV<real_t> kap;
  CV<const real_t> F, px, py, gm;
  real_t fxx, fyy, fxy;
  __device__ bool operator()(int i, int j, int k) const {
    const int w = ORDER == 2 ? 1 : 2;
    if (!kap.has(i, j, k) || !px.has(i, j, k) || !F.star2(i, j, k, w)) return false;
    if (ORDER == 4 && (!px.has(i, j - 2, k) || !px.has(i, j + 2, k))) return false;
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
    const real_t g = gm(i, j, k), den = g * g * g;
    real_t out = real_t(0);
    if (!(den < kZeroTol)) {
      const real_t x = px(i, j, k), y = py(i, j, k);
      out = (xx * y * y + yy * x * x - real_t(2) * xy * x * y) / den;
    }
    kap(i, j, k) = out;
    return true;
  }
};

// lsm_geometry2d_local.f.in:34-113 (SIGNED = 0) and :140-248 (SIGNED = 1)
template <int SIGNED>
struct NormalVecOp {
  CV<real_t> nx, ny;
  CV<const real_t> px, py, phi;
  real_t dx_sq;
  __device__ bool operator()(int i, int j, int k) const {
    if (!nx.has(i, j, k) || !px.has(i, j, k) || (SIGNED && !phi.has(i, j, k))) return false;
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
    return true;
  }
};

struct ImposeMaskOp {
  CV<real_t> dst;
  CV<const real_t> src, msk;
  __device__ bool operator()(int i, int j, int k) const {
    if (!dst.has(i, j, k) || !src.has(i, j, k) || !msk.has(i, j, k)) return false;
    dst(i, j, k) = maxr(msk(i, j, k), src(i, j, k));
    return true;
  }
};

struct CopyOp {
  CV<real_t> dst;
  CV<const real_t> src;
  __device__ bool operator()(int i, int j, int k) const {
    if (!dst.has(i, j, k) || !src.has(i, j, k)) return false;
    dst(i, j, k) = src(i, j, k);
    return true;
  }
};

// lsm2dMultiplyCutOffLSERHSLOCAL, lsm_localization2d.f:621-688
struct CutOffOp {
  CV<const real_t> phi;
  CV<real_t> rhs;
  real_t beta, gamma, c1, c2;
  __device__ bool operator()(int i, int j, int k) const {
    if (!phi.has(i, j, k) || !rhs.has(i, j, k)) return false;
    const real_t a = fabs(phi(i, j, k));
    real_t cut;
    if (a <= beta) {
      cut = real_t(1);
    } else if (a <= gamma) {
      const real_t t = a - gamma;
      cut = (t * t * (real_t(2) * a + c1)) / c2;
    } else {
      cut = real_t(0);
    }
    rhs(i, j, k) = cut * rhs(i, j, k);
    return true;
  }
};

struct MaxNormDiffOp {
  CV<const real_t> a, b;
  __device__ bool operator()(int i, int j, int k, double& v) const {
    if (!a.has(i, j, k) || !b.has(i, j, k)) return false;
    v = (double)fabs(a(i, j, k) - b(i, j, k));
    return true;
  }
};

struct AdvDtOp {
  CV<const real_t> u, v;
  real_t ix, iy;
  __device__ bool operator()(int i, int j, int k, double& r) const {
    if (!u.has(i, j, k)) return false;
    r = (double)(fabs(u(i, j, k)) * ix + fabs(v(i, j, k)) * iy);
    return true;
  }
};

// lsm_math_utils2d_local.f.in:246-462 (vn.v.p null: constant |speed| vc)
struct NormalDtOp {
  Four g;
  CV<const real_t> vn;
  real_t vc, mdx2, ix, iy;
  __device__ bool operator()(int i, int j, int k, double& r) const {
    if (!g.has(i, j, k) || (vn.v.p && !vn.has(i, j, k))) return false;
    const real_t av = vn.v.p ? fabs(vn(i, j, k)) : vc;
    const real_t x = maxr(fabs(g.xp(i, j, k)), fabs(g.xm(i, j, k)));
    const real_t y = maxr(fabs(g.yp(i, j, k)), fabs(g.ym(i, j, k)));
    r = (double)(av / sqrt(x * x + y * y + mdx2) * (x * ix + y * iy));
    return true;
  }
};

IndexList list2(const int* ix, const int* iy, int lo, int hi, int base) { return IndexList{ix, iy, nullptr, lo, hi, base}; }

}  // namespace

// ---------------------------------------------------------------------------
// Derivatives (src/toolbox/lsm_spatial_derivatives2d_local.h)
// ---------------------------------------------------------------------------
// hiD[o] is the last index of the levels over which the order-o differences
// are taken (levels 0..ORDER+1-o for ENO, 0..3 for WENO5's first differences).
template <int ORDER>
static void local_derivs2(const char* who, real_t* pxp, real_t* pyp, const Box& gp, real_t* pxm, real_t* pym,
                          const Box& gm, const real_t* phi, const Box& pg, real_t* D1, const Box& d1b, real_t* D2,
                          const Box& d2b, real_t* D3, const Box& d3b, real_t dx, real_t dy, const int* ix,
                          const int* iy, int lo0, int hi0, const int hiD[3], const unsigned char* nb, const Box& nbb,
                          const unsigned char* mark_fb, const unsigned char* const mark_d[3]) {
  const real_t inv[2] = {real_t(1.0) / dx, real_t(1.0) / dy};
  real_t* P[2] = {pxp, pyp};
  real_t* M[2] = {pxm, pym};
  const Mask nbm = make_mask(nb, nbb, nullptr, 0);
  const Mask fbm = make_mask(nb, nbb, mark_fb, 1);
  const int nd = ORDER == 5 ? 1 : ORDER;
  real_t* D[3] = {D1, D2, D3};
  const Box db[3] = {d1b, d2b, d3b};
  for (int d = 0; d < 2; d++) {
    for (int o = 0; o < nd; o++)
      dn_local(who, D[o], db[o], o == 0 ? phi : D[o - 1], o == 0 ? pg : db[o - 1], o + 1, d,
               list2(ix, iy, lo0, hiD[o], lo0), nbm, (signed char)*mark_d[o]);
    Deriv2Op<ORDER> op{cv(P[d], gp), cv(M[d], gm), cv((const real_t*)D1, d1b),
                       cv((const real_t*)(nd >= 2 ? D2 : D1), nd >= 2 ? d2b : d1b),
                       cv((const real_t*)(nd >= 3 ? D3 : D1), nd >= 3 ? d3b : d1b), d == 0, d == 1, inv[d]};
    launch_points(who, list2(ix, iy, lo0, hi0, lo0), fbm, op);
  }
  finish_local(who);
}

#define LPM2_ARGS                                                                                            \
  real_t *pxp, real_t *pyp, LSMHIP_BOXP2(gp), real_t *pxm, real_t *pym, LSMHIP_BOXP2(gm), const real_t *phi, \
      LSMHIP_BOXP2(pg)

LSMHIP_EXPORT void lsm2dhjeno1local_(LPM2_ARGS, real_t* D1, LSMHIP_BOXP2(d1), const real_t* dx, const real_t* dy,
                                     const int* ix, const int* iy, const int* nlo0, const int* nhi0, const int* nlo1,
                                     const int* nhi1, const unsigned char* nb, LSMHIP_BOXP2(nbb),
                                     const unsigned char* mark_fb, const unsigned char* mark_d1) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1, ix, iy, nb);
  (void)nlo1;
  const int hiD[3] = {*nhi1, *nhi1, *nhi1};
  const unsigned char* marks[3] = {mark_d1, mark_d1, mark_d1};
  const Box d1 = LSMHIP_BOX2(d1);
  local_derivs2<1>("lsm2dhjeno1local_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
                   D1, d1, nullptr, d1, nullptr, d1, *dx, *dy, ix, iy, *nlo0, *nhi0, hiD, nb, LSMHIP_BOX2(nbb),
                   mark_fb, marks);
}

LSMHIP_EXPORT void lsm2dhjeno2local_(LPM2_ARGS, real_t* D1, LSMHIP_BOXP2(d1), real_t* D2, LSMHIP_BOXP2(d2),
                                     const real_t* dx, const real_t* dy, const int* ix, const int* iy,
                                     const int* nlo0, const int* nhi0, const int* nlo1, const int* nhi1,
                                     const int* nlo2, const int* nhi2, const unsigned char* nb, LSMHIP_BOXP2(nbb),
                                     const unsigned char* mark_fb, const unsigned char* mark_d1,
                                     const unsigned char* mark_d2) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1, D2, ix, iy, nb);
  (void)nlo1;
  (void)nlo2;
  const int hiD[3] = {*nhi2, *nhi1, *nhi1};
  const unsigned char* marks[3] = {mark_d1, mark_d2, mark_d2};
  const Box d1 = LSMHIP_BOX2(d1);
  local_derivs2<2>("lsm2dhjeno2local_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
                   D1, d1, D2, LSMHIP_BOX2(d2), nullptr, d1, *dx, *dy, ix, iy, *nlo0, *nhi0, hiD, nb,
                   LSMHIP_BOX2(nbb), mark_fb, marks);
}

LSMHIP_EXPORT void lsm2dhjeno3local_(LPM2_ARGS, real_t* D1, LSMHIP_BOXP2(d1), real_t* D2, LSMHIP_BOXP2(d2),
                                     real_t* D3, LSMHIP_BOXP2(d3), const real_t* dx, const real_t* dy, const int* ix,
                                     const int* iy, const int* nlo0, const int* nhi0, const int* nlo1,
                                     const int* nhi1, const int* nlo2, const int* nhi2, const int* nlo3,
                                     const int* nhi3, const unsigned char* nb, LSMHIP_BOXP2(nbb),
                                     const unsigned char* mark_fb, const unsigned char* mark_d1,
                                     const unsigned char* mark_d2, const unsigned char* mark_d3) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1, D2, D3, ix, iy, nb);
  (void)nlo1;
  (void)nlo2;
  (void)nlo3;
  const int hiD[3] = {*nhi3, *nhi2, *nhi1};
  const unsigned char* marks[3] = {mark_d1, mark_d2, mark_d3};
  local_derivs2<3>("lsm2dhjeno3local_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
                   D1, LSMHIP_BOX2(d1), D2, LSMHIP_BOX2(d2), D3, LSMHIP_BOX2(d3), *dx, *dy, ix, iy, *nlo0, *nhi0,
                   hiD, nb, LSMHIP_BOX2(nbb), mark_fb, marks);
}

LSMHIP_EXPORT void lsm2dhjweno5local_(LPM2_ARGS, real_t* D1, LSMHIP_BOXP2(d1), const real_t* dx, const real_t* dy,
                                      const int* ix, const int* iy, const int* nlo0, const int* nhi0,
                                      const int* nlo1, const int* nhi1, const int* nlo2, const int* nhi2,
                                      const int* nlo3, const int* nhi3, const unsigned char* nb, LSMHIP_BOXP2(nbb),
                                      const unsigned char* mark_fb, const unsigned char* mark_d1) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, phi, D1, ix, iy, nb);
  (void)nlo1;
  (void)nhi1;
  (void)nlo2;
  (void)nhi2;
  (void)nlo3;
  const int hiD[3] = {*nhi3, *nhi3, *nhi3};
  const unsigned char* marks[3] = {mark_d1, mark_d1, mark_d1};
  const Box d1 = LSMHIP_BOX2(d1);
  local_derivs2<5>("lsm2dhjweno5local_", pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), phi, LSMHIP_BOX2(pg),
                   D1, d1, nullptr, d1, nullptr, d1, *dx, *dy, ix, iy, *nlo0, *nhi0, hiD, nb, LSMHIP_BOX2(nbb),
                   mark_fb, marks);
}

LSMHIP_EXPORT void lsm2dupwindhjeno2local_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi,
                                           LSMHIP_BOXP2(pg), const real_t* vx, const real_t* vy, LSMHIP_BOXP2(vg),
                                           real_t* D1, LSMHIP_BOXP2(d1), real_t* D2, LSMHIP_BOXP2(d2),
                                           const real_t* dx, const real_t* dy, const int* ix, const int* iy,
                                           const int* nlo0, const int* nhi0, const int* nlo1, const int* nhi1,
                                           const int* nlo2, const int* nhi2, const unsigned char* nb,
                                           LSMHIP_BOXP2(nbb), const unsigned char* mark_fb,
                                           const unsigned char* mark_d1, const unsigned char* mark_d2) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi, vx, vy, D1, D2, ix, iy, nb);
  const char* who = "lsm2dupwindhjeno2local_";
  (void)nlo1;
  (void)nlo2;
  const Box nbb = LSMHIP_BOX2(nbb), d1b = LSMHIP_BOX2(d1), d2b = LSMHIP_BOX2(d2), pg = LSMHIP_BOX2(pg);
  const Box g = LSMHIP_BOX2(g), vg = LSMHIP_BOX2(vg);
  const Mask nbm = make_mask(nb, nbb, nullptr, 0);
  const real_t inv[2] = {real_t(1.0) / *dx, real_t(1.0) / *dy};
  real_t* O[2] = {px, py};
  const real_t* V[2] = {vx, vy};
  for (int d = 0; d < 2; d++) {
    dn_local(who, D1, d1b, phi, pg, 1, d, list2(ix, iy, *nlo0, *nhi2, *nlo0), nbm, (signed char)*mark_d1);
    dn_local(who, D2, d2b, D1, d1b, 2, d, list2(ix, iy, *nlo0, *nhi1, *nlo0), nbm, (signed char)*mark_d2);
    launch_points(who, list2(ix, iy, *nlo0, *nhi0, *nlo0), make_mask(nb, nbb, mark_fb, 1),
                  Upwind2Op{cv(O[d], g), cv(V[d], vg), cv((const real_t*)D1, d1b), cv((const real_t*)D2, d2b),
                            d == 0, d == 1, inv[d]});
  }
  finish_local(who);
}

#define LOCAL2_TAIL                                                                                          \
  const int *ix, const int *iy, const int *nlo, const int *nhi, const unsigned char *nb, LSMHIP_BOXP2(nbb), \
      const unsigned char *mark_fb
#define LIST2 list2(ix, iy, *nlo, *nhi, *nlo)
#define FBMASK2 make_mask(nb, LSMHIP_BOX2(nbb), mark_fb, 1)

LSMHIP_EXPORT void lsm2dcentralgradorder2local_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi,
                                                LSMHIP_BOXP2(pg), const real_t* dx, const real_t* dy, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi, ix, iy, nb);
  const char* who = "lsm2dcentralgradorder2local_";
  const Box g = LSMHIP_BOX2(g);
  launch_points(who, LIST2, FBMASK2,
                Central2Op{cv(px, g), cv(py, g), cv(phi, LSMHIP_BOX2(pg)), real_t(0.5) / *dx, real_t(0.5) / *dy});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2dcentralgradorder4local_(real_t* px, real_t* py, LSMHIP_BOXP2(g), const real_t* phi,
                                                LSMHIP_BOXP2(pg), const real_t* dx, const real_t* dy, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(px, py, phi, ix, iy, nb);
  const char* who = "lsm2dcentralgradorder4local_";
  const Box g = LSMHIP_BOX2(g);
  const real_t c = real_t(0.0833333333333333333332);
  launch_points(who, LIST2, FBMASK2, Central4Op{cv(px, g), cv(py, g), cv(phi, LSMHIP_BOX2(pg)), c / *dx, c / *dy});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2dlaplacianorder2local_(real_t* lap, LSMHIP_BOXP2(lg), const real_t* phi, LSMHIP_BOXP2(pg),
                                              const real_t* dx, const real_t* dy, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(lap, phi, ix, iy, nb);
  const char* who = "lsm2dlaplacianorder2local_";
  launch_points(who, LIST2, FBMASK2,
                LaplacianOp{cv(lap, LSMHIP_BOX2(lg)), cv(phi, LSMHIP_BOX2(pg)), real_t(1.0) / *dx / *dx,
                            real_t(1.0) / *dy / *dy});
  finish_local(who);
}

// :1616-1683 -- mean of the central-difference |grad phi| over admissible
// points (0 when there are none).  The reference sums serially; here the sum
// is FP64 compensated in a fixed order.
LSMHIP_EXPORT void lsm2dcomputeavegradphilocal_(real_t* ave, const real_t* phi, LSMHIP_BOXP2(pg), const real_t* dx,
                                                const real_t* dy, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(phi, ix, iy, nb);
  points_mean("lsm2dcomputeavegradphilocal_", LIST2, FBMASK2,
              GradNormOp{cv(phi, LSMHIP_BOX2(pg)), real_t(0.5) / *dx, real_t(0.5) / *dy}, ave);
}

LSMHIP_EXPORT void lsm2dgradientmagnitudelocal_(const real_t* px, const real_t* py, real_t* mag, LSMHIP_BOXP2(g),
                                                LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(px, py, mag, ix, iy, nb);
  const char* who = "lsm2dgradientmagnitudelocal_";
  const Box g = LSMHIP_BOX2(g);
  launch_points(who, LIST2, FBMASK2, GradMagOp{cv(px, g), cv(py, g), cv(mag, g)});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// RHS terms (src/toolbox/lsm_level_set_evolution2d_local.h)
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm2dzerooutlevelseteqnrhslocal_(real_t* rhs, LSMHIP_BOXP2(rg), const int* ix, const int* iy,
                                                    const int* nlo, const int* nhi) {
  LSMHIP_REQUIRE_ARRAYS(rhs, ix, iy);
  const char* who = "lsm2dzerooutlevelseteqnrhslocal_";
  launch_points(who, LIST2, no_mask(), ZeroOp{cv(rhs, LSMHIP_BOX2(rg))});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2daddadvectiontermtolserhslocal_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* px,
                                                       const real_t* py, LSMHIP_BOXP2(g), const real_t* vx,
                                                       const real_t* vy, LSMHIP_BOXP2(vg), LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, vx, vy, ix, iy, nb);
  const char* who = "lsm2daddadvectiontermtolserhslocal_";
  const Box g = LSMHIP_BOX2(g), vg = LSMHIP_BOX2(vg);
  launch_points(who, LIST2, FBMASK2,
                AdvectOp{cv(rhs, LSMHIP_BOX2(rg)), cv(px, g), cv(py, g), cv(vx, vg), cv(vy, vg)});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2daddnormalveltermtolserhslocal_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp,
                                                       const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                       const real_t* pym, LSMHIP_BOXP2(gm), const real_t* vn,
                                                       LSMHIP_BOXP2(vg), LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym, vn, ix, iy, nb);
  const char* who = "lsm2daddnormalveltermtolserhslocal_";
  launch_points(who, LIST2, FBMASK2,
                NormalOp{cv(rhs, LSMHIP_BOX2(rg)), four(pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm)),
                         cv(vn, LSMHIP_BOX2(vg)), real_t(0)});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2daddconstnormalveltermtolserhslocal_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp,
                                                            const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                            const real_t* pym, LSMHIP_BOXP2(gm), const real_t* vn,
                                                            LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym, ix, iy, nb);
  const char* who = "lsm2daddconstnormalveltermtolserhslocal_";
  const Box none{0, -1, 0, -1, 0, -1};
  launch_points(who, LIST2, FBMASK2,
                NormalOp{cv(rhs, LSMHIP_BOX2(rg)), four(pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm)),
                         cv((const real_t*)nullptr, none), *vn});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2daddconstcurvtermtolserhslocal_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* px,
                                                       const real_t* py, LSMHIP_BOXP2(g), const real_t* xx,
                                                       const real_t* xy, const real_t* yy, LSMHIP_BOXP2(g2),
                                                       const real_t* b, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, xx, xy, yy, ix, iy, nb);
  const char* who = "lsm2daddconstcurvtermtolserhslocal_";
  const Box g = LSMHIP_BOX2(g), g2 = LSMHIP_BOX2(g2);
  launch_points(who, LIST2, FBMASK2,
                ConstCurvOp{cv(rhs, LSMHIP_BOX2(rg)), cv(px, g), cv(py, g), cv(xx, g2), cv(xy, g2), cv(yy, g2), *b});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2daddconstprecomputedcurvtermtolserhslocal_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* kappa,
                                                                  LSMHIP_BOXP2(kg), const real_t* gmag,
                                                                  LSMHIP_BOXP2(gg), const real_t* b, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, kappa, gmag, ix, iy, nb);
  const char* who = "lsm2daddconstprecomputedcurvtermtolserhslocal_";
  launch_points(who, LIST2, FBMASK2,
                PrecompCurvOp{cv(rhs, LSMHIP_BOX2(rg)), cv(kappa, LSMHIP_BOX2(kg)), cv(gmag, LSMHIP_BOX2(gg)), *b});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2daddexternalandnormalveltermtolserhslocal_(real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp,
                                                                  const real_t* pyp, LSMHIP_BOXP2(gp),
                                                                  const real_t* pxm, const real_t* pym,
                                                                  LSMHIP_BOXP2(gm), const real_t* vn,
                                                                  const real_t* vx, const real_t* vy,
                                                                  LSMHIP_BOXP2(vg), LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym, vn, vx, vy, ix, iy, nb);
  const char* who = "lsm2daddexternalandnormalveltermtolserhslocal_";
  const Box vg = LSMHIP_BOX2(vg);
  launch_points(who, LIST2, FBMASK2,
                ExtNormalOp{cv(rhs, LSMHIP_BOX2(rg)), four(pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm)),
                            cv(vn, vg), cv(vx, vg), cv(vy, vg)});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// TVD Runge-Kutta (src/toolbox/lsm_tvd_runge_kutta2d_local.h)
// ---------------------------------------------------------------------------
template <int KIND>
static void rk_local2(const char* who, real_t* out, const Box& ob, const real_t* us, const Box& sb, const real_t* uc,
                      const Box& cb, const real_t* rhs, const Box& rb, real_t dt, const IndexList& L, const Mask& M) {
  launch_points(who, L, M, RkOp<KIND>{cv(out, ob), cv(us, sb), cv(uc, cb), cv(rhs, rb), dt});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2drk1steplocal_(real_t* un, LSMHIP_BOXP2(ng), const real_t* uc, LSMHIP_BOXP2(cg),
                                      const real_t* rhs, LSMHIP_BOXP2(rg), const real_t* dt, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs, ix, iy, nb);
  rk_local2<1>("lsm2drk1steplocal_", un, LSMHIP_BOX2(ng), uc, LSMHIP_BOX2(cg), uc, LSMHIP_BOX2(cg), rhs,
               LSMHIP_BOX2(rg), *dt, LIST2, FBMASK2);
}
LSMHIP_EXPORT void lsm2dtvdrk2stage1local_(real_t* un, LSMHIP_BOXP2(ng), const real_t* uc, LSMHIP_BOXP2(cg),
                                           const real_t* rhs, LSMHIP_BOXP2(rg), const real_t* dt, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs, ix, iy, nb);
  rk_local2<1>("lsm2dtvdrk2stage1local_", un, LSMHIP_BOX2(ng), uc, LSMHIP_BOX2(cg), uc, LSMHIP_BOX2(cg), rhs,
               LSMHIP_BOX2(rg), *dt, LIST2, FBMASK2);
}
LSMHIP_EXPORT void lsm2dtvdrk2stage2local_(real_t* un, LSMHIP_BOXP2(ng), const real_t* u1, LSMHIP_BOXP2(g1),
                                           const real_t* uc, LSMHIP_BOXP2(cg), const real_t* rhs, LSMHIP_BOXP2(rg),
                                           const real_t* dt, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, u1, uc, rhs, ix, iy, nb);
  rk_local2<2>("lsm2dtvdrk2stage2local_", un, LSMHIP_BOX2(ng), u1, LSMHIP_BOX2(g1), uc, LSMHIP_BOX2(cg), rhs,
               LSMHIP_BOX2(rg), *dt, LIST2, FBMASK2);
}
LSMHIP_EXPORT void lsm2dtvdrk3stage1local_(real_t* un, LSMHIP_BOXP2(ng), const real_t* uc, LSMHIP_BOXP2(cg),
                                           const real_t* rhs, LSMHIP_BOXP2(rg), const real_t* dt, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs, ix, iy, nb);
  rk_local2<1>("lsm2dtvdrk3stage1local_", un, LSMHIP_BOX2(ng), uc, LSMHIP_BOX2(cg), uc, LSMHIP_BOX2(cg), rhs,
               LSMHIP_BOX2(rg), *dt, LIST2, FBMASK2);
}
LSMHIP_EXPORT void lsm2dtvdrk3stage2local_(real_t* u2, LSMHIP_BOXP2(g2), const real_t* u1, LSMHIP_BOXP2(g1),
                                           const real_t* uc, LSMHIP_BOXP2(cg), const real_t* rhs, LSMHIP_BOXP2(rg),
                                           const real_t* dt, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(u2, u1, uc, rhs, ix, iy, nb);
  rk_local2<32>("lsm2dtvdrk3stage2local_", u2, LSMHIP_BOX2(g2), u1, LSMHIP_BOX2(g1), uc, LSMHIP_BOX2(cg), rhs,
                LSMHIP_BOX2(rg), *dt, LIST2, FBMASK2);
}
LSMHIP_EXPORT void lsm2dtvdrk3stage3local_(real_t* un, LSMHIP_BOXP2(ng), const real_t* u2, LSMHIP_BOXP2(g2),
                                           const real_t* uc, LSMHIP_BOXP2(cg), const real_t* rhs, LSMHIP_BOXP2(rg),
                                           const real_t* dt, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, u2, uc, rhs, ix, iy, nb);
  rk_local2<33>("lsm2dtvdrk3stage3local_", un, LSMHIP_BOX2(ng), u2, LSMHIP_BOX2(g2), uc, LSMHIP_BOX2(cg), rhs,
                LSMHIP_BOX2(rg), *dt, LIST2, FBMASK2);
}

// ---------------------------------------------------------------------------
// Reinitialization, orthogonalization, curvature, normals
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm2dcomputereinitializationeqnrhslocal_(real_t* out, LSMHIP_BOXP2(rg), const real_t* phi,
                                                            LSMHIP_BOXP2(pg), const real_t* phi0, LSMHIP_BOXP2(p0g),
                                                            const real_t* pxp, const real_t* pyp, LSMHIP_BOXP2(gp),
                                                            const real_t* pxm, const real_t* pym, LSMHIP_BOXP2(gm),
                                                            const real_t* dx, const real_t* dy,
                                                            const int* use_phi0_for_sgn, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(out, phi, pxp, pyp, pxm, pym, ix, iy, nb);
  if (*use_phi0_for_sgn == 1) LSMHIP_REQUIRE_ARRAYS(phi0);
  const char* who = "lsm2dcomputereinitializationeqnrhslocal_";
  const real_t h = *dx > *dy ? *dx : *dy;
  launch_points(who, LIST2, FBMASK2,
                ReinitOp{cv(out, LSMHIP_BOX2(rg)), cv(phi, LSMHIP_BOX2(pg)), cv(phi0, LSMHIP_BOX2(p0g)),
                         four(pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm)), h * h, *use_phi0_for_sgn == 1});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2dcomputeorthogonalizationeqnrhslocal_(
    real_t* rhs, LSMHIP_BOXP2(rg), const real_t* pxp, const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
    const real_t* pym, LSMHIP_BOXP2(gm), const real_t* psi, LSMHIP_BOXP2(sg), const real_t* qxp, const real_t* qyp,
    LSMHIP_BOXP2(qp), const real_t* qxm, const real_t* qym, LSMHIP_BOXP2(qm), const real_t* dx, const real_t* dy,
    LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pxm, pym, psi, qxp, qyp, qxm, qym, ix, iy, nb);
  const char* who = "lsm2dcomputeorthogonalizationeqnrhslocal_";
  const real_t h = *dx > *dy ? *dx : *dy;
  launch_points(who, LIST2, FBMASK2,
                OrthoOp{cv(rhs, LSMHIP_BOX2(rg)), four(pxp, pyp, LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm)),
                        cv(psi, LSMHIP_BOX2(sg)), four(qxp, qyp, LSMHIP_BOX2(qp), qxm, qym, LSMHIP_BOX2(qm)), h * h});
  finish_local(who);
}

template <int ORDER>
static void mean_curv_local2(const char* who, real_t* kappa, const Box& kg, const real_t* phi, const Box& pg,
                             const real_t* px, const real_t* py, const real_t* gmag, const Box& gg, real_t dx,
                             real_t dy, const IndexList& L, const Mask& M) {
  const real_t c12 = real_t(0.0833333333333333333333);
  const real_t fxx = ORDER == 2 ? real_t(1.0) / dx / dx : c12 / dx / dx;
  const real_t fyy = ORDER == 2 ? real_t(1.0) / dy / dy : c12 / dy / dy;
  const real_t fxy = ORDER == 2 ? real_t(0.25) / dx / dy : c12 / dy;
  launch_points(who, L, M,
                MeanCurvOp<ORDER>{cv(kappa, kg), cv(phi, pg), cv(px, gg), cv(py, gg), cv(gmag, gg), fxx, fyy, fxy});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2dcomputemeancurvatureorder2local_(real_t* kappa, LSMHIP_BOXP2(kg), const real_t* phi,
                                                         LSMHIP_BOXP2(pg), const real_t* px, const real_t* py,
                                                         const real_t* gmag, LSMHIP_BOXP2(gg), const real_t* dx,
                                                         const real_t* dy, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, gmag, ix, iy, nb);
  mean_curv_local2<2>("lsm2dcomputemeancurvatureorder2local_", kappa, LSMHIP_BOX2(kg), phi, LSMHIP_BOX2(pg), px, py,
                      gmag, LSMHIP_BOX2(gg), *dx, *dy, LIST2, FBMASK2);
}

LSMHIP_EXPORT void lsm2dcomputemeancurvatureorder4local_(real_t* kappa, LSMHIP_BOXP2(kg), const real_t* phi,
                                                         LSMHIP_BOXP2(pg), const real_t* px, const real_t* py,
                                                         const real_t* gmag, LSMHIP_BOXP2(gg), const real_t* dx,
                                                         const real_t* dy, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, gmag, ix, iy, nb);
  mean_curv_local2<4>("lsm2dcomputemeancurvatureorder4local_", kappa, LSMHIP_BOX2(kg), phi, LSMHIP_BOX2(pg), px, py,
                      gmag, LSMHIP_BOX2(gg), *dx, *dy, LIST2, FBMASK2);
}

// dx and dy are unused, as in the reference
LSMHIP_EXPORT void lsm2dcomputeunitnormallocal_(real_t* nx, real_t* ny, LSMHIP_BOXP2(ng), const real_t* px,
                                                const real_t* py, LSMHIP_BOXP2(gg), const real_t* dx,
                                                const real_t* dy, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(nx, ny, px, py, ix, iy, nb);
  const char* who = "lsm2dcomputeunitnormallocal_";
  (void)dx;
  (void)dy;
  const Box ng = LSMHIP_BOX2(ng), gg = LSMHIP_BOX2(gg);
  launch_points(who, LIST2, FBMASK2,
                NormalVecOp<0>{cv(nx, ng), cv(ny, ng), cv(px, gg), cv(py, gg), cv(px, gg), real_t(0)});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2dcomputesignedunitnormallocal_(real_t* nx, real_t* ny, LSMHIP_BOXP2(ng), const real_t* px,
                                                      const real_t* py, LSMHIP_BOXP2(gg), const real_t* phi,
                                                      LSMHIP_BOXP2(pg), const real_t* dx, const real_t* dy,
                                                      LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(nx, ny, px, py, phi, ix, iy, nb);
  const char* who = "lsm2dcomputesignedunitnormallocal_";
  const Box ng = LSMHIP_BOX2(ng), gg = LSMHIP_BOX2(gg);
  const real_t h = *dx > *dy ? *dx : *dy;
  launch_points(who, LIST2, FBMASK2,
                NormalVecOp<1>{cv(nx, ng), cv(ny, ng), cv(px, gg), cv(py, gg), cv(phi, LSMHIP_BOX2(pg)), h * h});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// Mask / copy / cut-off (src/toolbox/lsm_localization2d.h)
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm2dimposemasklocal_(real_t* dst, LSMHIP_BOXP2(dg), const real_t* src, LSMHIP_BOXP2(sg),
                                         const real_t* msk, LSMHIP_BOXP2(mg), const int* ix, const int* iy,
                                         const int* nlo, const int* nhi) {
  LSMHIP_REQUIRE_ARRAYS(dst, src, msk, ix, iy);
  const char* who = "lsm2dimposemasklocal_";
  launch_points(who, LIST2, no_mask(),
                ImposeMaskOp{cv(dst, LSMHIP_BOX2(dg)), cv(src, LSMHIP_BOX2(sg)), cv(msk, LSMHIP_BOX2(mg))});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2dcopydatalocal_(real_t* dst, LSMHIP_BOXP2(dg), const real_t* src, LSMHIP_BOXP2(sg),
                                       const int* ix, const int* iy, const int* nlo, const int* nhi) {
  LSMHIP_REQUIRE_ARRAYS(dst, src, ix, iy);
  const char* who = "lsm2dcopydatalocal_";
  launch_points(who, LIST2, no_mask(), CopyOp{cv(dst, LSMHIP_BOX2(dg)), cv(src, LSMHIP_BOX2(sg))});
  finish_local(who);
}

LSMHIP_EXPORT void lsm2dmultiplycutofflserhslocal_(const real_t* phi, real_t* rhs, LSMHIP_BOXP2(g), const int* ix,
                                                   const int* iy, const int* nlo, const int* nhi,
                                                   const unsigned char* nb, LSMHIP_BOXP2(nbb),
                                                   const unsigned char* mark_fb, const real_t* beta,
                                                   const real_t* gamma) {
  LSMHIP_REQUIRE_ARRAYS(phi, rhs, ix, iy, nb);
  const char* who = "lsm2dmultiplycutofflserhslocal_";
  const Box g = LSMHIP_BOX2(g);
  const real_t c1 = *gamma - real_t(3) * *beta;
  real_t c2 = *gamma - *beta;
  c2 = c2 * c2 * c2;
  launch_points(who, LIST2, FBMASK2, CutOffOp{cv(phi, g), cv(rhs, g), *beta, *gamma, c1, c2});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// Reductions (src/toolbox/lsm_math_utils2d_local.h)
// ---------------------------------------------------------------------------
// :16-81 -- 0 when no point is admissible
LSMHIP_EXPORT void lsm2dmaxnormdifflocal_(real_t* out, const real_t* f1, LSMHIP_BOXP2(g1), const real_t* f2,
                                          LSMHIP_BOXP2(g2), LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(f1, f2, ix, iy, nb);
  const char* who = "lsm2dmaxnormdifflocal_";
  double m = 0.0;
  bool any;
  if (!points_max(who, LIST2, FBMASK2, MaxNormDiffOp{cv(f1, LSMHIP_BOX2(g1)), cv(f2, LSMHIP_BOX2(g2))}, &m, &any))
    return;
  *out = any ? real_t(m) : real_t(0);
}

// :165-241
LSMHIP_EXPORT void lsm2dcomputestableadvectiondtlocal_(real_t* dt, const real_t* vx, const real_t* vy,
                                                       LSMHIP_BOXP2(vg), const real_t* dx, const real_t* dy,
                                                       const real_t* cfl, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(vx, vy, ix, iy, nb);
  const char* who = "lsm2dcomputestableadvectiondtlocal_";
  const Box vg = LSMHIP_BOX2(vg);
  double m = -1.0;
  bool any;
  if (!points_max(who, LIST2, FBMASK2, AdvDtOp{cv(vx, vg), cv(vy, vg), real_t(1.0) / *dx, real_t(1.0) / *dy}, &m,
                  &any))
    return;
  if (!any) m = -1.0;
  *dt = *cfl / (real_t(m) + kTiny);
}

static void normal_dt_local2(const char* who, real_t* dt, const real_t* vn, const Box& vg, real_t vc,
                             const real_t* pxp, const real_t* pyp, const Box& gp, const real_t* pxm,
                             const real_t* pym, const Box& gm, real_t dx, real_t dy, real_t cfl, const IndexList& L,
                             const Mask& M) {
  const real_t h = dx > dy ? dx : dy;
  double m = -1.0;
  bool any;
  NormalDtOp op{four(pxp, pyp, gp, pxm, pym, gm), cv(vn, vg), vc, h * h, real_t(1.0) / dx, real_t(1.0) / dy};
  if (!points_max(who, L, M, op, &m, &any)) return;
  if (!any) m = -1.0;
  *dt = cfl / (real_t(m) + kTiny);
}

// :246-354
LSMHIP_EXPORT void lsm2dcomputestablenormalveldtlocal_(real_t* dt, const real_t* vn, LSMHIP_BOXP2(vg),
                                                       const real_t* pxp, const real_t* pyp, LSMHIP_BOXP2(gp),
                                                       const real_t* pxm, const real_t* pym, LSMHIP_BOXP2(gm),
                                                       const real_t* dx, const real_t* dy, const real_t* cfl,
                                                       LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(vn, pxp, pyp, pxm, pym, ix, iy, nb);
  normal_dt_local2("lsm2dcomputestablenormalveldtlocal_", dt, vn, LSMHIP_BOX2(vg), real_t(0), pxp, pyp,
                   LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), *dx, *dy, *cfl, LIST2, FBMASK2);
}

// :360-462
LSMHIP_EXPORT void lsm2dcomputestableconstnormalveldtlocal_(real_t* dt, const real_t* vn, const real_t* pxp,
                                                            const real_t* pyp, LSMHIP_BOXP2(gp), const real_t* pxm,
                                                            const real_t* pym, LSMHIP_BOXP2(gm), const real_t* dx,
                                                            const real_t* dy, const real_t* cfl, LOCAL2_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pxm, pym, ix, iy, nb);
  const Box none{0, -1, 0, -1, 0, -1};
  normal_dt_local2("lsm2dcomputestableconstnormalveldtlocal_", dt, nullptr, none, real_t(fabs(*vn)), pxp, pyp,
                   LSMHIP_BOX2(gp), pxm, pym, LSMHIP_BOX2(gm), *dx, *dy, *cfl, LIST2, FBMASK2);
}
