// local3d.hip -- narrow-band ("local") twins of the whole-grid kernels:
// HJ-ENO1/2 and HJ-WENO5 derivatives, central gradients, Laplacian,
// |grad phi|, RHS terms, TVD Runge-Kutta stages, reinitialization RHS, mean
// curvature, CFL steps, max-norm difference, averaged |grad phi|, mask/copy
// helpers and the RHS cut-off.
// References: src/toolbox/lsm_spatial_derivatives3d_local.f.in,
// src/toolbox/lsm_level_set_evolution3d_local.f.in, src/toolbox/lsm_tvd_runge_kutta3d_local.f,
// src/toolbox/lsm_math_utils3d_local.f.in, src/reinitialization/lsm_reinitialization3d_local.f.in,
// src/geometry/lsm_curvature3d_local.f.in, src/toolbox/lsm_localization3d.f:482-774.
//
// One thread per index-list entry l: (i,j,k) = (index_x, index_y, index_z)(l),
// where an index array declared index_x(nlo:nhi) in the reference is a
// pointer to its element nlo.  Points whose narrow_band mark (a signed byte)
// exceeds the fill-box mark are skipped, as in the reference.  The derivative
// routines keep the reference's use of the caller's D1/D2 scratch: differences
// are written at the listed points of levels 0..L (1e10 where the mark exceeds
// mark_D1) and then read at the level-0 points, so entries the lists do not
// cover keep whatever the caller's array held, exactly as in the reference.
// Each index and stencil neighbour is checked against its array's ghost box; a
// point that would leave it is skipped and raises the device flag
// (lsmhip_get_last_error then reports kErrIndex) where the reference would
// read or write out of bounds.
#include "local.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

struct Axis3 {
  CV<real_t> o[3];
};
Axis3 axis3(real_t* x, real_t* y, real_t* z, const Box& b) { return Axis3{{cv(x, b), cv(y, b), cv(z, b)}}; }

// Level-0 derivatives from the difference arrays along axis d.
template <int ORDER>
struct DerivOp {
  CV<real_t> p, m;
  CV<const real_t> d1, d2;
  int di, dj, dk;
  real_t inv;
  __device__ bool operator()(int i, int j, int k) const {
    if (!p.has(i, j, k) || !m.has(i, j, k)) return false;
    auto D1 = [&](int s) -> real_t { return d1(i + s * di, j + s * dj, k + s * dk); };
    auto D2 = [&](int s) -> real_t { return d2(i + s * di, j + s * dj, k + s * dk); };
    if (ORDER == 1) {
      if (!d1.has(i, j, k) || !d1.has(i + di, j + dj, k + dk)) return false;
      p(i, j, k) = D1(1) * inv;
      m(i, j, k) = D1(0) * inv;
    } else if (ORDER == 2) {
      if (!d1.has(i, j, k) || !d1.has(i + di, j + dj, k + dk) || !d2.has(i - di, j - dj, k - dk) ||
          !d2.has(i + di, j + dj, k + dk))
        return false;
      p(i, j, k) = (fabs(D2(0)) < fabs(D2(1)) ? D1(1) - real_t(0.5) * D2(0) : D1(1) - real_t(0.5) * D2(1)) * inv;
      m(i, j, k) = (fabs(D2(-1)) < fabs(D2(0)) ? D1(0) + real_t(0.5) * D2(-1) : D1(0) + real_t(0.5) * D2(0)) * inv;
    } else {
      if (!d1.has(i - 2 * di, j - 2 * dj, k - 2 * dk) || !d1.has(i + 3 * di, j + 3 * dj, k + 3 * dk)) return false;
      p(i, j, k) = stencil::weno5(D1(3) * inv, D1(2) * inv, D1(1) * inv, D1(0) * inv, D1(-1) * inv);
      m(i, j, k) = stencil::weno5(D1(-2) * inv, D1(-1) * inv, D1(0) * inv, D1(1) * inv, D1(2) * inv);
    }
    return true;
  }
};

// ----- per-point ops --------------------------------------------------------------------------
struct Six {
  CV<const real_t> xp, yp, zp, xm, ym, zm;
  __device__ bool has(int i, int j, int k) const { return xp.has(i, j, k) && xm.has(i, j, k); }
};
Six six(const real_t* xp, const real_t* yp, const real_t* zp, const Box& gp, const real_t* xm, const real_t* ym,
        const real_t* zm, const Box& gm) {
  return Six{cv(xp, gp), cv(yp, gp), cv(zp, gp), cv(xm, gm), cv(ym, gm), cv(zm, gm)};
}

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
  CV<const real_t> px, py, pz, vx, vy, vz;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !px.has(i, j, k) || !vx.has(i, j, k)) return false;
    r(i, j, k) = r(i, j, k) - (vx(i, j, k) * px(i, j, k) + vy(i, j, k) * py(i, j, k) + vz(i, j, k) * pz(i, j, k));
    return true;
  }
};

struct NormalOp {
  CV<real_t> r;
  Six g;
  CV<const real_t> vn;  // p null: constant speed vc
  real_t vc;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !g.has(i, j, k) || (vn.v.p && !vn.has(i, j, k))) return false;
    const real_t v = vn.v.p ? vn(i, j, k) : vc;
    const real_t s = stencil::godunov_sq(v > real_t(0), g.xp(i, j, k), g.yp(i, j, k), g.zp(i, j, k), g.xm(i, j, k),
                                         g.ym(i, j, k), g.zm(i, j, k));
    if (fabs(v) >= kZeroTol) r(i, j, k) = r(i, j, k) - v * sqrt(s);
    return true;
  }
};

struct ConstCurvOp {
  CV<real_t> r;
  CV<const real_t> px, py, pz, xx, xy, xz, yy, yz, zz;
  real_t b;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !px.has(i, j, k) || !xx.has(i, j, k)) return false;
    const real_t x = px(i, j, k), y = py(i, j, k), z = pz(i, j, k);
    const real_t g2 = x * x + y * y + z * z;
    real_t curv = real_t(0);
    if (!(g2 < kZeroTol))
      curv = stencil::curv_num(xx(i, j, k), xy(i, j, k), xz(i, j, k), yy(i, j, k), yz(i, j, k), zz(i, j, k), x, y,
                               z) / g2;
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

struct ExtNormalOp {
  CV<real_t> r;
  Six g;
  CV<const real_t> vn, vx, vy, vz;
  __device__ bool operator()(int i, int j, int k) const {
    if (!r.has(i, j, k) || !g.has(i, j, k) || !vn.has(i, j, k)) return false;
    const real_t n = vn(i, j, k), u = vx(i, j, k), v = vy(i, j, k), w = vz(i, j, k);
    const real_t px = stencil::ext_normal_pick(u, n, g.xp(i, j, k), g.xm(i, j, k));
    const real_t py = stencil::ext_normal_pick(v, n, g.yp(i, j, k), g.ym(i, j, k));
    const real_t pz = stencil::ext_normal_pick(w, n, g.zp(i, j, k), g.zm(i, j, k));
    const real_t s = px * px + py * py + pz * pz;
    if (fabs(n) >= kZeroTol && s >= kZeroTol)
      r(i, j, k) = r(i, j, k) - u * px - v * py - w * pz - n * sqrt(s);
    else
      r(i, j, k) = r(i, j, k) - u * px - v * py - w * pz;
    return true;
  }
};

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

struct ReinitOp {
  CV<real_t> out;
  CV<const real_t> phi, phi0;
  Six g;
  real_t dx_sq;
  int use0;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.has(i, j, k) || !g.has(i, j, k)) return false;
    if (use0 ? !phi0.has(i, j, k) : !phi.has(i, j, k)) return false;
    const real_t s = use0 ? phi0(i, j, k) : phi(i, j, k);
    out(i, j, k) = stencil::reinit_rhs(s, use0 != 0, g.xp(i, j, k), g.yp(i, j, k), g.zp(i, j, k), g.xm(i, j, k),
                                       g.ym(i, j, k), g.zm(i, j, k), dx_sq);
    return true;
  }
};

struct MeanCurvOp {
  CV<real_t> kap;
  CV<const real_t> F, px, py, pz, gm;
  real_t fxx, fyy, fzz, fxy, fxz, fyz;
  __device__ bool operator()(int i, int j, int k) const {
    if (!kap.has(i, j, k) || !px.has(i, j, k) || !F.star(i, j, k, 1)) return false;
    const real_t c = F(i, j, k);
    const real_t xx = (F(i + 1, j, k) - real_t(2) * c + F(i - 1, j, k)) * fxx;
    const real_t yy = (F(i, j + 1, k) - real_t(2) * c + F(i, j - 1, k)) * fyy;
    const real_t zz = (F(i, j, k + 1) - real_t(2) * c + F(i, j, k - 1)) * fzz;
    const real_t xy = (F(i + 1, j + 1, k) - F(i + 1, j - 1, k) - F(i - 1, j + 1, k) + F(i - 1, j - 1, k)) * fxy;
    const real_t xz = (F(i + 1, j, k + 1) - F(i + 1, j, k - 1) - F(i - 1, j, k + 1) + F(i - 1, j, k - 1)) * fxz;
    const real_t yz = (F(i, j + 1, k + 1) - F(i, j + 1, k - 1) - F(i, j - 1, k + 1) + F(i, j - 1, k - 1)) * fyz;
    const real_t g = gm(i, j, k), den = g * g * g;
    kap(i, j, k) =
        den < kZeroTol ? real_t(0) : stencil::curv_num(xx, xy, xz, yy, yz, zz, px(i, j, k), py(i, j, k), pz(i, j, k)) / den;
    return true;
  }
};

// Gaussian curvature (second order) on a band, lsm_curvature3d_local.f.in:151-280:
// the compact second differences of the mean curvature, the Gaussian numerator
// over |grad phi|^4, 0 where that is below zero_tol.
struct GaussCurv2Op {
  CV<real_t> kap;
  CV<const real_t> F, px, py, pz, gm;
  real_t fxx, fyy, fzz, fxy, fxz, fyz;
  __device__ bool operator()(int i, int j, int k) const {
    if (!kap.has(i, j, k) || !px.has(i, j, k) || !F.star(i, j, k, 1)) return false;
    const real_t c = F(i, j, k);
    const real_t xx = (F(i + 1, j, k) - real_t(2) * c + F(i - 1, j, k)) * fxx;
    const real_t yy = (F(i, j + 1, k) - real_t(2) * c + F(i, j - 1, k)) * fyy;
    const real_t zz = (F(i, j, k + 1) - real_t(2) * c + F(i, j, k - 1)) * fzz;
    const real_t xy = (F(i + 1, j + 1, k) - F(i + 1, j - 1, k) - F(i - 1, j + 1, k) + F(i - 1, j - 1, k)) * fxy;
    const real_t xz = (F(i + 1, j, k + 1) - F(i + 1, j, k - 1) - F(i - 1, j, k + 1) + F(i - 1, j, k - 1)) * fxz;
    const real_t yz = (F(i, j + 1, k + 1) - F(i, j + 1, k - 1) - F(i, j - 1, k + 1) + F(i, j - 1, k - 1)) * fyz;
    const real_t g = gm(i, j, k), den = g * g * g * g;
    kap(i, j, k) = den < kZeroTol ? real_t(0)
                                  : stencil::gauss_num(px(i, j, k), py(i, j, k), pz(i, j, k), xx, yy, zz, xy, xz, yz) /
                                        den;
    return true;
  }
};

// The fourth-order curvature of lsm_curvature3d_local.f.in:285-426 and :432-573:
// fourth-order second differences of phi, mixed derivatives as fourth-order
// differences of the given first derivatives (phi_xy = d/dy phi_x, phi_yz =
// d/dz phi_y, phi_zx = d/dx phi_z), the Gaussian numerator over |grad phi|^4.
// Both routines of the reference compute this same expression (the "mean"
// one included), and so do both entry points here.
struct Curv4Op {
  CV<real_t> kap;
  CV<const real_t> F, px, py, pz, gm;
  real_t f2x, f2y, f2z, f1x, f1y, f1z;
  __device__ bool operator()(int i, int j, int k) const {
    if (!kap.has(i, j, k) || !F.star(i, j, k, 2) || !px.star(i, j, k, 2)) return false;
    const real_t c = F(i, j, k);
    const real_t xx =
        (-F(i + 2, j, k) + real_t(16) * F(i + 1, j, k) - real_t(30) * c - F(i - 2, j, k) + real_t(16) * F(i - 1, j, k)) *
        f2x;
    const real_t yy =
        (-F(i, j + 2, k) + real_t(16) * F(i, j + 1, k) - real_t(30) * c - F(i, j - 2, k) + real_t(16) * F(i, j - 1, k)) *
        f2y;
    const real_t zz =
        (-F(i, j, k + 2) + real_t(16) * F(i, j, k + 1) - real_t(30) * c - F(i, j, k - 2) + real_t(16) * F(i, j, k - 1)) *
        f2z;
    const real_t xy =
        (-px(i, j + 2, k) + real_t(8) * px(i, j + 1, k) + px(i, j - 2, k) - real_t(8) * px(i, j - 1, k)) * f1y;
    const real_t yz =
        (-py(i, j, k + 2) + real_t(8) * py(i, j, k + 1) + py(i, j, k - 2) - real_t(8) * py(i, j, k - 1)) * f1z;
    const real_t zx =
        (-pz(i + 2, j, k) + real_t(8) * pz(i + 1, j, k) + pz(i - 2, j, k) - real_t(8) * pz(i - 1, j, k)) * f1x;
    const real_t g = gm(i, j, k), den = g * g * g * g;
    kap(i, j, k) = den < kZeroTol ? real_t(0)
                                  : stencil::gauss_num(px(i, j, k), py(i, j, k), pz(i, j, k), xx, yy, zz, xy, zx, yz) /
                                        den;
    return true;
  }
};

// Orthogonalization RHS on a band, lsm_reinitialization3d_local.f.in:326-522
// (the whole-grid formula of lsm_reinitialization3d.f.in:296-476 at each point).
struct OrthoOp {
  CV<real_t> rhs;
  CV<const real_t> xp, yp, zp, xm, ym, zm, psi, qxp, qyp, qzp, qxm, qym, qzm;
  real_t dx_sq;
  __device__ bool operator()(int i, int j, int k) const {
    if (!rhs.has(i, j, k) || !xp.has(i, j, k) || !xm.has(i, j, k) || !psi.has(i, j, k) || !qxp.has(i, j, k) ||
        !qxm.has(i, j, k))
      return false;
    const real_t g1 = real_t(0.5) * (qxp(i, j, k) + qxm(i, j, k));
    const real_t g2 = real_t(0.5) * (qyp(i, j, k) + qym(i, j, k));
    const real_t g3 = real_t(0.5) * (qzp(i, j, k) + qzm(i, j, k));
    const real_t ng = sqrt(g1 * g1 + g2 * g2 + g3 * g3);
    const real_t s = psi(i, j, k);
    real_t out = real_t(0);
    if (fabs(s) > kZeroTol && ng > kZeroTol) {
      const real_t sgn = s / sqrt(s * s + dx_sq * (ng * ng));
      const real_t a1 = sgn * g1, a2 = sgn * g2, a3 = sgn * g3, z = real_t(0);
      out = real_t(-1) / ng *
            (maxr(a1, z) * xm(i, j, k) + minr(a1, z) * xp(i, j, k) + maxr(a2, z) * ym(i, j, k) +
             minr(a2, z) * yp(i, j, k) + maxr(a3, z) * zm(i, j, k) + minr(a3, z) * zp(i, j, k));
    }
    rhs(i, j, k) = out;
    return true;
  }
};

struct Central2Op {
  Axis3 out;
  CV<const real_t> F;
  real_t fx, fy, fz;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.o[0].has(i, j, k) || !F.star(i, j, k, 1)) return false;
    out.o[0](i, j, k) = (F(i + 1, j, k) - F(i - 1, j, k)) * fx;
    out.o[1](i, j, k) = (F(i, j + 1, k) - F(i, j - 1, k)) * fy;
    out.o[2](i, j, k) = (F(i, j, k + 1) - F(i, j, k - 1)) * fz;
    return true;
  }
};

struct Central4Op {
  Axis3 out;
  CV<const real_t> F;
  real_t fx, fy, fz;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.o[0].has(i, j, k) || !F.star(i, j, k, 2)) return false;
    const real_t e = real_t(8);
    out.o[0](i, j, k) = (-F(i + 2, j, k) + e * F(i + 1, j, k) + F(i - 2, j, k) - e * F(i - 1, j, k)) * fx;
    out.o[1](i, j, k) = (-F(i, j + 2, k) + e * F(i, j + 1, k) + F(i, j - 2, k) - e * F(i, j - 1, k)) * fy;
    out.o[2](i, j, k) = (-F(i, j, k + 2) + e * F(i, j, k + 1) + F(i, j, k - 2) - e * F(i, j, k - 1)) * fz;
    return true;
  }
};

struct LaplacianOp {
  CV<real_t> out;
  CV<const real_t> F;
  real_t ax, ay, az;
  __device__ bool operator()(int i, int j, int k) const {
    if (!out.has(i, j, k) || !F.star(i, j, k, 1)) return false;
    const real_t c = F(i, j, k);
    out(i, j, k) = ax * (F(i + 1, j, k) - real_t(2) * c + F(i - 1, j, k)) +
                   ay * (F(i, j + 1, k) - real_t(2) * c + F(i, j - 1, k)) +
                   az * (F(i, j, k + 1) - real_t(2) * c + F(i, j, k - 1));
    return true;
  }
};

struct GradMagOp {
  CV<const real_t> x, y, z;
  CV<real_t> mag;
  __device__ bool operator()(int i, int j, int k) const {
    if (!mag.has(i, j, k)) return false;
    mag(i, j, k) = stencil::grad_magnitude(x(i, j, k), y(i, j, k), z(i, j, k), kZeroTol);
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

// lsm3dMultiplyCutOffLSERHSLOCAL, lsm_localization3d.f:482-561
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

}  // namespace

// ---------------------------------------------------------------------------
// Derivative entry points (src/toolbox/lsm_spatial_derivatives3d_local.h)
// ---------------------------------------------------------------------------
#define LPM_ARGS                                                                                             \
  real_t *pxp, real_t *pyp, real_t *pzp, LSMHIP_BOXP(gp), real_t *pxm, real_t *pym, real_t *pzm,            \
      LSMHIP_BOXP(gm), const real_t *phi, LSMHIP_BOXP(pg)

template <int ORDER>
static void local_derivs(const char* who, real_t* pxp, real_t* pyp, real_t* pzp, const Box& gp, real_t* pxm,
                         real_t* pym, real_t* pzm, const Box& gm, const real_t* phi, const Box& pg, real_t* D1,
                         const Box& d1b, real_t* D2, const Box& d2b, real_t dx, real_t dy, real_t dz, const int* ix,
                         const int* iy, const int* iz, int lo0, int hi0, int hi1, int hiD, const unsigned char* nb,
                         const Box& nbb, const unsigned char* mark_fb, const unsigned char* mark_d1,
                         const unsigned char* mark_d2) {
  const real_t inv[3] = {real_t(1.0) / dx, real_t(1.0) / dy, real_t(1.0) / dz};
  real_t* P[3] = {pxp, pyp, pzp};
  real_t* M[3] = {pxm, pym, pzm};
  const Mask nbm = make_mask(nb, nbb, nullptr, 0);
  const Mask fbm = make_mask(nb, nbb, mark_fb, 1);
  const IndexList all = list(ix, iy, iz, lo0, hiD, lo0);
  const IndexList lev0 = list(ix, iy, iz, lo0, hi0, lo0);
  for (int d = 0; d < 3; d++) {
    dn_local(who, D1, d1b, phi, pg, 1, d, all, nbm, (signed char)*mark_d1);
    if (ORDER == 2) dn_local(who, D2, d2b, D1, d1b, 2, d, list(ix, iy, iz, lo0, hi1, lo0), nbm, (signed char)*mark_d2);
    DerivOp<ORDER> op{cv(P[d], gp), cv(M[d], gm), cv((const real_t*)D1, d1b),
                      cv((const real_t*)(ORDER == 2 ? D2 : D1), ORDER == 2 ? d2b : d1b), d == 0, d == 1, d == 2,
                      inv[d]};
    launch_points(who, lev0, fbm, op);
  }
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dhjeno1local_(LPM_ARGS, real_t* D1, LSMHIP_BOXP(d1), const real_t* dx, const real_t* dy,
                                     const real_t* dz, const int* ix, const int* iy, const int* iz, const int* nlo0,
                                     const int* nhi0, const int* nlo1, const int* nhi1, const unsigned char* nb,
                                     LSMHIP_BOXP(nbb), const unsigned char* mark_fb, const unsigned char* mark_d1) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, phi, D1, ix, iy, iz, nb);
  (void)nlo1;
  local_derivs<1>("lsm3dhjeno1local_", pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm), phi,
                  LSMHIP_BOX(pg), D1, LSMHIP_BOX(d1), nullptr, LSMHIP_BOX(d1), *dx, *dy, *dz, ix, iy, iz, *nlo0,
                  *nhi0, *nhi1, *nhi1, nb, LSMHIP_BOX(nbb), mark_fb, mark_d1, nullptr);
}

LSMHIP_EXPORT void lsm3dhjeno2local_(LPM_ARGS, real_t* D1, LSMHIP_BOXP(d1), real_t* D2, LSMHIP_BOXP(d2),
                                     const real_t* dx, const real_t* dy, const real_t* dz, const int* ix,
                                     const int* iy, const int* iz, const int* nlo0, const int* nhi0, const int* nlo1,
                                     const int* nhi1, const int* nlo2, const int* nhi2, const unsigned char* nb,
                                     LSMHIP_BOXP(nbb), const unsigned char* mark_fb, const unsigned char* mark_d1,
                                     const unsigned char* mark_d2) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, phi, D1, D2, ix, iy, iz, nb);
  (void)nlo1;
  (void)nlo2;
  local_derivs<2>("lsm3dhjeno2local_", pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm), phi,
                  LSMHIP_BOX(pg), D1, LSMHIP_BOX(d1), D2, LSMHIP_BOX(d2), *dx, *dy, *dz, ix, iy, iz, *nlo0, *nhi0,
                  *nhi1, *nhi2, nb, LSMHIP_BOX(nbb), mark_fb, mark_d1, mark_d2);
}

// 3D HJ-WENO5 on a narrow band.  The reference has no 3D routine; this one
// follows lsm2dHJWENO5LOCAL (src/toolbox/lsm_spatial_derivatives2d_local.f.in:816-1154):
// D1 over levels 0..3 filtered by mark_D1, WENO5 of the 3D whole-grid routine
// on level 0 filtered by mark_fb (with the 3D templated tiny number 1e-36).
LSMHIP_EXPORT void lsm3dhjweno5local_(LPM_ARGS, real_t* D1, LSMHIP_BOXP(d1), const real_t* dx, const real_t* dy,
                                      const real_t* dz, const int* ix, const int* iy, const int* iz,
                                      const int* nlo0, const int* nhi0, const int* nlo1, const int* nhi1,
                                      const int* nlo2, const int* nhi2, const int* nlo3, const int* nhi3,
                                      const unsigned char* nb, LSMHIP_BOXP(nbb), const unsigned char* mark_fb,
                                      const unsigned char* mark_d1) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, phi, D1, ix, iy, iz, nb);
  (void)nlo1;
  (void)nhi1;
  (void)nlo2;
  (void)nhi2;
  (void)nlo3;
  local_derivs<5>("lsm3dhjweno5local_", pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm), phi,
                  LSMHIP_BOX(pg), D1, LSMHIP_BOX(d1), nullptr, LSMHIP_BOX(d1), *dx, *dy, *dz, ix, iy, iz, *nlo0,
                  *nhi0, *nhi3, *nhi3, nb, LSMHIP_BOX(nbb), mark_fb, mark_d1, nullptr);
}

#define LOCAL_TAIL                                                                                           \
  const int *ix, const int *iy, const int *iz, const int *nlo, const int *nhi, const unsigned char *nb,      \
      LSMHIP_BOXP(nbb), const unsigned char *mark_fb
#define LIST list(ix, iy, iz, *nlo, *nhi, *nlo)
#define FBMASK make_mask(nb, LSMHIP_BOX(nbb), mark_fb, 1)

LSMHIP_EXPORT void lsm3dcentralgradorder2local_(real_t* px, real_t* py, real_t* pz, LSMHIP_BOXP(g), const real_t* phi,
                                                LSMHIP_BOXP(pg), const real_t* dx, const real_t* dy, const real_t* dz,
                                                LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi, ix, iy, iz, nb);
  const char* who = "lsm3dcentralgradorder2local_";
  launch_points(who, LIST, FBMASK,
                Central2Op{axis3(px, py, pz, LSMHIP_BOX(g)), cv(phi, LSMHIP_BOX(pg)), real_t(0.5) / *dx,
                           real_t(0.5) / *dy, real_t(0.5) / *dz});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dcentralgradorder4local_(real_t* px, real_t* py, real_t* pz, LSMHIP_BOXP(g), const real_t* phi,
                                                LSMHIP_BOXP(pg), const real_t* dx, const real_t* dy, const real_t* dz,
                                                LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, phi, ix, iy, iz, nb);
  const char* who = "lsm3dcentralgradorder4local_";
  const real_t c = real_t(0.0833333333333333333333);
  launch_points(who, LIST, FBMASK,
                Central4Op{axis3(px, py, pz, LSMHIP_BOX(g)), cv(phi, LSMHIP_BOX(pg)), c / *dx, c / *dy, c / *dz});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dlaplacianorder2local_(real_t* lap, LSMHIP_BOXP(lg), const real_t* phi, LSMHIP_BOXP(pg),
                                              const real_t* dx, const real_t* dy, const real_t* dz, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(lap, phi, ix, iy, iz, nb);
  const char* who = "lsm3dlaplacianorder2local_";
  launch_points(who, LIST, FBMASK,
                LaplacianOp{cv(lap, LSMHIP_BOX(lg)), cv(phi, LSMHIP_BOX(pg)), real_t(1.0) / *dx / *dx,
                            real_t(1.0) / *dy / *dy, real_t(1.0) / *dz / *dz});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dgradientmagnitudelocal_(const real_t* px, const real_t* py, const real_t* pz, real_t* mag,
                                                LSMHIP_BOXP(g), LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(px, py, pz, mag, ix, iy, iz, nb);
  const char* who = "lsm3dgradientmagnitudelocal_";
  const Box g = LSMHIP_BOX(g);
  launch_points(who, LIST, FBMASK, GradMagOp{cv(px, g), cv(py, g), cv(pz, g), cv(mag, g)});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// RHS terms (src/toolbox/lsm_level_set_evolution3d_local.h)
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3dzerooutlevelseteqnrhslocal_(real_t* rhs, LSMHIP_BOXP(rg), const int* ix, const int* iy,
                                                    const int* iz, const int* nlo, const int* nhi) {
  LSMHIP_REQUIRE_ARRAYS(rhs, ix, iy, iz);
  const char* who = "lsm3dzerooutlevelseteqnrhslocal_";
  launch_points(who, LIST, no_mask(), ZeroOp{cv(rhs, LSMHIP_BOX(rg))});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3daddadvectiontermtolserhslocal_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* px,
                                                       const real_t* py, const real_t* pz, LSMHIP_BOXP(g),
                                                       const real_t* vx, const real_t* vy, const real_t* vz,
                                                       LSMHIP_BOXP(vg), LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, pz, vx, vy, vz, ix, iy, iz, nb);
  const char* who = "lsm3daddadvectiontermtolserhslocal_";
  const Box g = LSMHIP_BOX(g), vg = LSMHIP_BOX(vg);
  launch_points(who, LIST, FBMASK,
                AdvectOp{cv(rhs, LSMHIP_BOX(rg)), cv(px, g), cv(py, g), cv(pz, g), cv(vx, vg), cv(vy, vg), cv(vz, vg)});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3daddnormalveltermtolserhslocal_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp,
                                                       const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                       const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                       LSMHIP_BOXP(gm), const real_t* vn, LSMHIP_BOXP(vg),
                                                       LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm, vn, ix, iy, iz, nb);
  const char* who = "lsm3daddnormalveltermtolserhslocal_";
  launch_points(who, LIST, FBMASK,
                NormalOp{cv(rhs, LSMHIP_BOX(rg)), six(pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm)),
                         cv(vn, LSMHIP_BOX(vg)), real_t(0)});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3daddconstnormalveltermtolserhslocal_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp,
                                                            const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                            const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                            LSMHIP_BOXP(gm), const real_t* vn, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm, ix, iy, iz, nb);
  const char* who = "lsm3daddconstnormalveltermtolserhslocal_";
  Box none{0, -1, 0, -1, 0, -1};
  launch_points(who, LIST, FBMASK,
                NormalOp{cv(rhs, LSMHIP_BOX(rg)), six(pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm)),
                         cv((const real_t*)nullptr, none), *vn});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3daddconstcurvtermtolserhslocal_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* px,
                                                       const real_t* py, const real_t* pz, LSMHIP_BOXP(g),
                                                       const real_t* xx, const real_t* xy, const real_t* xz,
                                                       const real_t* yy, const real_t* yz, const real_t* zz,
                                                       LSMHIP_BOXP(g2), const real_t* b, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, pz, xx, xy, xz, yy, yz, zz, ix, iy, iz, nb);
  const char* who = "lsm3daddconstcurvtermtolserhslocal_";
  const Box g = LSMHIP_BOX(g), g2 = LSMHIP_BOX(g2);
  launch_points(who, LIST, FBMASK,
                ConstCurvOp{cv(rhs, LSMHIP_BOX(rg)), cv(px, g), cv(py, g), cv(pz, g), cv(xx, g2), cv(xy, g2),
                            cv(xz, g2), cv(yy, g2), cv(yz, g2), cv(zz, g2), *b});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3daddconstprecomputedcurvtermtolserhslocal_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* kappa,
                                                                  LSMHIP_BOXP(kg), const real_t* gmag,
                                                                  LSMHIP_BOXP(gg), const real_t* b, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, kappa, gmag, ix, iy, iz, nb);
  const char* who = "lsm3daddconstprecomputedcurvtermtolserhslocal_";
  launch_points(who, LIST, FBMASK,
                PrecompCurvOp{cv(rhs, LSMHIP_BOX(rg)), cv(kappa, LSMHIP_BOX(kg)), cv(gmag, LSMHIP_BOX(gg)), *b});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3daddexternalandnormalveltermtolserhslocal_(
    real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp, const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
    const real_t* pxm, const real_t* pym, const real_t* pzm, LSMHIP_BOXP(gm), const real_t* vn, const real_t* vx,
    const real_t* vy, const real_t* vz, LSMHIP_BOXP(vg), LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm, vn, vx, vy, vz, ix, iy, iz, nb);
  const char* who = "lsm3daddexternalandnormalveltermtolserhslocal_";
  const Box vg = LSMHIP_BOX(vg);
  launch_points(who, LIST, FBMASK,
                ExtNormalOp{cv(rhs, LSMHIP_BOX(rg)),
                            six(pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm)), cv(vn, vg), cv(vx, vg),
                            cv(vy, vg), cv(vz, vg)});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// TVD Runge-Kutta (src/toolbox/lsm_tvd_runge_kutta3d_local.h)
// ---------------------------------------------------------------------------
template <int KIND>
static void rk_local(const char* who, real_t* out, const Box& ob, const real_t* us, const Box& sb, const real_t* uc,
                     const Box& cb, const real_t* rhs, const Box& rb, real_t dt, const IndexList& L, const Mask& M) {
  launch_points(who, L, M, RkOp<KIND>{cv(out, ob), cv(us, sb), cv(uc, cb), cv(rhs, rb), dt});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3drk1steplocal_(real_t* un, LSMHIP_BOXP(ng), const real_t* uc, LSMHIP_BOXP(cg),
                                      const real_t* rhs, LSMHIP_BOXP(rg), const real_t* dt, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs, ix, iy, iz, nb);
  rk_local<1>("lsm3drk1steplocal_", un, LSMHIP_BOX(ng), uc, LSMHIP_BOX(cg), uc, LSMHIP_BOX(cg), rhs, LSMHIP_BOX(rg),
              *dt, LIST, FBMASK);
}
LSMHIP_EXPORT void lsm3dtvdrk2stage1local_(real_t* un, LSMHIP_BOXP(ng), const real_t* uc, LSMHIP_BOXP(cg),
                                           const real_t* rhs, LSMHIP_BOXP(rg), const real_t* dt, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs, ix, iy, iz, nb);
  rk_local<1>("lsm3dtvdrk2stage1local_", un, LSMHIP_BOX(ng), uc, LSMHIP_BOX(cg), uc, LSMHIP_BOX(cg), rhs,
              LSMHIP_BOX(rg), *dt, LIST, FBMASK);
}
LSMHIP_EXPORT void lsm3dtvdrk2stage2local_(real_t* un, LSMHIP_BOXP(ng), const real_t* u1, LSMHIP_BOXP(g1),
                                           const real_t* uc, LSMHIP_BOXP(cg), const real_t* rhs, LSMHIP_BOXP(rg),
                                           const real_t* dt, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, u1, uc, rhs, ix, iy, iz, nb);
  rk_local<2>("lsm3dtvdrk2stage2local_", un, LSMHIP_BOX(ng), u1, LSMHIP_BOX(g1), uc, LSMHIP_BOX(cg), rhs,
              LSMHIP_BOX(rg), *dt, LIST, FBMASK);
}
LSMHIP_EXPORT void lsm3dtvdrk3stage1local_(real_t* un, LSMHIP_BOXP(ng), const real_t* uc, LSMHIP_BOXP(cg),
                                           const real_t* rhs, LSMHIP_BOXP(rg), const real_t* dt, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs, ix, iy, iz, nb);
  rk_local<1>("lsm3dtvdrk3stage1local_", un, LSMHIP_BOX(ng), uc, LSMHIP_BOX(cg), uc, LSMHIP_BOX(cg), rhs,
              LSMHIP_BOX(rg), *dt, LIST, FBMASK);
}
LSMHIP_EXPORT void lsm3dtvdrk3stage2local_(real_t* u2, LSMHIP_BOXP(g2), const real_t* u1, LSMHIP_BOXP(g1),
                                           const real_t* uc, LSMHIP_BOXP(cg), const real_t* rhs, LSMHIP_BOXP(rg),
                                           const real_t* dt, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(u2, u1, uc, rhs, ix, iy, iz, nb);
  rk_local<32>("lsm3dtvdrk3stage2local_", u2, LSMHIP_BOX(g2), u1, LSMHIP_BOX(g1), uc, LSMHIP_BOX(cg), rhs,
               LSMHIP_BOX(rg), *dt, LIST, FBMASK);
}
LSMHIP_EXPORT void lsm3dtvdrk3stage3local_(real_t* un, LSMHIP_BOXP(ng), const real_t* u2, LSMHIP_BOXP(g2),
                                           const real_t* uc, LSMHIP_BOXP(cg), const real_t* rhs, LSMHIP_BOXP(rg),
                                           const real_t* dt, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(un, u2, uc, rhs, ix, iy, iz, nb);
  rk_local<33>("lsm3dtvdrk3stage3local_", un, LSMHIP_BOX(ng), u2, LSMHIP_BOX(g2), uc, LSMHIP_BOX(cg), rhs,
               LSMHIP_BOX(rg), *dt, LIST, FBMASK);
}

// ---------------------------------------------------------------------------
// Reinitialization, curvature
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3dcomputereinitializationeqnrhslocal_(
    real_t* out, LSMHIP_BOXP(rg), const real_t* phi, LSMHIP_BOXP(pg), const real_t* phi0, LSMHIP_BOXP(p0g),
    const real_t* pxp, const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp), const real_t* pxm, const real_t* pym,
    const real_t* pzm, LSMHIP_BOXP(gm), const real_t* dx, const real_t* dy, const real_t* dz,
    const int* use_phi0_for_sgn, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(out, phi, pxp, pyp, pzp, pxm, pym, pzm, ix, iy, iz, nb);
  if (*use_phi0_for_sgn == 1) LSMHIP_REQUIRE_ARRAYS(phi0);
  const char* who = "lsm3dcomputereinitializationeqnrhslocal_";
  real_t h = *dx > *dy ? *dx : *dy;
  h = h > *dz ? h : *dz;
  launch_points(who, LIST, FBMASK,
                ReinitOp{cv(out, LSMHIP_BOX(rg)), cv(phi, LSMHIP_BOX(pg)), cv(phi0, LSMHIP_BOX(p0g)),
                         six(pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm)), h * h,
                         *use_phi0_for_sgn == 1});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dcomputemeancurvatureorder2local_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                         LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                         const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                         const real_t* dx, const real_t* dy, const real_t* dz,
                                                         LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag, ix, iy, iz, nb);
  const char* who = "lsm3dcomputemeancurvatureorder2local_";
  const Box gg = LSMHIP_BOX(gg);
  launch_points(who, LIST, FBMASK,
                MeanCurvOp{cv(kappa, LSMHIP_BOX(kg)), cv(phi, LSMHIP_BOX(pg)), cv(px, gg), cv(py, gg), cv(pz, gg),
                           cv(gmag, gg), real_t(1.0) / *dx / *dx, real_t(1.0) / *dy / *dy, real_t(1.0) / *dz / *dz,
                           real_t(0.25) / *dx / *dy, real_t(0.25) / *dx / *dz, real_t(0.25) / *dy / *dz});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dcomputegaussiancurvatureorder2local_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                             LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                             const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                             const real_t* dx, const real_t* dy, const real_t* dz,
                                                             LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag, ix, iy, iz, nb);
  const char* who = "lsm3dcomputegaussiancurvatureorder2local_";
  const Box gg = LSMHIP_BOX(gg);
  launch_points(who, LIST, FBMASK,
                GaussCurv2Op{cv(kappa, LSMHIP_BOX(kg)), cv(phi, LSMHIP_BOX(pg)), cv(px, gg), cv(py, gg), cv(pz, gg),
                             cv(gmag, gg), real_t(1.0) / *dx / *dx, real_t(1.0) / *dy / *dy, real_t(1.0) / *dz / *dz,
                             real_t(0.25) / *dx / *dy, real_t(0.25) / *dx / *dz, real_t(0.25) / *dy / *dz});
  finish_local(who);
}

namespace {
void curv4_local(const char* who, real_t* kappa, const Box& kg, const real_t* phi, const Box& pg, const real_t* px,
                 const real_t* py, const real_t* pz, const real_t* gmag, const Box& gg, real_t dx, real_t dy,
                 real_t dz, const IndexList& L, const Mask& M) {
  const real_t c12 = real_t(0.0833333333333333333333);
  launch_points(who, L, M,
                Curv4Op{cv(kappa, kg), cv(phi, pg), cv(px, gg), cv(py, gg), cv(pz, gg), cv(gmag, gg), c12 / dx / dx,
                        c12 / dy / dy, c12 / dz / dz, c12 / dx, c12 / dy, c12 / dz});
  finish_local(who);
}
}  // namespace

// The reference's lsm3dComputeMeanCurvatureOrder4Local computes the Gaussian
// curvature of lsm3dComputeGaussianCurvatureOrder4Local (identical bodies,
// lsm_curvature3d_local.f.in:352-418 and :499-565); kept as it is.
LSMHIP_EXPORT void lsm3dcomputemeancurvatureorder4local_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                         LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                         const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                         const real_t* dx, const real_t* dy, const real_t* dz,
                                                         LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag, ix, iy, iz, nb);
  curv4_local("lsm3dcomputemeancurvatureorder4local_", kappa, LSMHIP_BOX(kg), phi, LSMHIP_BOX(pg), px, py, pz, gmag,
              LSMHIP_BOX(gg), *dx, *dy, *dz, LIST, FBMASK);
}
LSMHIP_EXPORT void lsm3dcomputegaussiancurvatureorder4local_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                             LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                             const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                             const real_t* dx, const real_t* dy, const real_t* dz,
                                                             LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag, ix, iy, iz, nb);
  curv4_local("lsm3dcomputegaussiancurvatureorder4local_", kappa, LSMHIP_BOX(kg), phi, LSMHIP_BOX(pg), px, py, pz,
              gmag, LSMHIP_BOX(gg), *dx, *dy, *dz, LIST, FBMASK);
}

LSMHIP_EXPORT void lsm3dcomputeorthogonalizationeqnrhslocal_(
    real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp, const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
    const real_t* pxm, const real_t* pym, const real_t* pzm, LSMHIP_BOXP(gm), const real_t* psi, LSMHIP_BOXP(sg),
    const real_t* qxp, const real_t* qyp, const real_t* qzp, LSMHIP_BOXP(qp), const real_t* qxm, const real_t* qym,
    const real_t* qzm, LSMHIP_BOXP(qm), const real_t* dx, const real_t* dy, const real_t* dz, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm, psi, qxp, qyp, qzp, qxm, qym, qzm, ix, iy, iz, nb);
  const char* who = "lsm3dcomputeorthogonalizationeqnrhslocal_";
  const Box gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm), qp = LSMHIP_BOX(qp), qm = LSMHIP_BOX(qm);
  real_t h = *dx > *dy ? *dx : *dy;
  h = h > *dz ? h : *dz;
  launch_points(who, LIST, FBMASK,
                OrthoOp{cv(rhs, LSMHIP_BOX(rg)), cv(pxp, gp), cv(pyp, gp), cv(pzp, gp), cv(pxm, gm), cv(pym, gm),
                        cv(pzm, gm), cv(psi, LSMHIP_BOX(sg)), cv(qxp, qp), cv(qyp, qp), cv(qzp, qp), cv(qxm, qm),
                        cv(qym, qm), cv(qzm, qm), h * h});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// Mask / copy / cut-off (src/toolbox/lsm_localization3d.h)
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3dimposemasklocal_(real_t* dst, LSMHIP_BOXP(dg), const real_t* src, LSMHIP_BOXP(sg),
                                         const real_t* msk, LSMHIP_BOXP(mg), const int* ix, const int* iy,
                                         const int* iz, const int* nlo, const int* nhi) {
  LSMHIP_REQUIRE_ARRAYS(dst, src, msk, ix, iy, iz);
  const char* who = "lsm3dimposemasklocal_";
  launch_points(who, LIST, no_mask(),
                ImposeMaskOp{cv(dst, LSMHIP_BOX(dg)), cv(src, LSMHIP_BOX(sg)), cv(msk, LSMHIP_BOX(mg))});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dcopydatalocal_(real_t* dst, LSMHIP_BOXP(dg), const real_t* src, LSMHIP_BOXP(sg),
                                       const int* ix, const int* iy, const int* iz, const int* nlo, const int* nhi) {
  LSMHIP_REQUIRE_ARRAYS(dst, src, ix, iy, iz);
  const char* who = "lsm3dcopydatalocal_";
  launch_points(who, LIST, no_mask(), CopyOp{cv(dst, LSMHIP_BOX(dg)), cv(src, LSMHIP_BOX(sg))});
  finish_local(who);
}

LSMHIP_EXPORT void lsm3dmultiplycutofflserhslocal_(const real_t* phi, real_t* rhs, LSMHIP_BOXP(g), const int* ix,
                                                   const int* iy, const int* iz, const int* nlo, const int* nhi,
                                                   const unsigned char* nb, LSMHIP_BOXP(nbb),
                                                   const unsigned char* mark_fb, const real_t* beta,
                                                   const real_t* gamma) {
  LSMHIP_REQUIRE_ARRAYS(phi, rhs, ix, iy, iz, nb);
  const char* who = "lsm3dmultiplycutofflserhslocal_";
  const Box g = LSMHIP_BOX(g);
  const real_t c1 = *gamma - real_t(3) * *beta;
  real_t c2 = *gamma - *beta;
  c2 = c2 * c2 * c2;
  launch_points(who, LIST, FBMASK, CutOffOp{cv(phi, g), cv(rhs, g), *beta, *gamma, c1, c2});
  finish_local(who);
}

// ---------------------------------------------------------------------------
// Reductions (src/toolbox/lsm_math_utils3d_local.h)
// ---------------------------------------------------------------------------
namespace {
struct MaxNormDiffOp {
  CV<const real_t> a, b;
  __device__ bool operator()(int i, int j, int k, double& v) const {
    if (!a.has(i, j, k) || !b.has(i, j, k)) return false;
    v = (double)fabs(a(i, j, k) - b(i, j, k));
    return true;
  }
};
struct AdvDtOp {
  CV<const real_t> u, v, w;
  real_t ix, iy, iz;
  __device__ bool operator()(int i, int j, int k, double& r) const {
    if (!u.has(i, j, k)) return false;
    r = (double)(fabs(u(i, j, k)) * ix + fabs(v(i, j, k)) * iy + fabs(w(i, j, k)) * iz);
    return true;
  }
};
// skip_small: the variable-speed local routine only admits points whose
// |grad phi| exceeds zero_tol (lsm_math_utils3d_local.f.in:405-416).
struct NormalDtOp {
  Six g;
  CV<const real_t> vn;
  real_t vc, mdx2, ix, iy, iz;
  int skip_small;
  __device__ bool operator()(int i, int j, int k, double& r) const {
    if (!g.has(i, j, k) || (vn.v.p && !vn.has(i, j, k))) return false;
    const real_t av = vn.v.p ? fabs(vn(i, j, k)) : vc;
    real_t n;
    const real_t h = stencil::normal_h_over_dx(av, g.xp(i, j, k), g.yp(i, j, k), g.zp(i, j, k), g.xm(i, j, k),
                                               g.ym(i, j, k), g.zm(i, j, k), mdx2, ix, iy, iz, &n);
    r = (skip_small && !(n > kZeroTol)) ? -1.0 : (double)h;
    return true;
  }
};
}  // namespace

LSMHIP_EXPORT void lsm3dmaxnormdifflocal_(real_t* out, const real_t* f1, LSMHIP_BOXP(g1), const real_t* f2,
                                          LSMHIP_BOXP(g2), LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(f1, f2, ix, iy, iz, nb);
  const char* who = "lsm3dmaxnormdifflocal_";
  double m = 0.0;
  bool any;
  if (!points_max(who, LIST, FBMASK, MaxNormDiffOp{cv(f1, LSMHIP_BOX(g1)), cv(f2, LSMHIP_BOX(g2))}, &m, &any)) return;
  *out = any ? real_t(m) : real_t(0);
}

LSMHIP_EXPORT void lsm3dcomputestableadvectiondtlocal_(real_t* dt, const real_t* vx, const real_t* vy,
                                                       const real_t* vz, LSMHIP_BOXP(vg), const real_t* dx,
                                                       const real_t* dy, const real_t* dz, const real_t* cfl,
                                                       LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(vx, vy, vz, ix, iy, iz, nb);
  const char* who = "lsm3dcomputestableadvectiondtlocal_";
  const Box vg = LSMHIP_BOX(vg);
  double m = -1.0;
  bool any;
  if (!points_max(who, LIST, FBMASK,
                  AdvDtOp{cv(vx, vg), cv(vy, vg), cv(vz, vg), real_t(1.0) / *dx, real_t(1.0) / *dy,
                          real_t(1.0) / *dz},
                  &m, &any))
    return;
  if (!any) m = -1.0;
  *dt = *cfl / (real_t(m) + kTiny);
}

static void normal_dt_local(const char* who, real_t* dt, const real_t* vn, const Box& vg, real_t vc,
                            const real_t* pxp, const real_t* pyp, const real_t* pzp, const Box& gp, const real_t* pxm,
                            const real_t* pym, const real_t* pzm, const Box& gm, real_t dx, real_t dy, real_t dz,
                            real_t cfl, const IndexList& L, const Mask& M, int skip_small) {
  real_t h = dx > dy ? dx : dy;
  h = h > dz ? h : dz;
  double m = -1.0;
  bool any;
  NormalDtOp op{six(pxp, pyp, pzp, gp, pxm, pym, pzm, gm), cv(vn, vg), vc, h * h, real_t(1.0) / dx,
                real_t(1.0) / dy, real_t(1.0) / dz, skip_small};
  if (!points_max(who, L, M, op, &m, &any)) return;
  if (!any) m = -1.0;
  *dt = cfl / (real_t(m) + kTiny);
}

LSMHIP_EXPORT void lsm3dcomputestablenormalveldtlocal_(real_t* dt, const real_t* vn, LSMHIP_BOXP(vg),
                                                       const real_t* pxp, const real_t* pyp, const real_t* pzp,
                                                       LSMHIP_BOXP(gp), const real_t* pxm, const real_t* pym,
                                                       const real_t* pzm, LSMHIP_BOXP(gm), const real_t* dx,
                                                       const real_t* dy, const real_t* dz, const real_t* cfl,
                                                       LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(vn, pxp, pyp, pzp, pxm, pym, pzm, ix, iy, iz, nb);
  normal_dt_local("lsm3dcomputestablenormalveldtlocal_", dt, vn, LSMHIP_BOX(vg), real_t(0), pxp, pyp, pzp,
                  LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm), *dx, *dy, *dz, *cfl, LIST, FBMASK, 1);
}

// The reference reads two uninitialised locals here
// (src/toolbox/lsm_math_utils3d_local.f.in:628-751); this computes the
// evident intent, the whole-grid formula of lsm_math_utils3d.f.in:519-564.
LSMHIP_EXPORT void lsm3dcomputestableconstnormalveldtlocal_(real_t* dt, const real_t* vn, const real_t* pxp,
                                                            const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                            const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                            LSMHIP_BOXP(gm), const real_t* dx, const real_t* dy,
                                                            const real_t* dz, const real_t* cfl, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, ix, iy, iz, nb);
  Box none{0, -1, 0, -1, 0, -1};
  normal_dt_local("lsm3dcomputestableconstnormalveldtlocal_", dt, nullptr, none, real_t(fabs(*vn)), pxp, pyp, pzp,
                  LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm), *dx, *dy, *dz, *cfl, LIST, FBMASK, 0);
}

namespace {
struct GradNormOp {
  CV<const real_t> F;
  real_t fx, fy, fz;
  __device__ bool operator()(int i, int j, int k, double& v) const {
    if (!F.star(i, j, k, 1)) return false;
    const real_t x = (F(i + 1, j, k) - F(i - 1, j, k)) * fx;
    const real_t y = (F(i, j + 1, k) - F(i, j - 1, k)) * fy;
    const real_t z = (F(i, j, k + 1) - F(i, j, k - 1)) * fz;
    v = (double)sqrt(x * x + y * y + z * z);
    return true;
  }
};
}  // namespace

// lsm3dComputeAveGradPhiLocal (lsm_spatial_derivatives3d_local.f.in:1089-1167):
// mean of the central-difference |grad phi| over admissible points.  The
// reference sums serially; here the sum is FP64 compensated in a fixed order.
LSMHIP_EXPORT void lsm3dcomputeavegradphilocal_(real_t* ave, const real_t* phi, LSMHIP_BOXP(pg), const real_t* dx,
                                                const real_t* dy, const real_t* dz, LOCAL_TAIL) {
  LSMHIP_REQUIRE_ARRAYS(phi, ix, iy, iz, nb);
  points_mean("lsm3dcomputeavegradphilocal_", LIST, FBMASK,
              GradNormOp{cv(phi, LSMHIP_BOX(pg)), real_t(0.5) / *dx, real_t(0.5) / *dy, real_t(0.5) / *dz}, ave);
}
