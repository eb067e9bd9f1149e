// geometry3d.hip -- the kernels either side of the curvature / reinitialization
// path (whole grid): unit and signed unit normals, mean curvature from a
// precomputed Hessian, Gaussian curvature (second and fourth order),
// fourth-order mean curvature, the orthogonalization RHS and the field-extension
// RHS.
// References: src/geometry/lsm_geometry3d_fort.f.in, src/geometry/lsm_curvature3d.f.in,
// src/reinitialization/lsm_reinitialization3d.f.in, src/field_extension/lsm_field_extension3d.f.
//
// One thread per cell of the fill box; every kernel reads each input once per
// cell (stencil neighbours come from L1/L2), so all of them are bound by HBM
// bandwidth.  In FP64 the operation order is the reference's, without
// contraction, and every result is bitwise the reference's except the mean
// curvature's |grad phi|^3 = (|grad phi|^2)^1.5 (see pow_three_halves).
#include "lsmhip_core.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

bool inside(const char* who, const Box& fb, std::initializer_list<std::pair<Box, const char*>> arrays) {
  for (auto& a : arrays)
    if (!check_box(a.first, fb, who, a.second) || !check_size(a.first, who)) return false;
  return true;
}

// An array read at offsets -w..w along each axis separately (a cross-shaped
// stencil): each widened box must lie inside the array's ghost box.
bool cross_inside(const char* who, const Box& fb, const Box& g, int w, const char* what) {
  return check_box(g, fb.widen(w, 0, 0), who, what) && check_box(g, fb.widen(0, w, 0), who, what) &&
         check_box(g, fb.widen(0, 0, w), who, what) && check_size(g, who);
}

real_t max3(real_t a, real_t b, real_t c) {
  const real_t m = a > b ? a : b;
  return m > c ? m : c;
}

// x^1.5 for x >= 1e-11.  FP64: the exact product x * sqrt(x) carried as a
// double-double (sqrt's remainder x - s^2 and the product's rounding error are
// both exact with FMA) and rounded once, i.e. correctly rounded except within
// about 2^-100 of a rounding boundary -- the value a correctly rounded pow(x,
// 1.5) returns, which is what the C library's pow gives the reference's
// compiled Fortran.  FP32: x * sqrt(x).
__device__ __forceinline__ real_t pow_three_halves(real_t x) {
  if constexpr (kDouble) {
    const double s = sqrt(x);
    const double p = x * s;
    if (__builtin_isinf(p)) return p;
    const double e = fma(x, s, -p);       // x s - p
    const double r = fma(-s, s, x);       // x - s^2
    const double t = r / (2.0 * s);       // sqrt(x) - s to second order
    return p + fma(x, t, e);
  } else {
    return x * sqrt(x);
  }
}

// ---------------------------------------------------------------------------
// Normals, lsm_geometry3d_fort.f.in:27-272
// ---------------------------------------------------------------------------
struct Vec3Out {
  View<real_t> x, y, z;
};
struct Vec3In {
  View<const real_t> x, y, z;
};

__global__ void __launch_bounds__(256) unit_normal_kernel(Box fb, Vec3Out n, Vec3In g) {
  LSMHIP_CELL(fb);
  const real_t x = g.x(i, j, k), y = g.y(i, j, k), z = g.z(i, j, k);
  const real_t norm = sqrt(x * x + y * y + z * z);
  real_t ox = real_t(1), oy = real_t(0), oz = real_t(0);
  if (norm >= kZeroTol) {
    const real_t inv = real_t(1) / norm;
    ox = x * inv;
    oy = y * inv;
    oz = z * inv;
  }
  n.x(i, j, k) = ox;
  n.y(i, j, k) = oy;
  n.z(i, j, k) = oz;
}

__global__ void __launch_bounds__(256) signed_unit_normal_kernel(Box fb, Vec3Out n, Vec3In g,
                                                                 View<const real_t> phi, real_t dx_sq) {
  LSMHIP_CELL(fb);
  const real_t c = phi(i, j, k);
  real_t ox = real_t(0), oy = real_t(0), oz = real_t(0);
  if (fabs(c) > kZeroTol) {
    const real_t x = g.x(i, j, k), y = g.y(i, j, k), z = g.z(i, j, k);
    const real_t ngs = x * x + y * y + z * z;
    if (ngs >= kZeroTol) {
      const real_t sgn = c / sqrt(c * c + ngs * dx_sq);
      const real_t inv = real_t(1) / sqrt(ngs);
      ox = sgn * x * inv;
      oy = sgn * y * inv;
      oz = sgn * z * inv;
    } else {
      ox = real_t(1);
    }
  }
  n.x(i, j, k) = ox;
  n.y(i, j, k) = oy;
  n.z(i, j, k) = oz;
}

// ---------------------------------------------------------------------------
// Curvatures
// ---------------------------------------------------------------------------
struct Hessian {
  View<const real_t> xx, xy, xz, yy, yz, zz;
};

// lsm_geometry3d_fort.f.in:274-394 -- numerator terms in that routine's order.
__global__ void __launch_bounds__(256) mean_curvature_hessian_kernel(Box fb, View<real_t> kappa, Vec3In g,
                                                                     Hessian h) {
  LSMHIP_CELL(fb);
  const real_t x = g.x(i, j, k), y = g.y(i, j, k), z = g.z(i, j, k);
  const real_t ngs = x * x + y * y + z * z;
  real_t out = real_t(0);
  if (!(ngs < kZeroTol)) {
    const real_t xx = h.xx(i, j, k), xy = h.xy(i, j, k), xz = h.xz(i, j, k), yy = h.yy(i, j, k),
                 yz = h.yz(i, j, k), zz = h.zz(i, j, k);
    const real_t num = xx * y * y - real_t(2) * xy * x * y + yy * x * x + xx * z * z - real_t(2) * xz * x * z +
                       zz * x * x + yy * z * z - real_t(2) * yz * y * z + zz * y * y;
    out = num / pow_three_halves(ngs);
  }
  kappa(i, j, k) = out;
}

// Gaussian numerator, lsm_curvature3d.f.in:230-237 (and :490-497 with the x-z
// mixed derivative taken from phi_z along x).
__device__ __forceinline__ real_t gauss_num(real_t x, real_t y, real_t z, real_t xx, real_t yy, real_t zz, real_t xy,
                                            real_t xz, real_t yz) {
  return x * x * (yy * zz - yz * yz) + y * y * (xx * zz - xz * xz) + z * z * (xx * yy - xy * xy) +
         real_t(2) * (x * y * (xz * yz - xy * zz) + y * z * (xy * xz - yz * xx) + x * z * (xy * yz - xz * yy));
}

struct Factors2 {
  real_t xx, yy, zz, xy, xz, yz;
};

// lsm_curvature3d.f.in:136-248: second differences of phi (2-point cross
// stencil for the mixed terms), numerator over |grad phi|^4.
__global__ void __launch_bounds__(256) gaussian_order2_kernel(Box fb, View<real_t> kappa, View<const real_t> F,
                                                              Vec3In g, View<const real_t> gmag, Factors2 f) {
  LSMHIP_CELL(fb);
  const real_t c = F(i, j, k);
  const real_t xx = (F(i + 1, j, k) - real_t(2) * c + F(i - 1, j, k)) * f.xx;
  const real_t yy = (F(i, j + 1, k) - real_t(2) * c + F(i, j - 1, k)) * f.yy;
  const real_t zz = (F(i, j, k + 1) - real_t(2) * c + F(i, j, k - 1)) * f.zz;
  const real_t xy = (F(i + 1, j + 1, k) - F(i + 1, j - 1, k) - F(i - 1, j + 1, k) + F(i - 1, j - 1, k)) * f.xy;
  const real_t xz = (F(i + 1, j, k + 1) - F(i + 1, j, k - 1) - F(i - 1, j, k + 1) + F(i - 1, j, k - 1)) * f.xz;
  const real_t yz = (F(i, j + 1, k + 1) - F(i, j + 1, k - 1) - F(i, j - 1, k + 1) + F(i, j - 1, k - 1)) * f.yz;
  const real_t gm = gmag(i, j, k);
  const real_t den = gm * gm * gm * gm;
  real_t out = real_t(0);
  if (!(den < kZeroTol)) out = gauss_num(g.x(i, j, k), g.y(i, j, k), g.z(i, j, k), xx, yy, zz, xy, xz, yz) / den;
  kappa(i, j, k) = out;
}

struct Factors4 {
  real_t xx, yy, zz;  // 1/(12 h^2)
  real_t x, y, z;     // 1/(12 h)
};

// lsm_curvature3d.f.in:253-507.  Fourth-order second differences of phi
// (:330-341), mixed derivatives as fourth-order differences of the given first
// derivatives (:342-351): phi_xy = d/dy phi_x, phi_yz = d/dz phi_y,
// phi_zx = d/dx phi_z.  GAUSS = 0: mean curvature over |grad phi|^3 (numerator
// order of :358-366); GAUSS = 1: Gaussian curvature over |grad phi|^4.
template <int GAUSS>
__global__ void __launch_bounds__(256) curvature_order4_kernel(Box fb, View<real_t> kappa, View<const real_t> F,
                                                               Vec3In g, View<const real_t> gmag, Factors4 f) {
  LSMHIP_CELL(fb);
  const real_t c = F(i, j, k);
  const real_t xx = (-F(i + 2, j, k) + real_t(16) * F(i + 1, j, k) - real_t(30) * c - F(i - 2, j, k) +
                     real_t(16) * F(i - 1, j, k)) * f.xx;
  const real_t yy = (-F(i, j + 2, k) + real_t(16) * F(i, j + 1, k) - real_t(30) * c - F(i, j - 2, k) +
                     real_t(16) * F(i, j - 1, k)) * f.yy;
  const real_t zz = (-F(i, j, k + 2) + real_t(16) * F(i, j, k + 1) - real_t(30) * c - F(i, j, k - 2) +
                     real_t(16) * F(i, j, k - 1)) * f.zz;
  const real_t xy = (-g.x(i, j + 2, k) + real_t(8) * g.x(i, j + 1, k) + g.x(i, j - 2, k) -
                     real_t(8) * g.x(i, j - 1, k)) * f.y;
  const real_t yz = (-g.y(i, j, k + 2) + real_t(8) * g.y(i, j, k + 1) + g.y(i, j, k - 2) -
                     real_t(8) * g.y(i, j, k - 1)) * f.z;
  const real_t zx = (-g.z(i + 2, j, k) + real_t(8) * g.z(i + 1, j, k) + g.z(i - 2, j, k) -
                     real_t(8) * g.z(i - 1, j, k)) * f.x;
  const real_t gm = gmag(i, j, k);
  const real_t den = GAUSS ? gm * gm * gm * gm : gm * gm * gm;
  real_t out = real_t(0);
  if (!(den < kZeroTol)) {
    const real_t x = g.x(i, j, k), y = g.y(i, j, k), z = g.z(i, j, k);
    const real_t num = GAUSS ? gauss_num(x, y, z, xx, yy, zz, xy, zx, yz)
                             : stencil::curv_num(xx, xy, zx, yy, yz, zz, x, y, z);
    out = num / den;
  }
  kappa(i, j, k) = out;
}

// ---------------------------------------------------------------------------
// Orthogonalization RHS, lsm_reinitialization3d.f.in:296-476
// ---------------------------------------------------------------------------
struct Sided {
  Vec3In plus, minus;
};

__global__ void __launch_bounds__(256) orthogonalization_kernel(Box fb, View<real_t> rhs, Sided phi_d,
                                                                View<const real_t> psi, Sided psi_d,
                                                                real_t dx_sq) {
  LSMHIP_CELL(fb);
  const real_t g1 = real_t(0.5) * (psi_d.plus.x(i, j, k) + psi_d.minus.x(i, j, k));
  const real_t g2 = real_t(0.5) * (psi_d.plus.y(i, j, k) + psi_d.minus.y(i, j, k));
  const real_t g3 = real_t(0.5) * (psi_d.plus.z(i, j, k) + psi_d.minus.z(i, j, k));
  const real_t ng = sqrt(g1 * g1 + g2 * g2 + g3 * g3);
  const real_t s = psi(i, j, k);
  real_t out = real_t(0);
  if (fabs(s) > kZeroTol && ng > kZeroTol) {
    const real_t sgn = s / sqrt(s * s + dx_sq * (ng * ng));
    const real_t a1 = sgn * g1, a2 = sgn * g2, a3 = sgn * g3;
    const real_t z = real_t(0);
    out = real_t(-1) / ng *
          (maxr(a1, z) * phi_d.minus.x(i, j, k) + minr(a1, z) * phi_d.plus.x(i, j, k) +
           maxr(a2, z) * phi_d.minus.y(i, j, k) + minr(a2, z) * phi_d.plus.y(i, j, k) +
           maxr(a3, z) * phi_d.minus.z(i, j, k) + minr(a3, z) * phi_d.plus.z(i, j, k));
  }
  rhs(i, j, k) = out;
}

// ---------------------------------------------------------------------------
// Field extension RHS, lsm_field_extension3d.f:31-156
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) field_extension_kernel(Box fb, View<real_t> rhs, View<const real_t> P,
                                                              Vec3In grad_s, Vec3In normal, real_t cutoff) {
  LSMHIP_CELL(fb);
  const real_t c = P(i, j, k);
  const bool away = fabs(c) > cutoff ||
                    (c * P(i - 1, j, k) > real_t(0) && c * P(i + 1, j, k) > real_t(0) &&
                     c * P(i, j - 1, k) > real_t(0) && c * P(i, j + 1, k) > real_t(0) &&
                     c * P(i, j, k - 1) > real_t(0) && c * P(i, j, k + 1) > real_t(0));
  real_t out = real_t(0);
  if (away)
    out = -(normal.x(i, j, k) * grad_s.x(i, j, k) + normal.y(i, j, k) * grad_s.y(i, j, k) +
            normal.z(i, j, k) * grad_s.z(i, j, k));
  rhs(i, j, k) = out;
}

Vec3In vin(const real_t* x, const real_t* y, const real_t* z, const Box& b) {
  return Vec3In{view(x, b), view(y, b), view(z, b)};
}
Vec3Out vout(real_t* x, real_t* y, real_t* z, const Box& b) { return Vec3Out{view(x, b), view(y, b), view(z, b)}; }

}  // namespace

// ---------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3dcomputeunitnormal_(real_t* nx, real_t* ny, real_t* nz, LSMHIP_BOXP(ng), const real_t* px,
                                           const real_t* py, const real_t* pz, LSMHIP_BOXP(gg), LSMHIP_BOXP(fb)) {
  const char* who = "lsm3dcomputeunitnormal_";
  Box fb = LSMHIP_BOX(fb), ng = LSMHIP_BOX(ng), gg = LSMHIP_BOX(gg);
  if (fb.empty() || !inside(who, fb, {{ng, "normal"}, {gg, "grad_phi"}})) return;
  hipLaunchKernelGGL(unit_normal_kernel, grid_for(fb), block2d(), 0, stream(), fb, vout(nx, ny, nz, ng),
                     vin(px, py, pz, gg));
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputesignedunitnormal_(real_t* nx, real_t* ny, real_t* nz, LSMHIP_BOXP(ng),
                                                 const real_t* px, const real_t* py, const real_t* pz,
                                                 LSMHIP_BOXP(gg), const real_t* phi, LSMHIP_BOXP(pg),
                                                 LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                 const real_t* dz) {
  const char* who = "lsm3dcomputesignedunitnormal_";
  Box fb = LSMHIP_BOX(fb), ng = LSMHIP_BOX(ng), gg = LSMHIP_BOX(gg), pg = LSMHIP_BOX(pg);
  if (fb.empty() || !inside(who, fb, {{ng, "normal"}, {gg, "grad_phi"}, {pg, "phi"}})) return;
  const real_t h = max3(*dx, *dy, *dz);
  hipLaunchKernelGGL(signed_unit_normal_kernel, grid_for(fb), block2d(), 0, stream(), fb, vout(nx, ny, nz, ng),
                     vin(px, py, pz, gg), view(phi, pg), h * h);
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputemeancurvature_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* px, const real_t* py,
                                              const real_t* pz, LSMHIP_BOXP(gg), const real_t* pxx,
                                              const real_t* pxy, const real_t* pxz, const real_t* pyy,
                                              const real_t* pyz, const real_t* pzz, LSMHIP_BOXP(hg),
                                              LSMHIP_BOXP(fb)) {
  const char* who = "lsm3dcomputemeancurvature_";
  Box fb = LSMHIP_BOX(fb), kg = LSMHIP_BOX(kg), gg = LSMHIP_BOX(gg), hg = LSMHIP_BOX(hg);
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}, {gg, "grad_phi"}, {hg, "hessian_phi"}})) return;
  Hessian h{view(pxx, hg), view(pxy, hg), view(pxz, hg), view(pyy, hg), view(pyz, hg), view(pzz, hg)};
  hipLaunchKernelGGL(mean_curvature_hessian_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(kappa, kg),
                     vin(px, py, pz, gg), h);
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputegaussiancurvatureorder2_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                        LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                        const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                        LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                        const real_t* dz) {
  const char* who = "lsm3dcomputegaussiancurvatureorder2_";
  Box fb = LSMHIP_BOX(fb), kg = LSMHIP_BOX(kg), pg = LSMHIP_BOX(pg), gg = LSMHIP_BOX(gg);
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}, {gg, "grad_phi"}}) ||
      !check_box(pg, fb.widen(1, 1, 1), who, "phi") || !check_size(pg, who))
    return;
  Factors2 f{real_t(1.0) / *dx / *dx, real_t(1.0) / *dy / *dy, real_t(1.0) / *dz / *dz,
             real_t(0.25) / *dx / *dy, real_t(0.25) / *dx / *dz, real_t(0.25) / *dy / *dz};
  hipLaunchKernelGGL(gaussian_order2_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(kappa, kg),
                     view(phi, pg), vin(px, py, pz, gg), view(gmag, gg), f);
  finish(who);
}

template <int GAUSS>
static void curvature_order4(const char* who, real_t* kappa, const Box& kg, const real_t* phi, const Box& pg,
                             const real_t* px, const real_t* py, const real_t* pz, const real_t* gmag,
                             const Box& gg, const Box& fb, real_t dx, real_t dy, real_t dz) {
  if (fb.empty() || !inside(who, fb, {{kg, "kappa"}}) || !cross_inside(who, fb, gg, 2, "grad_phi") ||
      !cross_inside(who, fb, pg, 2, "phi"))
    return;
  const real_t c = real_t(0.0833333333333333333333);
  Factors4 f{c / dx / dx, c / dy / dy, c / dz / dz, c / dx, c / dy, c / dz};
  hipLaunchKernelGGL(curvature_order4_kernel<GAUSS>, grid_for(fb), block2d(), 0, stream(), fb, view(kappa, kg),
                     view(phi, pg), vin(px, py, pz, gg), view(gmag, gg), f);
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputemeancurvatureorder4_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                    LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                    const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                    LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                    const real_t* dz) {
  curvature_order4<0>("lsm3dcomputemeancurvatureorder4_", kappa, LSMHIP_BOX(kg), phi, LSMHIP_BOX(pg), px, py, pz,
                      gmag, LSMHIP_BOX(gg), LSMHIP_BOX(fb), *dx, *dy, *dz);
}

LSMHIP_EXPORT void lsm3dcomputegaussiancurvatureorder4_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                        LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                        const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                        LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                        const real_t* dz) {
  curvature_order4<1>("lsm3dcomputegaussiancurvatureorder4_", kappa, LSMHIP_BOX(kg), phi, LSMHIP_BOX(pg), px, py,
                      pz, gmag, LSMHIP_BOX(gg), LSMHIP_BOX(fb), *dx, *dy, *dz);
}

LSMHIP_EXPORT void lsm3dcomputeorthogonalizationeqnrhs_(
    real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp, const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
    const real_t* pxm, const real_t* pym, const real_t* pzm, LSMHIP_BOXP(gm), const real_t* psi, LSMHIP_BOXP(sg),
    const real_t* qxp, const real_t* qyp, const real_t* qzp, LSMHIP_BOXP(qp), const real_t* qxm, const real_t* qym,
    const real_t* qzm, LSMHIP_BOXP(qm), LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy, const real_t* dz) {
  const char* who = "lsm3dcomputeorthogonalizationeqnrhs_";
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm), sg = LSMHIP_BOX(sg),
      qp = LSMHIP_BOX(qp), qm = LSMHIP_BOX(qm);
  if (fb.empty() || !inside(who, fb,
                            {{rg, "ortho_rhs"},
                             {gp, "grad_phi_plus"},
                             {gm, "grad_phi_minus"},
                             {sg, "psi"},
                             {qp, "grad_psi_plus"},
                             {qm, "grad_psi_minus"}}))
    return;
  const real_t h = max3(*dx, *dy, *dz);
  Sided phi_d{vin(pxp, pyp, pzp, gp), vin(pxm, pym, pzm, gm)};
  Sided psi_d{vin(qxp, qyp, qzp, qp), vin(qxm, qym, qzm, qm)};
  hipLaunchKernelGGL(orthogonalization_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), phi_d,
                     view(psi, sg), psi_d, h * h);
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputefieldextensioneqnrhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* S, LSMHIP_BOXP(sg),
                                                     const real_t* phi, LSMHIP_BOXP(pg), const real_t* sx,
                                                     const real_t* sy, const real_t* sz, LSMHIP_BOXP(ug),
                                                     const real_t* nx, const real_t* ny, const real_t* nz,
                                                     LSMHIP_BOXP(ng), LSMHIP_BOXP(fb), const real_t* dx,
                                                     const real_t* dy, const real_t* dz) {
  const char* who = "lsm3dcomputefieldextensioneqnrhs_";
  (void)S;  // part of the reference signature; the RHS reads only its upwind gradient
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), pg = LSMHIP_BOX(pg), ug = LSMHIP_BOX(ug), ng = LSMHIP_BOX(ng);
  if (fb.empty() || !inside(who, fb, {{rg, "rhs"}, {ug, "grad_S_upwind"}, {ng, "signed_normal"}}) ||
      !cross_inside(who, fb, pg, 1, "phi"))
    return;
  const real_t cutoff = real_t(3.0) * max3(*dx, *dy, *dz);
  hipLaunchKernelGGL(field_extension_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(phi, pg),
                     vin(sx, sy, sz, ug), vin(nx, ny, nz, ng), cutoff);
  finish(who);
}
