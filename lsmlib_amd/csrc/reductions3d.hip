// reductions3d.hip -- CFL time steps, max-norm difference, smoothed volume and
// surface integrals (src/toolbox/lsm_math_utils3d.f.in,
// src/geometry/lsm_geometry3d_fort.f.in).
//
// Maxima are order-independent, so the CFL steps and the max-norm difference
// equal the reference's bitwise.  The reference accumulates integrals with a
// naive serial sum; here each cell's term is computed as the reference
// computes it and the terms are summed in FP64 with a fixed-order
// compensated (Neumaier) reduction: deterministic, and at least as accurate.
// Both precisions accumulate in FP64.
#include <mutex>

#include "lsmhip_core.hpp"
#include "reduce.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

constexpr double kPi = 3.14159265358979323846;
// Truncated 1/pi of the reference's Fortran Heaviside (lsm_geometry3d_fort.f.in:432-433).
constexpr double kOneOverPi = 0.31830988618379;

// Ops with branches on their loaded values are split: operator()(i, j, k) =
// eval(load(i, j, k), i, j, k), where load reads what every cell needs and
// eval may read more where the term needs it.  The two branch-free maxima
// below are faster called whole.
struct MaxNormDiff {
  View<const real_t> a, b;
  __device__ double operator()(int i, int j, int k) const { return (double)fabs(a(i, j, k) - b(i, j, k)); }
};

struct AdvectionCfl {
  View<const real_t> u, v, w;
  real_t ix, iy, iz;
  __device__ double operator()(int i, int j, int k) const {
    return (double)(fabs(u(i, j, k)) * ix + fabs(v(i, j, k)) * iy + fabs(w(i, j, k)) * iz);
  }
};

struct NormalCfl {
  View<const real_t> vn, xp, yp, zp, xm, ym, zm;
  real_t vconst, mdx2, ix, iy, iz;
  struct In {
    real_t av, xp, yp, zp, xm, ym, zm;
  };
  __device__ In load(int i, int j, int k) const {
    return In{vn.p ? fabs(vn(i, j, k)) : vconst, xp(i, j, k), yp(i, j, k), zp(i, j, k), xm(i, j, k), ym(i, j, k),
              zm(i, j, k)};
  }
  __device__ double eval(const In& a, int, int, int) const {
    return (double)stencil::normal_h_over_dx(a.av, a.xp, a.yp, a.zp, a.xm, a.ym, a.zm, mdx2, ix, iy, iz);
  }
  __device__ double operator()(int i, int j, int k) const { return eval(load(i, j, k), i, j, k); }
};

// sign = -1: smoothed volume of {phi < 0}; sign = +1: of {phi > 0}.  F.p null means F = 1.
struct VolumeTerm {
  View<const real_t> phi, F;
  real_t eps, dV;
  int sign;
  struct In {
    real_t p, f;
  };
  __device__ In load(int i, int j, int k) const { return In{phi(i, j, k), F.p ? F(i, j, k) : real_t(1)}; }
  __device__ double eval(const In& v, int, int, int) const {
    // q = phi / eps only inside the smoothing band, where it is used
    const real_t p = v.p, f = v.f;
    if (sign < 0) {
      if (p < -eps) return (double)(F.p ? f * dV : dV);
      if (p < eps) {
        const real_t q = p / eps;
        const real_t h = real_t(0.5) * (real_t(1) - q - real_t(kOneOverPi) * sin(real_t(kPi) * q));
        return (double)(F.p ? h * f * dV : h * dV);
      }
    } else {
      if (p > eps) return (double)(F.p ? f * dV : dV);
      if (p > -eps) {
        const real_t q = p / eps;
        const real_t h = real_t(0.5) * (real_t(1) + q + real_t(kOneOverPi) * sin(real_t(kPi) * q));
        return (double)(F.p ? h * f * dV : h * dV);
      }
    }
    return 0.0;
  }
  __device__ double operator()(int i, int j, int k) const { return eval(load(i, j, k), i, j, k); }
};

struct SurfaceTerm {
  View<const real_t> phi, F, x, y, z;
  real_t eps, ie, dV;
  struct In {
    real_t p;
  };
  __device__ In load(int i, int j, int k) const { return In{phi(i, j, k)}; }
  // the gradient (and F) only in the smoothing band
  __device__ double eval(const In& v, int i, int j, int k) const {
    const real_t p = v.p;
    if (!(fabs(p) < eps)) return 0.0;
    const real_t delta = real_t(0.5) * ie * (real_t(1) + cos(real_t(kPi) * p * ie));
    const real_t gx = x(i, j, k), gy = y(i, j, k), gz = z(i, j, k);
    const real_t n = sqrt(gx * gx + gy * gy + gz * gz);
    return (double)(F.p ? delta * n * F(i, j, k) * dV : delta * n * dV);
  }
  __device__ double operator()(int i, int j, int k) const { return eval(load(i, j, k), i, j, k); }
};

// The cells of a control volume (src/toolbox/lsm_math_utils3d.f.in:852-1844):
// control_vol > 0 when control_vol_sgn > 0, control_vol < 0 otherwise.
struct CvMask {
  View<const real_t> cv;
  int sgn;
  __device__ bool in(int i, int j, int k) const {
    const real_t c = cv(i, j, k);
    return sgn > 0 ? c > real_t(0) : c < real_t(0);
  }
};
// A sum term on the control volume's cells only.
template <class Op>
struct CvSum {
  Op op;
  CvMask m;
  __device__ double operator()(int i, int j, int k) const { return m.in(i, j, k) ? op(i, j, k) : 0.0; }
};
// A maximum term on the control volume's cells only: -1 elsewhere, for the
// reduction that reports the reference's initial -1 when no cell takes part.
template <class Op>
struct CvMax {
  Op op;
  CvMask m;
  __device__ double operator()(int i, int j, int k) const { return m.in(i, j, k) ? op(i, j, k) : -1.0; }
};
// The max-norm difference on a control volume starts from the first cell of
// the box whatever its control volume (lsm_math_utils3d.f.in:887-888), so that
// cell always takes part.
struct MaxNormDiffCv {
  MaxNormDiff d;
  CvMask m;
  int ci, cj, ck;
  __device__ double operator()(int i, int j, int k) const {
    return m.in(i, j, k) || (i == ci && j == cj && k == ck) ? d(i, j, k) : 0.0;
  }
};
// 1 where phi > 0 (sign > 0) or phi < 0 (sign < 0): the voxel counts; on a
// control volume these take control_vol >= 0 (sgn > 0) or <= 0
// (lsm_math_utils3d.f.in:1849-2007).
struct SignCount {
  View<const real_t> phi, cv;
  int sign, sgn;
  struct In {
    real_t p;
  };
  __device__ In load(int i, int j, int k) const { return In{phi(i, j, k)}; }
  __device__ double eval(const In& v, int i, int j, int k) const {
    const real_t p = v.p;
    if (!(sign > 0 ? p > real_t(0) : p < real_t(0))) return 0.0;
    if (cv.p) {
      const real_t c = cv(i, j, k);
      if (!(sgn > 0 ? c >= real_t(0) : c <= real_t(0))) return 0.0;
    }
    return 1.0;
  }
  __device__ double operator()(int i, int j, int k) const { return eval(load(i, j, k), i, j, k); }
};
// delta_phi |grad phi| dV where a given delta function is positive
// (lsm_geometry3d_fort.f.in:669-745).
struct DeltaAreaTerm {
  View<const real_t> d, x, y, z;
  real_t dV;
  struct In {
    real_t dp;
  };
  __device__ In load(int i, int j, int k) const { return In{d(i, j, k)}; }
  // the gradient only where the delta function is positive
  __device__ double eval(const In& v, int i, int j, int k) const {
    const real_t dp = v.dp;
    if (!(dp > real_t(0))) return 0.0;
    const real_t gx = x(i, j, k), gy = y(i, j, k), gz = z(i, j, k);
    return (double)(dp * sqrt(gx * gx + gy * gy + gz * gz) * dV);
  }
  __device__ double operator()(int i, int j, int k) const { return eval(load(i, j, k), i, j, k); }
};

bool ok(const char* who, const Box& ib, std::initializer_list<std::pair<Box, const char*>> arrays) {
  for (auto& a : arrays)
    if (!check_box(a.first, ib, who, a.second) || !check_size(a.first, who)) return false;
  return true;
}

real_t cfl_dt(double m, real_t cfl) { return cfl / (real_t(m) + kTiny); }

}  // namespace

// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3dmaxnormdiff_(real_t* out, const real_t* f1, LSMHIP_BOXP(g1), const real_t* f2,
                                     LSMHIP_BOXP(g2), LSMHIP_BOXP(ib)) {
  LSMHIP_REQUIRE_ARRAYS(f1, f2);
  const char* who = "lsm3dmaxnormdiff_";
  Box ib = LSMHIP_BOX(ib), g1 = LSMHIP_BOX(g1), g2 = LSMHIP_BOX(g2);
  *out = real_t(0);
  if (ib.empty() || !ok(who, ib, {{g1, "field1"}, {g2, "field2"}})) return;
  double m = 0;
  if (reduce::box_max(who, ib, MaxNormDiff{view(f1, g1), view(f2, g2)}, &m)) *out = real_t(m);
}

LSMHIP_EXPORT void lsm3dcomputestableadvectiondt_(real_t* dt, const real_t* vx, const real_t* vy, const real_t* vz,
                                                  LSMHIP_BOXP(vg), LSMHIP_BOXP(ib), const real_t* dx,
                                                  const real_t* dy, const real_t* dz, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(vx, vy, vz);
  const char* who = "lsm3dcomputestableadvectiondt_";
  Box ib = LSMHIP_BOX(ib), vg = LSMHIP_BOX(vg);
  double m = -1.0;
  if (!ib.empty()) {
    if (!ok(who, ib, {{vg, "vel"}})) return;
    AdvectionCfl op{view(vx, vg), view(vy, vg), view(vz, vg), real_t(1.0) / *dx, real_t(1.0) / *dy,
                    real_t(1.0) / *dz};
    if (!reduce::box_max(who, ib, op, &m)) return;
  }
  *dt = cfl_dt(m, *cfl);
}

// cv: the control volume whose cells take part (all of ib when null).
static void normal_dt(const char* who, real_t* dt, const real_t* vn, const Box& vg, real_t vconst, const real_t* pxp,
                      const real_t* pyp, const real_t* pzp, const Box& gp, const real_t* pxm, const real_t* pym,
                      const real_t* pzm, const Box& gm, const Box& ib, real_t dx, real_t dy, real_t dz, real_t cfl,
                      const CvMask* cv = nullptr) {
  double m = -1.0;
  if (!ib.empty()) {
    if (!ok(who, ib, {{gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}})) return;
    if (vn && !ok(who, ib, {{vg, "vel_n"}})) return;
    real_t h = dx > dy ? dx : dy;
    h = h > dz ? h : dz;
    NormalCfl op{vn ? view(vn, vg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0},
                 view(pxp, gp), view(pyp, gp), view(pzp, gp), view(pxm, gm), view(pym, gm), view(pzm, gm),
                 vconst, h * h, real_t(1.0) / dx, real_t(1.0) / dy, real_t(1.0) / dz};
    if (cv ? !reduce::box_max<CvMax<NormalCfl>, true>(who, ib, CvMax<NormalCfl>{op, *cv}, &m)
           : !reduce::box_max(who, ib, op, &m))
      return;
  }
  *dt = cfl_dt(m, cfl);
}

LSMHIP_EXPORT void lsm3dcomputestablenormalveldt_(real_t* dt, const real_t* vn, LSMHIP_BOXP(vg), const real_t* pxp,
                                                  const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                  const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                  LSMHIP_BOXP(gm), LSMHIP_BOXP(ib), const real_t* dx,
                                                  const real_t* dy, const real_t* dz, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(vn, pxp, pyp, pzp, pxm, pym, pzm);
  normal_dt("lsm3dcomputestablenormalveldt_", dt, vn, LSMHIP_BOX(vg), real_t(0), pxp, pyp, pzp, LSMHIP_BOX(gp),
            pxm, pym, pzm, LSMHIP_BOX(gm), LSMHIP_BOX(ib), *dx, *dy, *dz, *cfl);
}

LSMHIP_EXPORT void lsm3dcomputestableconstnormalveldt_(real_t* dt, const real_t* vel_n, const real_t* pxp,
                                                       const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                       const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                       LSMHIP_BOXP(gm), LSMHIP_BOXP(ib), const real_t* dx,
                                                       const real_t* dy, const real_t* dz, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm);
  Box none{0, -1, 0, -1, 0, -1};
  normal_dt("lsm3dcomputestableconstnormalveldt_", dt, nullptr, none, real_t(fabs(*vel_n)), pxp, pyp, pzp,
            LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm), LSMHIP_BOX(ib), *dx, *dy, *dz, *cfl);
}

// ----- integrals -----------------------------------------------------------
// buf: device memory for the sum's partials (2 * reduce::kMaxBlocks + 2
// doubles), or null for the library workspace
// A sum over ib of op, or of op on the control volume's cells when cv is set.
template <class Op>
static bool sum_over(const char* who, const Box& ib, const Op& op, const CvMask* cv, double* s, double* buf = nullptr) {
  if (cv) return reduce::box_sum(who, ib, CvSum<Op>{op, *cv}, s, buf);
  return reduce::box_sum(who, ib, op, s, buf);
}

// A control volume's mask after its array has been checked against ib.
static bool cv_mask(const char* who, const Box& ib, const real_t* cv, const Box& cg, int sgn, CvMask* m) {
  if (!ok(who, ib, {{cg, "control_vol"}})) return false;
  *m = CvMask{view(cv, cg), sgn};
  return true;
}

static double volume(const char* who, const real_t* F, const Box& fg, const real_t* phi, const Box& pg,
                     const Box& ib, real_t dx, real_t dy, real_t dz, real_t eps, int sign, bool* good,
                     double* buf = nullptr, const CvMask* cv = nullptr) {
  *good = false;
  if (ib.empty()) {
    *good = true;
    return 0.0;
  }
  if (!ok(who, ib, {{pg, "phi"}})) return 0.0;
  if (F && !ok(who, ib, {{fg, "F"}})) return 0.0;
  VolumeTerm op{view(phi, pg), F ? view(F, fg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0}, eps, dx * dy * dz,
                sign};
  double s = 0;
  *good = sum_over(who, ib, op, cv, &s, buf);
  return s;
}

static double surface(const char* who, const real_t* F, const Box& fg, const real_t* phi, const Box& pg,
                      const real_t* px, const real_t* py, const real_t* pz, const Box& g, const Box& ib, real_t dx,
                      real_t dy, real_t dz, real_t eps, bool* good, const CvMask* cv = nullptr) {
  *good = false;
  if (ib.empty()) {
    *good = true;
    return 0.0;
  }
  if (!ok(who, ib, {{pg, "phi"}, {g, "grad_phi"}})) return 0.0;
  if (F && !ok(who, ib, {{fg, "F"}})) return 0.0;
  SurfaceTerm op{view(phi, pg), F ? view(F, fg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0}, view(px, g),
                 view(py, g), view(pz, g), eps, real_t(1.0) / eps, dx * dy * dz};
  double s = 0;
  *good = sum_over(who, ib, op, cv, &s);
  return s;
}

LSMHIP_EXPORT void lsm3dvolumeintegralphilessthanzero_(real_t* out, const real_t* F, LSMHIP_BOXP(fg),
                                                       const real_t* phi, LSMHIP_BOXP(pg), LSMHIP_BOXP(ib),
                                                       const real_t* dx, const real_t* dy, const real_t* dz,
                                                       const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi);
  bool good;
  double s = volume("lsm3dvolumeintegralphilessthanzero_", F, LSMHIP_BOX(fg), phi, LSMHIP_BOX(pg), LSMHIP_BOX(ib),
                    *dx, *dy, *dz, *eps, -1, &good);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dvolumeintegralphigreaterthanzero_(real_t* out, const real_t* F, LSMHIP_BOXP(fg),
                                                          const real_t* phi, LSMHIP_BOXP(pg), LSMHIP_BOXP(ib),
                                                          const real_t* dx, const real_t* dy, const real_t* dz,
                                                          const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi);
  bool good;
  double s = volume("lsm3dvolumeintegralphigreaterthanzero_", F, LSMHIP_BOX(fg), phi, LSMHIP_BOX(pg),
                    LSMHIP_BOX(ib), *dx, *dy, *dz, *eps, +1, &good);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dsurfaceintegral_(real_t* out, const real_t* F, LSMHIP_BOXP(fg), const real_t* phi,
                                         LSMHIP_BOXP(pg), const real_t* px, const real_t* py, const real_t* pz,
                                         LSMHIP_BOXP(g), LSMHIP_BOXP(ib), const real_t* dx, const real_t* dy,
                                         const real_t* dz, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi, px, py, pz);
  bool good;
  double s = surface("lsm3dsurfaceintegral_", F, LSMHIP_BOX(fg), phi, LSMHIP_BOX(pg), px, py, pz, LSMHIP_BOX(g),
                     LSMHIP_BOX(ib), *dx, *dy, *dz, *eps, &good);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dvolumeregionphilessthanzero_(real_t* out, const real_t* phi, LSMHIP_BOXP(pg),
                                                     LSMHIP_BOXP(ib), const real_t* dx, const real_t* dy,
                                                     const real_t* dz, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  bool good;
  Box none{0, -1, 0, -1, 0, -1};
  double s = volume("lsm3dvolumeregionphilessthanzero_", nullptr, none, phi, LSMHIP_BOX(pg), LSMHIP_BOX(ib), *dx,
                    *dy, *dz, *eps, -1, &good);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dvolumeregionphigreaterthanzero_(real_t* out, const real_t* phi, LSMHIP_BOXP(pg),
                                                        LSMHIP_BOXP(ib), const real_t* dx, const real_t* dy,
                                                        const real_t* dz, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi);
  bool good;
  Box none{0, -1, 0, -1, 0, -1};
  double s = volume("lsm3dvolumeregionphigreaterthanzero_", nullptr, none, phi, LSMHIP_BOX(pg), LSMHIP_BOX(ib), *dx,
                    *dy, *dz, *eps, +1, &good);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dsurfaceareazerolevelset_(real_t* out, const real_t* phi, LSMHIP_BOXP(pg), const real_t* px,
                                                 const real_t* py, const real_t* pz, LSMHIP_BOXP(g), LSMHIP_BOXP(ib),
                                                 const real_t* dx, const real_t* dy, const real_t* dz,
                                                 const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi, px, py, pz);
  bool good;
  Box none{0, -1, 0, -1, 0, -1};
  double s = surface("lsm3dsurfaceareazerolevelset_", nullptr, none, phi, LSMHIP_BOX(pg), px, py, pz, LSMHIP_BOX(g),
                     LSMHIP_BOX(ib), *dx, *dy, *dz, *eps, &good);
  if (good) *out = real_t(s);
}

// ---------------------------------------------------------------------------
// Average difference, voxel counts, area from a delta function
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3daveabsdiff_(real_t* out, const real_t* f1, LSMHIP_BOXP(g1), const real_t* f2, LSMHIP_BOXP(g2),
                                    LSMHIP_BOXP(ib)) {
  LSMHIP_REQUIRE_ARRAYS(f1, f2);
  const char* who = "lsm3daveabsdiff_";
  const Box ib = LSMHIP_BOX(ib), g1 = LSMHIP_BOX(g1), g2 = LSMHIP_BOX(g2);
  if (ib.empty()) {
    *out = real_t(0);
    return;
  }
  if (!ok(who, ib, {{g1, "field1"}, {g2, "field2"}})) return;
  double s = 0;
  if (!reduce::box_sum(who, ib, MaxNormDiff{view(f1, g1), view(f2, g2)}, &s)) return;
  // the sum rounded to the storage precision over the (exact) number of cells
  *out = real_t(s) / real_t((double)ib.nx() * ib.ny() * ib.nz());
}

static void voxel_count(const char* who, int* count, const real_t* phi, const Box& pg, const real_t* cv,
                        const Box& cg, int sgn, const Box& fb, int sign) {
  if (fb.empty()) {
    *count = 0;
    return;
  }
  if (!ok(who, fb, {{pg, "phi"}})) return;
  if (cv && !ok(who, fb, {{cg, "control_vol"}})) return;
  SignCount op{view(phi, pg), cv ? view(cv, cg) : View<const real_t>{nullptr, 0, 0, 0, 0, 0}, sign, sgn};
  double s = 0;
  if (reduce::box_sum(who, fb, op, &s)) *count = (int)s;
}
LSMHIP_EXPORT void lsm3dvoxelcountgreaterthanzero_(int* count, const real_t* phi, LSMHIP_BOXP(pg), LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(count, phi);
  voxel_count("lsm3dvoxelcountgreaterthanzero_", count, phi, LSMHIP_BOX(pg), nullptr, Box{0, -1, 0, -1, 0, -1}, 0,
              LSMHIP_BOX(fb), +1);
}
LSMHIP_EXPORT void lsm3dvoxelcountlessthanzero_(int* count, const real_t* phi, LSMHIP_BOXP(pg), LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(count, phi);
  voxel_count("lsm3dvoxelcountlessthanzero_", count, phi, LSMHIP_BOX(pg), nullptr, Box{0, -1, 0, -1, 0, -1}, 0,
              LSMHIP_BOX(fb), -1);
}

static void delta_area(const char* who, real_t* out, const real_t* dphi, const Box& pg, const real_t* px,
                       const real_t* py, const real_t* pz, const Box& g, const Box& ib, real_t dx, real_t dy,
                       real_t dz, const CvMask* cv) {
  double s = 0;
  if (!ib.empty()) {
    if (!ok(who, ib, {{pg, "delta_phi"}, {g, "grad_phi"}})) return;
    DeltaAreaTerm op{view(dphi, pg), view(px, g), view(py, g), view(pz, g), dx * dy * dz};
    if (!sum_over(who, ib, op, cv, &s)) return;
  }
  *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dsurfaceareazerolevelsetdelta_(real_t* out, const real_t* dphi, LSMHIP_BOXP(pg),
                                                      const real_t* px, const real_t* py, const real_t* pz,
                                                      LSMHIP_BOXP(g), LSMHIP_BOXP(ib), const real_t* dx,
                                                      const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(dphi, px, py, pz);
  delta_area("lsm3dsurfaceareazerolevelsetdelta_", out, dphi, LSMHIP_BOX(pg), px, py, pz, LSMHIP_BOX(g),
             LSMHIP_BOX(ib), *dx, *dy, *dz, nullptr);
}

// ---------------------------------------------------------------------------
// Control-volume forms (src/toolbox/lsm_math_utils3d.f.in:852-2007,
// src/geometry/lsm_geometry3d_fort.f.in:770-1329): the whole-box computation
// on the cells of ib inside the control volume.
// ---------------------------------------------------------------------------
#define CV_ARGS const real_t *cv, LSMHIP_BOXP(cg), const int *cvs
#define CV_MASK(who)                                                                    \
  CvMask m{};                                                                           \
  const Box ib_ = LSMHIP_BOX(ib);                                                       \
  if (!ib_.empty() && !cv_mask(who, ib_, cv, LSMHIP_BOX(cg), *cvs, &m)) return

LSMHIP_EXPORT void lsm3dmaxnormdiffcontrolvolume_(real_t* out, const real_t* f1, LSMHIP_BOXP(g1), const real_t* f2,
                                                  LSMHIP_BOXP(g2), CV_ARGS, LSMHIP_BOXP(ib)) {
  LSMHIP_REQUIRE_ARRAYS(f1, f2, cv);
  const char* who = "lsm3dmaxnormdiffcontrolvolume_";
  const Box ib = LSMHIP_BOX(ib), g1 = LSMHIP_BOX(g1), g2 = LSMHIP_BOX(g2);
  *out = real_t(0);
  if (ib.empty() || !ok(who, ib, {{g1, "field1"}, {g2, "field2"}})) return;
  CvMask m;
  if (!cv_mask(who, ib, cv, LSMHIP_BOX(cg), *cvs, &m)) return;
  double r = 0;
  MaxNormDiffCv op{MaxNormDiff{view(f1, g1), view(f2, g2)}, m, ib.ilo, ib.jlo, ib.klo};
  if (reduce::box_max(who, ib, op, &r)) *out = real_t(r);
}

LSMHIP_EXPORT void lsm3dcomputestableadvectiondtcontrolvolume_(real_t* dt, const real_t* vx, const real_t* vy,
                                                               const real_t* vz, LSMHIP_BOXP(vg), CV_ARGS,
                                                               LSMHIP_BOXP(ib), const real_t* dx, const real_t* dy,
                                                               const real_t* dz, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(vx, vy, vz, cv);
  const char* who = "lsm3dcomputestableadvectiondtcontrolvolume_";
  const Box ib = LSMHIP_BOX(ib), vg = LSMHIP_BOX(vg);
  double m = -1.0;
  if (!ib.empty()) {
    if (!ok(who, ib, {{vg, "vel"}})) return;
    CvMask cm;
    if (!cv_mask(who, ib, cv, LSMHIP_BOX(cg), *cvs, &cm)) return;
    AdvectionCfl op{view(vx, vg), view(vy, vg), view(vz, vg), real_t(1.0) / *dx, real_t(1.0) / *dy,
                    real_t(1.0) / *dz};
    if (!reduce::box_max<CvMax<AdvectionCfl>, true>(who, ib, CvMax<AdvectionCfl>{op, cm}, &m)) return;
  }
  *dt = cfl_dt(m, *cfl);
}

LSMHIP_EXPORT void lsm3dcomputestablenormalveldtcontrolvolume_(
    real_t* dt, const real_t* vn, LSMHIP_BOXP(vg), const real_t* pxp, const real_t* pyp, const real_t* pzp,
    LSMHIP_BOXP(gp), const real_t* pxm, const real_t* pym, const real_t* pzm, LSMHIP_BOXP(gm), CV_ARGS,
    LSMHIP_BOXP(ib), const real_t* dx, const real_t* dy, const real_t* dz, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(vn, pxp, pyp, pzp, pxm, pym, pzm, cv);
  const char* who = "lsm3dcomputestablenormalveldtcontrolvolume_";
  CV_MASK(who);
  normal_dt(who, dt, vn, LSMHIP_BOX(vg), real_t(0), pxp, pyp, pzp, LSMHIP_BOX(gp), pxm, pym, pzm, LSMHIP_BOX(gm),
            ib_, *dx, *dy, *dz, *cfl, &m);
}

LSMHIP_EXPORT void lsm3dcomputestableconstnormalveldtcontrolvolume_(
    real_t* dt, const real_t* vel_n, const real_t* pxp, const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
    const real_t* pxm, const real_t* pym, const real_t* pzm, LSMHIP_BOXP(gm), CV_ARGS, LSMHIP_BOXP(ib),
    const real_t* dx, const real_t* dy, const real_t* dz, const real_t* cfl) {
  LSMHIP_REQUIRE_ARRAYS(pxp, pyp, pzp, pxm, pym, pzm, cv);
  const char* who = "lsm3dcomputestableconstnormalveldtcontrolvolume_";
  CV_MASK(who);
  normal_dt(who, dt, nullptr, Box{0, -1, 0, -1, 0, -1}, real_t(fabs(*vel_n)), pxp, pyp, pzp, LSMHIP_BOX(gp), pxm,
            pym, pzm, LSMHIP_BOX(gm), ib_, *dx, *dy, *dz, *cfl, &m);
}

LSMHIP_EXPORT void lsm3dvolumeintegralphilessthanzerocontrolvolume_(real_t* out, const real_t* F, LSMHIP_BOXP(fg),
                                                                    const real_t* phi, LSMHIP_BOXP(pg), CV_ARGS,
                                                                    LSMHIP_BOXP(ib), const real_t* dx,
                                                                    const real_t* dy, const real_t* dz,
                                                                    const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi, cv);
  const char* who = "lsm3dvolumeintegralphilessthanzerocontrolvolume_";
  CV_MASK(who);
  bool good;
  double s = volume(who, F, LSMHIP_BOX(fg), phi, LSMHIP_BOX(pg), ib_, *dx, *dy, *dz, *eps, -1, &good, nullptr, &m);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dvolumeintegralphigreaterthanzerocontrolvolume_(
    real_t* out, const real_t* F, LSMHIP_BOXP(fg), const real_t* phi, LSMHIP_BOXP(pg), CV_ARGS, LSMHIP_BOXP(ib),
    const real_t* dx, const real_t* dy, const real_t* dz, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi, cv);
  const char* who = "lsm3dvolumeintegralphigreaterthanzerocontrolvolume_";
  CV_MASK(who);
  bool good;
  double s = volume(who, F, LSMHIP_BOX(fg), phi, LSMHIP_BOX(pg), ib_, *dx, *dy, *dz, *eps, +1, &good, nullptr, &m);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dsurfaceintegralcontrolvolume_(real_t* out, const real_t* F, LSMHIP_BOXP(fg),
                                                      const real_t* phi, LSMHIP_BOXP(pg), const real_t* px,
                                                      const real_t* py, const real_t* pz, LSMHIP_BOXP(g), CV_ARGS,
                                                      LSMHIP_BOXP(ib), const real_t* dx, const real_t* dy,
                                                      const real_t* dz, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(F, phi, px, py, pz, cv);
  const char* who = "lsm3dsurfaceintegralcontrolvolume_";
  CV_MASK(who);
  bool good;
  double s = surface(who, F, LSMHIP_BOX(fg), phi, LSMHIP_BOX(pg), px, py, pz, LSMHIP_BOX(g), ib_, *dx, *dy, *dz,
                     *eps, &good, &m);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dvoxelcountgreaterthanzerocontrolvolume_(int* count, const real_t* phi, LSMHIP_BOXP(pg),
                                                                CV_ARGS, LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(count, phi, cv);
  voxel_count("lsm3dvoxelcountgreaterthanzerocontrolvolume_", count, phi, LSMHIP_BOX(pg), cv, LSMHIP_BOX(cg), *cvs,
              LSMHIP_BOX(fb), +1);
}
LSMHIP_EXPORT void lsm3dvoxelcountlessthanzerocontrolvolume_(int* count, const real_t* phi, LSMHIP_BOXP(pg),
                                                             CV_ARGS, LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(count, phi, cv);
  voxel_count("lsm3dvoxelcountlessthanzerocontrolvolume_", count, phi, LSMHIP_BOX(pg), cv, LSMHIP_BOX(cg), *cvs,
              LSMHIP_BOX(fb), -1);
}
LSMHIP_EXPORT void lsm3dvolumeregionphilessthanzerocontrolvolume_(real_t* out, const real_t* phi, LSMHIP_BOXP(pg),
                                                                  CV_ARGS, LSMHIP_BOXP(ib), const real_t* dx,
                                                                  const real_t* dy, const real_t* dz,
                                                                  const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi, cv);
  const char* who = "lsm3dvolumeregionphilessthanzerocontrolvolume_";
  CV_MASK(who);
  bool good;
  double s = volume(who, nullptr, Box{0, -1, 0, -1, 0, -1}, phi, LSMHIP_BOX(pg), ib_, *dx, *dy, *dz, *eps, -1, &good,
                    nullptr, &m);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dvolumeregionphigreaterthanzerocontrolvolume_(real_t* out, const real_t* phi,
                                                                     LSMHIP_BOXP(pg), CV_ARGS, LSMHIP_BOXP(ib),
                                                                     const real_t* dx, const real_t* dy,
                                                                     const real_t* dz, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi, cv);
  const char* who = "lsm3dvolumeregionphigreaterthanzerocontrolvolume_";
  CV_MASK(who);
  bool good;
  double s = volume(who, nullptr, Box{0, -1, 0, -1, 0, -1}, phi, LSMHIP_BOX(pg), ib_, *dx, *dy, *dz, *eps, +1, &good,
                    nullptr, &m);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dsurfaceareazerolevelsetcontrolvolume_(real_t* out, const real_t* phi, LSMHIP_BOXP(pg),
                                                              const real_t* px, const real_t* py, const real_t* pz,
                                                              LSMHIP_BOXP(g), CV_ARGS, LSMHIP_BOXP(ib),
                                                              const real_t* dx, const real_t* dy, const real_t* dz,
                                                              const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS(phi, px, py, pz, cv);
  const char* who = "lsm3dsurfaceareazerolevelsetcontrolvolume_";
  CV_MASK(who);
  bool good;
  double s = surface(who, nullptr, Box{0, -1, 0, -1, 0, -1}, phi, LSMHIP_BOX(pg), px, py, pz, LSMHIP_BOX(g), ib_,
                     *dx, *dy, *dz, *eps, &good, &m);
  if (good) *out = real_t(s);
}
LSMHIP_EXPORT void lsm3dsurfaceareazerolevelsetdeltacontrolvolume_(real_t* out, const real_t* dphi, LSMHIP_BOXP(pg),
                                                                   const real_t* px, const real_t* py,
                                                                   const real_t* pz, LSMHIP_BOXP(g), CV_ARGS,
                                                                   LSMHIP_BOXP(ib), const real_t* dx,
                                                                   const real_t* dy, const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(dphi, px, py, pz, cv);
  const char* who = "lsm3dsurfaceareazerolevelsetdeltacontrolvolume_";
  CV_MASK(who);
  delta_area(who, out, dphi, LSMHIP_BOX(pg), px, py, pz, LSMHIP_BOX(g), ib_, *dx, *dy, *dz, &m);
}
#undef CV_MASK
#undef CV_ARGS

// FP64-accumulated volume of {phi < 0} returned in double whatever the storage
// precision (used by the FP32 path for its interface-volume error).
LSMHIP_EXPORT int lsmhip_volume_phi_lt0_f64(double* out, const real_t* phi, LSMHIP_BOXP(pg), LSMHIP_BOXP(ib),
                                            const double* dx, const double* dy, const double* dz,
                                            const double* eps) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, out, phi);
  bool good;
  Box none{0, -1, 0, -1, 0, -1};
  double s = volume("lsmhip_volume_phi_lt0_f64", nullptr, none, phi, LSMHIP_BOX(pg), LSMHIP_BOX(ib), real_t(*dx),
                    real_t(*dy), real_t(*dz), real_t(*eps), -1, &good);
  if (good) *out = s;
  return good ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Allocation-free volume of {phi < 0} (no counterpart in the reference): the
// computation of lsm3dvolumeregionphilessthanzero_ -- the same kernels and
// block count, hence the same bits -- with the per-block partial sums and the
// result kept in `scratch`, a file-scope device array, instead of the
// library's workspace, which may call hipMalloc.  For callers that cannot
// allocate device memory (code under HIP graph capture, serving loops).
//
// Each GPU holds its own copy of scratch (2 * reduce::kMaxBlocks + 2 doubles,
// 64 KiB) until the process ends.  The sum kernel never runs more than
// reduce::kMaxBlocks blocks, so scratch covers every box the library takes
// (fewer than 2^31 cells); a larger box is refused before anything is
// launched.  Every slot the finishing kernel reads is written earlier in the
// same call, so nothing needs clearing between calls; calls that overlap on
// one GPU (several host threads or streams) race on scratch.
// ---------------------------------------------------------------------------
namespace {
constexpr int kScratchDoubles = 2 * reduce::kMaxBlocks + 2;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxScratchDevices] = {};

// scratch's device address on the current device, looked up on the first call
// made with that device current (host code must not use the name itself)
double* scratch_address(const char* who) {
  int d = 0;
  if (!check_hip(hipGetDevice(&d), who)) return nullptr;
  if (d < 0 || d >= kMaxScratchDevices) {
    set_error(kErrArg, "%s: device %d beyond the %d this entry point serves", who, d, kMaxScratchDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[d]) {
    void* p = nullptr;
    if (!check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), who)) return nullptr;
    g_scratch[d] = static_cast<double*>(p);
  }
  return g_scratch[d];
}
}  // namespace

LSMHIP_EXPORT int lsmhip_volume_region_phi_lt0_noalloc(real_t* out, const real_t* phi, LSMHIP_BOXP(pg),
                                                       LSMHIP_BOXP(ib), const real_t* dx, const real_t* dy,
                                                       const real_t* dz, const real_t* eps) {
  LSMHIP_REQUIRE_ARRAYS_OR(-1, phi);
  const char* who = "lsmhip_volume_region_phi_lt0_noalloc";
  if (!out) {
    set_error(kErrArg, "%s: out is NULL", who);
    return -1;
  }
  const Box pgb = LSMHIP_BOX(pg), ibb = LSMHIP_BOX(ib);
  // the sizes are checked before any device work (volume() checks them again)
  if (!ok(who, ibb, {{pgb, "phi"}})) return -1;
  static_assert(kScratchDoubles >= 2 * reduce::kMaxBlocks + 2, "scratch holds every block's partial pair");
  double* buf = scratch_address(who);
  if (!buf) return -1;
  bool good;
  Box none{0, -1, 0, -1, 0, -1};
  double s = volume(who, nullptr, none, phi, pgb, ibb, *dx, *dy, *dz, *eps, -1, &good, buf);
  if (good) *out = real_t(s);
  return good ? 0 : -1;
}

// Smoothed Heaviside and delta (src/toolbox/lsm_calculus_toolbox.f:25-92); host scalars.
LSMHIP_EXPORT void lsmheaviside_(real_t* H, const real_t* x, const real_t* epsilon) {
  const real_t e = *epsilon, v = *x;
  if (v < -e) *H = real_t(0);
  else if (v > e) *H = real_t(1);
  else {
    const real_t ie = real_t(1.0) / e;
    *H = real_t(0.5) * (real_t(1) + v * ie + real_t(kOneOverPi) * std::sin(real_t(kPi) * v * ie));
  }
}
LSMHIP_EXPORT void lsmdeltafunction_(real_t* delta, const real_t* x, const real_t* epsilon) {
  const real_t e = *epsilon, v = *x;
  if (fabs(v) > e) *delta = real_t(0);
  else {
    const real_t ie = real_t(1.0) / e;
    *delta = real_t(0.5) * ie * (real_t(1) + std::cos(real_t(kPi) * v * ie));
  }
}
