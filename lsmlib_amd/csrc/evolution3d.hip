// evolution3d.hip -- level-set equation RHS terms, TVD Runge-Kutta stages,
// reinitialization RHS and second-order mean curvature (whole grid).
// References: src/toolbox/lsm_level_set_evolution3d.f.in,
// src/toolbox/lsm_tvd_runge_kutta3d.f, src/reinitialization/lsm_reinitialization3d.f.in,
// src/geometry/lsm_curvature3d.f.in.
#include "lsmhip_core.hpp"
#include "march.hpp"
#include "stencil.hpp"

using namespace lsmhip;

namespace {

struct Six {
  View<const real_t> xp, yp, zp, xm, ym, zm;
};
Six six(const real_t* xp, const real_t* yp, const real_t* zp, const Box& gp, const real_t* xm, const real_t* ym,
        const real_t* zm, const Box& gm) {
  return Six{view(xp, gp), view(yp, gp), view(zp, gp), view(xm, gm), view(ym, gm), view(zm, gm)};
}

__global__ void __launch_bounds__(256) advection_kernel(Box fb, View<real_t> rhs, View<const real_t> px,
                                                        View<const real_t> py, View<const real_t> pz,
                                                        View<const real_t> vx, View<const real_t> vy,
                                                        View<const real_t> vz) {
  LSMHIP_CELLS(fb, kPointNK);
  struct V {
    real_t r, u, x, v, y, w, z;
  };
  pair_cells(
      k0, nk,
      [&](int k) { return V{rhs(i, j, k), vx(i, j, k), px(i, j, k), vy(i, j, k), py(i, j, k), vz(i, j, k), pz(i, j, k)}; },
      [&](int k, const V& a) { rhs(i, j, k) = a.r - (a.u * a.x + a.v * a.y + a.w * a.z); });
}

// vconst is used when vn.p is null (constant normal speed).
__global__ void __launch_bounds__(256) normal_kernel(Box fb, View<real_t> rhs, Six g, View<const real_t> vn,
                                                     real_t vconst) {
  LSMHIP_CELLS(fb, kPointNK);
  // every input of both cells is loaded first (also where the speed is below
  // zero_tol and the cell is left as it is)
  struct V {
    real_t v, xp, yp, zp, xm, ym, zm, r;
  };
  pair_cells(
      k0, nk,
      [&](int k) {
        return V{vn.p ? vn(i, j, k) : vconst, g.xp(i, j, k), g.yp(i, j, k), g.zp(i, j, k), g.xm(i, j, k),
                 g.ym(i, j, k), g.zm(i, j, k), rhs(i, j, k)};
      },
      [&](int k, const V& a) {
        if (fabs(a.v) >= kZeroTol)
          rhs(i, j, k) = a.r - a.v * sqrt(stencil::godunov_sq(a.v > real_t(0), a.xp, a.yp, a.zp, a.xm, a.ym, a.zm));
      });
}

__global__ void __launch_bounds__(256) const_curv_kernel(Box fb, View<real_t> rhs, View<const real_t> px,
                                                         View<const real_t> py, View<const real_t> pz,
                                                         View<const real_t> xx, View<const real_t> xy,
                                                         View<const real_t> xz, View<const real_t> yy,
                                                         View<const real_t> yz, View<const real_t> zz, real_t b) {
  LSMHIP_CELL(fb);
  const real_t x = px(i, j, k), y = py(i, j, k), z = pz(i, j, k);
  const real_t g2 = x * x + y * y + z * z;
  real_t curv;
  if (g2 < kZeroTol) {
    curv = real_t(0);
  } else {
    curv = stencil::curv_num(xx(i, j, k), xy(i, j, k), xz(i, j, k), yy(i, j, k), yz(i, j, k), zz(i, j, k), x, y, z);
    curv = curv / g2;
  }
  rhs(i, j, k) = rhs(i, j, k) + b * curv;
}

__global__ void __launch_bounds__(256) precomputed_curv_kernel(Box lb, View<real_t> rhs, View<const real_t> kappa,
                                                               View<const real_t> gmag, real_t b) {
  LSMHIP_CELL(lb);
  rhs(i, j, k) = rhs(i, j, k) + b * kappa(i, j, k) * gmag(i, j, k);
}

__global__ void __launch_bounds__(256) ext_normal_kernel(Box lb, View<real_t> rhs, Six g, View<const real_t> vn,
                                                         View<const real_t> vx, View<const real_t> vy,
                                                         View<const real_t> vz) {
  LSMHIP_CELL(lb);
  const real_t n = vn(i, j, k), u = vx(i, j, k), v = vy(i, j, k), w = vz(i, j, k);
  const real_t px = stencil::ext_normal_pick(u, n, g.xp(i, j, k), g.xm(i, j, k));
  const real_t py = stencil::ext_normal_pick(v, n, g.yp(i, j, k), g.ym(i, j, k));
  const real_t pz = stencil::ext_normal_pick(w, n, g.zp(i, j, k), g.zm(i, j, k));
  const real_t s = px * px + py * py + pz * pz;
  if (fabs(n) >= kZeroTol && s >= kZeroTol)
    rhs(i, j, k) = rhs(i, j, k) - u * px - v * py - w * pz - n * sqrt(s);
  else
    rhs(i, j, k) = rhs(i, j, k) - u * px - v * py - w * pz;
}

__global__ void __launch_bounds__(256) reinit_kernel(Box fb, View<real_t> out, View<const real_t> phi,
                                                     View<const real_t> phi0, Six g, real_t dx_sq, int use_phi0) {
  LSMHIP_CELLS(fb, kPointNK);
  struct V {
    real_t s, xp, yp, zp, xm, ym, zm;
  };
  pair_cells(
      k0, nk,
      [&](int k) {
        return V{use_phi0 ? phi0(i, j, k) : phi(i, j, k), g.xp(i, j, k), g.yp(i, j, k), g.zp(i, j, k), g.xm(i, j, k),
                 g.ym(i, j, k), g.zm(i, j, k)};
      },
      [&](int k, const V& a) {
        out(i, j, k) = stencil::reinit_rhs(a.s, use_phi0 != 0, a.xp, a.yp, a.zp, a.xm, a.ym, a.zm, dx_sq);
      });
}

struct CurvFactors {
  real_t xx, yy, zz, xy, xz, yz;
};

// Second-order mean curvature (src/geometry/lsm_curvature3d.f.in:15-128): the inputs at
// one cell.
struct CurvInputs {
  real_t c, xp, xm, yp, ym, zp, zm;  // F at the cell and its 6 neighbours
  real_t xy[4], xz[4], yz[4];        // F at (+,+), (+,-), (-,+), (-,-) of each plane pair
  real_t x, y, z, gm;
};
// Compact second differences of F, numerator over |grad phi|^3 (zero when that
// is below zero_tol), in the reference's operation order.
__device__ __forceinline__ real_t mean_curvature_of(const CurvInputs& a, const CurvFactors& f) {
  const real_t xx = (a.xp - real_t(2) * a.c + a.xm) * f.xx;
  const real_t yy = (a.yp - real_t(2) * a.c + a.ym) * f.yy;
  const real_t zz = (a.zp - real_t(2) * a.c + a.zm) * f.zz;
  const real_t xy = (a.xy[0] - a.xy[1] - a.xy[2] + a.xy[3]) * f.xy;
  const real_t xz = (a.xz[0] - a.xz[1] - a.xz[2] + a.xz[3]) * f.xz;
  const real_t yz = (a.yz[0] - a.yz[1] - a.yz[2] + a.yz[3]) * f.yz;
  const real_t den = a.gm * a.gm * a.gm;
  if (den < kZeroTol) return real_t(0);
  return stencil::curv_num(xx, xy, xz, yy, yz, zz, a.x, a.y, a.z) / den;
}

// The mean curvature on the plane march: the 19-point cross from the march,
// the gradient and its magnitude loaded per cell.
struct MeanCurvMarch {
  View<real_t> kappa;
  View<const real_t> px, py, pz, gmag;
  CurvFactors f;
  struct In {
    real_t x, y, z, gm;
  };
  __device__ In load(int i, int j, int k) const { return In{px(i, j, k), py(i, j, k), pz(i, j, k), gmag(i, j, k)}; }
  __device__ void store(int i, int j, int k, const march::Cross19& c, const In& v) const {
    const CurvInputs a{c.c,
                       c.xp,
                       c.xm,
                       c.yp,
                       c.ym,
                       c.zp,
                       c.zm,
                       {c.xy[0], c.xy[1], c.xy[2], c.xy[3]},
                       {c.xz[0], c.xz[1], c.xz[2], c.xz[3]},
                       {c.yz[0], c.yz[1], c.yz[2], c.yz[3]},
                       v.x,
                       v.y,
                       v.z,
                       v.gm};
    kappa(i, j, k) = mean_curvature_of(a, f);
  }
};

// RK stage kinds: 1 = u + dt L;  2 = 0.5 (u + u1 + dt L);  32 = 0.75 u + 0.25 (u1 + dt L);
// 33 = 1/3 u + 2/3 (u2 + dt L)  (lsm_tvd_runge_kutta3d.f:78, :240, :403, :490).
template <int KIND>
__global__ void __launch_bounds__(256) rk_kernel(Box fb, View<real_t> out, View<const real_t> us,
                                                 View<const real_t> uc, View<const real_t> rhs, real_t dt) {
  LSMHIP_CELLS(fb, kPointNK);
  struct V {
    real_t r, c, s;  // rhs, u_cur, u_stage (not read for RK1)
  };
  pair_cells(
      k0, nk, [&](int k) { return V{rhs(i, j, k), uc(i, j, k), KIND == 1 ? real_t(0) : us(i, j, k)}; },
      [&](int k, const V& a) {
        if (KIND == 1) out(i, j, k) = a.c + dt * a.r;
        if (KIND == 2) out(i, j, k) = real_t(0.5) * (a.c + a.s + dt * a.r);
        if (KIND == 32) out(i, j, k) = real_t(0.75) * a.c + real_t(0.25) * (a.s + dt * a.r);
        if (KIND == 33) out(i, j, k) = real_t(1.0 / 3.0) * a.c + real_t(2.0 / 3.0) * (a.s + dt * a.r);
      });
}

bool boxes_ok(const char* who, const Box& fb, std::initializer_list<std::pair<Box, const char*>> arrays) {
  for (auto& a : arrays)
    if (!check_box(a.first, fb, who, a.second) || !check_size(a.first, who)) return false;
  return true;
}

template <int KIND>
void rk_stage(const char* who, real_t* out, const Box& og, const real_t* us, const Box& sg, const real_t* uc,
              const Box& cg, const real_t* rhs, const Box& rg, const Box& fb, real_t dt) {
  if (fb.empty()) return;
  if (!boxes_ok(who, fb, {{og, "u_next"}, {cg, "u_cur"}, {rg, "rhs"}})) return;
  if (KIND != 1 && !boxes_ok(who, fb, {{sg, "u_stage"}})) return;
  hipLaunchKernelGGL(rk_kernel<KIND>, grid_for_cells(fb, kPointNK), block2d(), 0, stream(), fb, view(out, og), view(us, sg),
                     view(uc, cg), view(rhs, rg), dt);
  finish(who);
}

}  // namespace

// ---------------------------------------------------------------------------
// RHS terms (src/toolbox/lsm_level_set_evolution3d.h)
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3dzerooutlevelseteqnrhs_(real_t* rhs, LSMHIP_BOXP(rg)) {
  LSMHIP_REQUIRE_ARRAYS(rhs);
  const char* who = "lsm3dzerooutlevelseteqnrhs_";
  Box rg = LSMHIP_BOX(rg);
  if (rg.empty() || !check_size(rg, who)) return;
  // the whole ghost box, as the reference zeroes it: rg.volume() contiguous scalars from rhs
  if (fill_contiguous(rhs, rg.volume(), real_t(0), who)) finish(who);
}

LSMHIP_EXPORT void lsm3daddadvectiontermtolserhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* px, const real_t* py,
                                                  const real_t* pz, LSMHIP_BOXP(g), const real_t* vx,
                                                  const real_t* vy, const real_t* vz, LSMHIP_BOXP(vg),
                                                  LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, pz, vx, vy, vz);
  const char* who = "lsm3daddadvectiontermtolserhs_";
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), g = LSMHIP_BOX(g), vg = LSMHIP_BOX(vg);
  if (fb.empty() || !boxes_ok(who, fb, {{rg, "lse_rhs"}, {g, "grad_phi"}, {vg, "vel"}})) return;
  hipLaunchKernelGGL(advection_kernel, grid_for_cells(fb, kPointNK), block2d(), 0, stream(), fb, view(rhs, rg), view(px, g),
                     view(py, g), view(pz, g), view(vx, vg), view(vy, vg), view(vz, vg));
  finish(who);
}

LSMHIP_EXPORT void lsm3daddnormalveltermtolserhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp,
                                                  const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                  const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                  LSMHIP_BOXP(gm), const real_t* vn, LSMHIP_BOXP(vg),
                                                  LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm, vn);
  const char* who = "lsm3daddnormalveltermtolserhs_";
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm), vg = LSMHIP_BOX(vg);
  if (fb.empty() ||
      !boxes_ok(who, fb, {{rg, "lse_rhs"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}, {vg, "vel_n"}}))
    return;
  hipLaunchKernelGGL(normal_kernel, grid_for_cells(fb, kPointNK), block2d(), 0, stream(), fb, view(rhs, rg),
                     six(pxp, pyp, pzp, gp, pxm, pym, pzm, gm), view(vn, vg), real_t(0));
  finish(who);
}

LSMHIP_EXPORT void lsm3daddconstnormalveltermtolserhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp,
                                                       const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                       const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                       LSMHIP_BOXP(gm), const real_t* vel_n, LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm);
  const char* who = "lsm3daddconstnormalveltermtolserhs_";
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm);
  if (fb.empty() || !(fabs(*vel_n) >= kZeroTol)) return;
  if (!boxes_ok(who, fb, {{rg, "lse_rhs"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}})) return;
  View<const real_t> none{nullptr, 0, 0, 0, 0, 0};
  hipLaunchKernelGGL(normal_kernel, grid_for_cells(fb, kPointNK), block2d(), 0, stream(), fb, view(rhs, rg),
                     six(pxp, pyp, pzp, gp, pxm, pym, pzm, gm), none, *vel_n);
  finish(who);
}

LSMHIP_EXPORT void lsm3daddconstcurvtermtolserhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* px, const real_t* py,
                                                  const real_t* pz, LSMHIP_BOXP(g), const real_t* xx,
                                                  const real_t* xy, const real_t* xz, const real_t* yy,
                                                  const real_t* yz, const real_t* zz, LSMHIP_BOXP(g2),
                                                  const real_t* b, LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, px, py, pz, xx, xy, xz, yy, yz, zz);
  const char* who = "lsm3daddconstcurvtermtolserhs_";
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), g = LSMHIP_BOX(g), g2 = LSMHIP_BOX(g2);
  if (fb.empty() || !boxes_ok(who, fb, {{rg, "lse_rhs"}, {g, "grad_phi"}, {g2, "grad2_phi"}})) return;
  hipLaunchKernelGGL(const_curv_kernel, grid_for(fb), block2d(), 0, stream(), fb, view(rhs, rg), view(px, g),
                     view(py, g), view(pz, g), view(xx, g2), view(xy, g2), view(xz, g2), view(yy, g2), view(yz, g2),
                     view(zz, g2), *b);
  finish(who);
}

// The reference loops k = klo_fb..khi_rhs_gb, j = jlo_fb..jhi_rhs_gb,
// i = ilo_rhs_gb..ihi_rhs_gb (lsm_level_set_evolution3d.f.in:514-516, :620-622);
// the same cells are updated here.
static Box quirky_loop_box(const Box& rg, const Box& fb) {
  return Box{rg.ilo, rg.ihi, fb.jlo, rg.jhi, fb.klo, rg.khi};
}

LSMHIP_EXPORT void lsm3daddconstprecomputedcurvtermtolserhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* kappa,
                                                             LSMHIP_BOXP(kg), const real_t* gmag, LSMHIP_BOXP(gg),
                                                             const real_t* b, LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, kappa, gmag);
  const char* who = "lsm3daddconstprecomputedcurvtermtolserhs_";
  Box rg = LSMHIP_BOX(rg), kg = LSMHIP_BOX(kg), gg = LSMHIP_BOX(gg);
  Box lb = quirky_loop_box(rg, LSMHIP_BOX(fb));
  if (lb.empty() || !boxes_ok(who, lb, {{rg, "lse_rhs"}, {kg, "kappa"}, {gg, "grad_phi_mag"}})) return;
  hipLaunchKernelGGL(precomputed_curv_kernel, grid_for(lb), block2d(), 0, stream(), lb, view(rhs, rg),
                     view(kappa, kg), view(gmag, gg), *b);
  finish(who);
}

LSMHIP_EXPORT void lsm3daddexternalandnormalveltermtolserhs_(real_t* rhs, LSMHIP_BOXP(rg), const real_t* pxp,
                                                             const real_t* pyp, const real_t* pzp, LSMHIP_BOXP(gp),
                                                             const real_t* pxm, const real_t* pym, const real_t* pzm,
                                                             LSMHIP_BOXP(gm), const real_t* vn, const real_t* vx,
                                                             const real_t* vy, const real_t* vz, LSMHIP_BOXP(vg),
                                                             LSMHIP_BOXP(fb)) {
  LSMHIP_REQUIRE_ARRAYS(rhs, pxp, pyp, pzp, pxm, pym, pzm, vn, vx, vy, vz);
  const char* who = "lsm3daddexternalandnormalveltermtolserhs_";
  Box rg = LSMHIP_BOX(rg), gp = LSMHIP_BOX(gp), gm = LSMHIP_BOX(gm), vg = LSMHIP_BOX(vg);
  Box lb = quirky_loop_box(rg, LSMHIP_BOX(fb));
  if (lb.empty() ||
      !boxes_ok(who, lb, {{rg, "lse_rhs"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}, {vg, "vel"}}))
    return;
  hipLaunchKernelGGL(ext_normal_kernel, grid_for(lb), block2d(), 0, stream(), lb, view(rhs, rg),
                     six(pxp, pyp, pzp, gp, pxm, pym, pzm, gm), view(vn, vg), view(vx, vg), view(vy, vg),
                     view(vz, vg));
  finish(who);
}

// ---------------------------------------------------------------------------
// Reinitialization and curvature
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3dcomputereinitializationeqnrhs_(real_t* out, LSMHIP_BOXP(rg), const real_t* phi,
                                                       LSMHIP_BOXP(pg), const real_t* phi0, LSMHIP_BOXP(p0g),
                                                       const real_t* pxp, const real_t* pyp, const real_t* pzp,
                                                       LSMHIP_BOXP(gp), const real_t* pxm, const real_t* pym,
                                                       const real_t* pzm, LSMHIP_BOXP(gm), LSMHIP_BOXP(fb),
                                                       const real_t* dx, const real_t* dy, const real_t* dz,
                                                       const int* use_phi0_for_sgn) {
  LSMHIP_REQUIRE_ARRAYS(out, phi, pxp, pyp, pzp, pxm, pym, pzm);
  if (*use_phi0_for_sgn == 1) LSMHIP_REQUIRE_ARRAYS(phi0);
  const char* who = "lsm3dcomputereinitializationeqnrhs_";
  Box fb = LSMHIP_BOX(fb), rg = LSMHIP_BOX(rg), pg = LSMHIP_BOX(pg), p0g = LSMHIP_BOX(p0g), gp = LSMHIP_BOX(gp),
      gm = LSMHIP_BOX(gm);
  if (fb.empty()) return;
  const int use0 = *use_phi0_for_sgn == 1;
  if (!boxes_ok(who, fb, {{rg, "reinit_rhs"}, {pg, "phi"}, {gp, "grad_phi_plus"}, {gm, "grad_phi_minus"}})) return;
  if (use0 && !boxes_ok(who, fb, {{p0g, "phi0"}})) return;
  real_t h = *dx > *dy ? *dx : *dy;
  h = h > *dz ? h : *dz;
  hipLaunchKernelGGL(reinit_kernel, grid_for_cells(fb, kPointNK), block2d(), 0, stream(), fb, view(out, rg), view(phi, pg),
                     view(phi0, p0g), six(pxp, pyp, pzp, gp, pxm, pym, pzm, gm), h * h, use0);
  finish(who);
}

LSMHIP_EXPORT void lsm3dcomputemeancurvatureorder2_(real_t* kappa, LSMHIP_BOXP(kg), const real_t* phi,
                                                    LSMHIP_BOXP(pg), const real_t* px, const real_t* py,
                                                    const real_t* pz, const real_t* gmag, LSMHIP_BOXP(gg),
                                                    LSMHIP_BOXP(fb), const real_t* dx, const real_t* dy,
                                                    const real_t* dz) {
  LSMHIP_REQUIRE_ARRAYS(kappa, phi, px, py, pz, gmag);
  const char* who = "lsm3dcomputemeancurvatureorder2_";
  Box fb = LSMHIP_BOX(fb), kg = LSMHIP_BOX(kg), pg = LSMHIP_BOX(pg), gg = LSMHIP_BOX(gg);
  if (fb.empty()) return;
  if (!boxes_ok(who, fb, {{kg, "kappa"}, {gg, "grad_phi"}}) || !check_box(pg, fb.widen(1, 1, 1), who, "phi") ||
      !check_size(pg, who))
    return;
  CurvFactors f{real_t(1.0) / *dx / *dx, real_t(1.0) / *dy / *dy, real_t(1.0) / *dz / *dz,
                real_t(0.25) / *dx / *dy, real_t(0.25) / *dx / *dz, real_t(0.25) / *dy / *dz};
  const march::Tiles tl = march::tiles_for(fb);
  hipLaunchKernelGGL(march::cross19_march_kernel<MeanCurvMarch>, march::grid_for(tl), march::block(), 0, stream(), fb,
                     view(phi, pg), tl,
                     MeanCurvMarch{view(kappa, kg), view(px, gg), view(py, gg), view(pz, gg), view(gmag, gg), f});
  finish(who);
}

// ---------------------------------------------------------------------------
// TVD Runge-Kutta (src/toolbox/lsm_tvd_runge_kutta3d.h)
// ---------------------------------------------------------------------------
LSMHIP_EXPORT void lsm3drk1step_(real_t* un, LSMHIP_BOXP(ng), const real_t* uc, LSMHIP_BOXP(cg), const real_t* rhs,
                                 LSMHIP_BOXP(rg), LSMHIP_BOXP(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs);
  rk_stage<1>("lsm3drk1step_", un, LSMHIP_BOX(ng), nullptr, LSMHIP_BOX(ng), uc, LSMHIP_BOX(cg), rhs, LSMHIP_BOX(rg),
              LSMHIP_BOX(fb), *dt);
}
LSMHIP_EXPORT void lsm3dtvdrk2stage1_(real_t* un, LSMHIP_BOXP(ng), const real_t* uc, LSMHIP_BOXP(cg),
                                      const real_t* rhs, LSMHIP_BOXP(rg), LSMHIP_BOXP(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs);
  rk_stage<1>("lsm3dtvdrk2stage1_", un, LSMHIP_BOX(ng), nullptr, LSMHIP_BOX(ng), uc, LSMHIP_BOX(cg), rhs,
              LSMHIP_BOX(rg), LSMHIP_BOX(fb), *dt);
}
LSMHIP_EXPORT void lsm3dtvdrk2stage2_(real_t* un, LSMHIP_BOXP(ng), const real_t* u1, LSMHIP_BOXP(g1),
                                      const real_t* uc, LSMHIP_BOXP(cg), const real_t* rhs, LSMHIP_BOXP(rg),
                                      LSMHIP_BOXP(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, u1, uc, rhs);
  rk_stage<2>("lsm3dtvdrk2stage2_", un, LSMHIP_BOX(ng), u1, LSMHIP_BOX(g1), uc, LSMHIP_BOX(cg), rhs, LSMHIP_BOX(rg),
              LSMHIP_BOX(fb), *dt);
}
LSMHIP_EXPORT void lsm3dtvdrk3stage1_(real_t* un, LSMHIP_BOXP(ng), const real_t* uc, LSMHIP_BOXP(cg),
                                      const real_t* rhs, LSMHIP_BOXP(rg), LSMHIP_BOXP(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, uc, rhs);
  rk_stage<1>("lsm3dtvdrk3stage1_", un, LSMHIP_BOX(ng), nullptr, LSMHIP_BOX(ng), uc, LSMHIP_BOX(cg), rhs,
              LSMHIP_BOX(rg), LSMHIP_BOX(fb), *dt);
}
LSMHIP_EXPORT void lsm3dtvdrk3stage2_(real_t* u2, LSMHIP_BOXP(g2), const real_t* u1, LSMHIP_BOXP(g1),
                                      const real_t* uc, LSMHIP_BOXP(cg), const real_t* rhs, LSMHIP_BOXP(rg),
                                      LSMHIP_BOXP(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(u2, u1, uc, rhs);
  rk_stage<32>("lsm3dtvdrk3stage2_", u2, LSMHIP_BOX(g2), u1, LSMHIP_BOX(g1), uc, LSMHIP_BOX(cg), rhs,
               LSMHIP_BOX(rg), LSMHIP_BOX(fb), *dt);
}
LSMHIP_EXPORT void lsm3dtvdrk3stage3_(real_t* un, LSMHIP_BOXP(ng), const real_t* u2, LSMHIP_BOXP(g2),
                                      const real_t* uc, LSMHIP_BOXP(cg), const real_t* rhs, LSMHIP_BOXP(rg),
                                      LSMHIP_BOXP(fb), const real_t* dt) {
  LSMHIP_REQUIRE_ARRAYS(un, u2, uc, rhs);
  rk_stage<33>("lsm3dtvdrk3stage3_", un, LSMHIP_BOX(ng), u2, LSMHIP_BOX(g2), uc, LSMHIP_BOX(cg), rhs,
               LSMHIP_BOX(rg), LSMHIP_BOX(fb), *dt);
}
