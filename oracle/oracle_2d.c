/*
 * oracle_2d.c -- TEST INFRASTRUCTURE ONLY (see lsm_oracle.h).
 *
 * The 2D twins of the whole-grid path: HJ-ENO1/2/3 and HJ-WENO5 derivatives
 * (plain and velocity-upwinded), central gradients, Laplacian, averaged and
 * magnitude of the gradient, the level-set RHS terms, TVD Runge-Kutta stages,
 * reinitialization and orthogonalization RHS, mean curvature (orders 2 and 4,
 * and from a given Hessian), unit normals, field-extension RHS, CFL time steps,
 * max-norm difference, smoothed area / perimeter / volume and surface
 * integrals, and the ghost-cell fills.  Each function cites the reference
 * routine whose operation order it keeps.  2D arrays are column-major over
 * (i, j); they are viewed here as 3D arrays of one plane (k = 0), so the
 * per-axis stencil drivers of oracle_wholegrid.c serve both dimensions.
 */
#include "lsm_oracle.h"

#define BOXP2(n) const int *ilo_##n, const int *ihi_##n, const int *jlo_##n, const int *jhi_##n

static orc_arr v2(const double *p, const int *ilo, const int *ihi, const int *jlo, const int *jhi) {
  orc_arr a;
  a.p = (double *)p;
  a.ilo = *ilo;
  a.jlo = *jlo;
  a.klo = 0;
  a.nx = (long)(*ihi - *ilo + 1);
  a.nxy = a.nx * (long)(*jhi - *jlo + 1);
  return a;
}
static obox b2(const int *ilo, const int *ihi, const int *jlo, const int *jhi) {
  obox b = {*ilo, *ihi, *jlo, *jhi, 0, 0};
  return b;
}
#define V2(p, n) v2((p), ilo_##n, ihi_##n, jlo_##n, jhi_##n)
#define B2(n) b2(ilo_##n, ihi_##n, jlo_##n, jhi_##n)
#define A2(a, i, j) AT(a, i, j, 0)
#define FOR2(b)                                \
  for (int j = (b).jlo; j <= (b).jhi; j++)     \
    for (int i = (b).ilo; i <= (b).ihi; i++)

/* ---------------------------------------------------------------------------
 * Spatial derivatives, src/toolbox/lsm_spatial_derivatives2d.f.in
 * ------------------------------------------------------------------------- */
#define INV2(inv) double inv[3] = {1.0 / *dx, 1.0 / *dy, 0.0}

/* lsm2dHJENO1 :165-272 */
void oracle_lsm2dhjeno1(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                        const double *phi, BOXP2(pg), double *D1, BOXP2(d1), BOXP2(fb), const double *dx,
                        const double *dy) {
  INV2(inv);
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)}, none = {0};
  obox nb = {0};
  orc_hj_plus_minus(2, 1, P, M, V2(phi, pg), V2(D1, d1), B2(d1), none, nb, none, nb, B2(fb), inv);
}
/* lsm2dHJENO2 :296-454 */
void oracle_lsm2dhjeno2(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                        const double *phi, BOXP2(pg), double *D1, BOXP2(d1), double *D2, BOXP2(d2), BOXP2(fb),
                        const double *dx, const double *dy) {
  INV2(inv);
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)}, none = {0};
  obox nb = {0};
  orc_hj_plus_minus(2, 2, P, M, V2(phi, pg), V2(D1, d1), B2(d1), V2(D2, d2), B2(d2), none, nb, B2(fb), inv);
}
/* lsm2dHJENO3 :479-733 */
void oracle_lsm2dhjeno3(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                        const double *phi, BOXP2(pg), double *D1, BOXP2(d1), double *D2, BOXP2(d2), double *D3,
                        BOXP2(d3), BOXP2(fb), const double *dx, const double *dy) {
  INV2(inv);
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)};
  orc_hj_plus_minus(2, 3, P, M, V2(phi, pg), V2(D1, d1), B2(d1), V2(D2, d2), B2(d2), V2(D3, d3), B2(d3), B2(fb),
                    inv);
}
/* lsm2dHJWENO5 :756-1051 */
void oracle_lsm2dhjweno5(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                         const double *phi, BOXP2(pg), double *D1, BOXP2(d1), BOXP2(fb), const double *dx,
                         const double *dy) {
  INV2(inv);
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)}, none = {0};
  obox nb = {0};
  orc_hj_plus_minus(2, 5, P, M, V2(phi, pg), V2(D1, d1), B2(d1), none, nb, none, nb, B2(fb), inv);
}

/* lsm2dUpwindHJENO1 :1069-1195 */
void oracle_lsm2dupwindhjeno1(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg), const double *vx,
                              const double *vy, BOXP2(vg), double *D1, BOXP2(d1), BOXP2(fb), const double *dx,
                              const double *dy) {
  INV2(inv);
  orc_arr O[2] = {V2(px, g), V2(py, g)}, V[2] = {V2(vx, vg), V2(vy, vg)}, none = {0};
  obox nb = {0};
  orc_hj_upwind(2, 1, O, V2(phi, pg), V, V2(D1, d1), B2(d1), none, nb, none, nb, B2(fb), inv);
}
/* lsm2dUpwindHJENO2 :1214-1396 */
void oracle_lsm2dupwindhjeno2(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg), const double *vx,
                              const double *vy, BOXP2(vg), double *D1, BOXP2(d1), double *D2, BOXP2(d2),
                              BOXP2(fb), const double *dx, const double *dy) {
  INV2(inv);
  orc_arr O[2] = {V2(px, g), V2(py, g)}, V[2] = {V2(vx, vg), V2(vy, vg)}, none = {0};
  obox nb = {0};
  orc_hj_upwind(2, 2, O, V2(phi, pg), V, V2(D1, d1), B2(d1), V2(D2, d2), B2(d2), none, nb, B2(fb), inv);
}
/* lsm2dUpwindHJENO3 :1416-1681 */
void oracle_lsm2dupwindhjeno3(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg), const double *vx,
                              const double *vy, BOXP2(vg), double *D1, BOXP2(d1), double *D2, BOXP2(d2),
                              double *D3, BOXP2(d3), BOXP2(fb), const double *dx, const double *dy) {
  INV2(inv);
  orc_arr O[2] = {V2(px, g), V2(py, g)}, V[2] = {V2(vx, vg), V2(vy, vg)};
  orc_hj_upwind(2, 3, O, V2(phi, pg), V, V2(D1, d1), B2(d1), V2(D2, d2), B2(d2), V2(D3, d3), B2(d3), B2(fb),
                inv);
}
/* lsm2dUpwindHJWENO5 :1699-1941 */
void oracle_lsm2dupwindhjweno5(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg), const double *vx,
                               const double *vy, BOXP2(vg), double *D1, BOXP2(d1), BOXP2(fb), const double *dx,
                               const double *dy) {
  INV2(inv);
  orc_arr O[2] = {V2(px, g), V2(py, g)}, V[2] = {V2(vx, vg), V2(vy, vg)}, none = {0};
  obox nb = {0};
  orc_hj_upwind(2, 5, O, V2(phi, pg), V, V2(D1, d1), B2(d1), none, nb, none, nb, B2(fb), inv);
}

/* lsm2dCentralGradOrder2 :1957-2005 */
void oracle_lsm2dcentralgradorder2(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg), BOXP2(fb),
                                   const double *dx, const double *dy) {
  orc_arr X = V2(px, g), Y = V2(py, g), F = V2(phi, pg);
  double fx = 0.5 / *dx, fy = 0.5 / *dy;
  obox b = B2(fb);
  FOR2(b) {
    A2(X, i, j) = (A2(F, i + 1, j) - A2(F, i - 1, j)) * fx;
    A2(Y, i, j) = (A2(F, i, j + 1) - A2(F, i, j - 1)) * fy;
  }
}
/* lsm2dCentralGradOrder4 :2021-2075 */
void oracle_lsm2dcentralgradorder4(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg), BOXP2(fb),
                                   const double *dx, const double *dy) {
  orc_arr X = V2(px, g), Y = V2(py, g), F = V2(phi, pg);
  double fx = 0.0833333333333333333333 / *dx, fy = 0.0833333333333333333333 / *dy;
  obox b = B2(fb);
  FOR2(b) {
    A2(X, i, j) = (-A2(F, i + 2, j) + 8.0 * A2(F, i + 1, j) + A2(F, i - 2, j) - 8.0 * A2(F, i - 1, j)) * fx;
    A2(Y, i, j) = (-A2(F, i, j + 2) + 8.0 * A2(F, i, j + 1) + A2(F, i, j - 2) - 8.0 * A2(F, i, j - 1)) * fy;
  }
}
/* lsm2dLaplacianOrder2 :2091-2140 */
void oracle_lsm2dlaplacianorder2(double *lap, BOXP2(lg), const double *phi, BOXP2(pg), BOXP2(fb), const double *dx,
                                 const double *dy) {
  orc_arr L = V2(lap, lg), F = V2(phi, pg);
  double ax = 1.0 / *dx / *dx, ay = 1.0 / *dy / *dy;
  obox b = B2(fb);
  FOR2(b) {
    double c = A2(F, i, j);
    A2(L, i, j) = ax * (A2(F, i + 1, j) - 2.0 * c + A2(F, i - 1, j)) + ay * (A2(F, i, j + 1) - 2.0 * c + A2(F, i, j - 1));
  }
}
/* lsm2dAverageGradPhi :2322-2381 */
void oracle_lsm2daveragegradphi(double *px, double *py, BOXP2(g), const double *pxp, const double *pyp, BOXP2(gp),
                                const double *pxm, const double *pym, BOXP2(gm), BOXP2(fb)) {
  orc_arr X = V2(px, g), Y = V2(py, g), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  obox b = B2(fb);
  FOR2(b) {
    A2(X, i, j) = 0.5 * (A2(XP, i, j) + A2(XM, i, j));
    A2(Y, i, j) = 0.5 * (A2(YP, i, j) + A2(YM, i, j));
  }
}
/* lsm2dGradientMagnitude :2395-2448 -- squared components below zero_tol are
 * dropped and a magnitude below zero_tol is 0 (the 3D routine uses 1e-8). */
void oracle_lsm2dgradientmagnitude(const double *px, const double *py, double *mag, BOXP2(g), BOXP2(fb)) {
  orc_arr X = V2(px, g), Y = V2(py, g), M = V2(mag, g);
  obox b = B2(fb);
  FOR2(b) {
    double xs = A2(X, i, j) * A2(X, i, j), ys = A2(Y, i, j) * A2(Y, i, j);
    if (xs < ORC_ZERO_TOL) xs = 0.0;
    if (ys < ORC_ZERO_TOL) ys = 0.0;
    double t = sqrt(xs + ys);
    A2(M, i, j) = t < ORC_ZERO_TOL ? 0.0 : t;
  }
}

/* ---------------------------------------------------------------------------
 * Level-set RHS terms, src/toolbox/lsm_level_set_evolution2d.f.in
 * ------------------------------------------------------------------------- */
/* :16-42 -- the whole ghost box */
void oracle_lsm2dzerooutlevelseteqnrhs(double *rhs, BOXP2(rg)) {
  orc_arr R = V2(rhs, rg);
  obox b = B2(rg);
  FOR2(b) A2(R, i, j) = 0.0;
}
/* :47-96 */
void oracle_lsm2daddadvectiontermtolserhs(double *rhs, BOXP2(rg), const double *px, const double *py, BOXP2(g),
                                          const double *vx, const double *vy, BOXP2(vg), BOXP2(fb)) {
  orc_arr R = V2(rhs, rg), X = V2(px, g), Y = V2(py, g), U = V2(vx, vg), V = V2(vy, vg);
  obox b = B2(fb);
  FOR2(b) A2(R, i, j) = A2(R, i, j) - (A2(U, i, j) * A2(X, i, j) + A2(V, i, j) * A2(Y, i, j));
}
/* Godunov |grad phi|^2 of :157-166 and :240-249 (V > 0 and V <= 0 branches). */
static double godunov2(int vpos, double xp, double yp, double xm, double ym) {
  if (vpos)
    return orc_max(orc_sq(orc_max(xm, 0.0)), orc_sq(orc_min(xp, 0.0))) +
           orc_max(orc_sq(orc_max(ym, 0.0)), orc_sq(orc_min(yp, 0.0)));
  return orc_max(orc_sq(orc_min(xm, 0.0)), orc_sq(orc_max(xp, 0.0))) +
         orc_max(orc_sq(orc_min(ym, 0.0)), orc_sq(orc_max(yp, 0.0)));
}
/* :101-184 */
void oracle_lsm2daddnormalveltermtolserhs(double *rhs, BOXP2(rg), const double *pxp, const double *pyp, BOXP2(gp),
                                          const double *pxm, const double *pym, BOXP2(gm), const double *vn,
                                          BOXP2(vg), BOXP2(fb)) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm), N = V2(vn, vg);
  obox b = B2(fb);
  FOR2(b) {
    double v = A2(N, i, j);
    if (fabs(v) >= ORC_ZERO_TOL) {
      double s = godunov2(v > 0.0, A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j));
      A2(R, i, j) = A2(R, i, j) - v * sqrt(s);
    }
  }
}
/* :189-268 */
void oracle_lsm2daddconstnormalveltermtolserhs(double *rhs, BOXP2(rg), const double *pxp, const double *pyp,
                                               BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm),
                                               const double *vn, BOXP2(fb)) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  double v = *vn;
  obox b = B2(fb);
  if (fabs(v) < ORC_ZERO_TOL) return;
  FOR2(b) {
    double s = godunov2(v > 0.0, A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j));
    A2(R, i, j) = A2(R, i, j) - v * sqrt(s);
  }
}
/* :273-341 -- b * kappa |grad phi| with kappa |grad phi| = numerator / |grad phi|^2 */
void oracle_lsm2daddconstcurvtermtolserhs(double *rhs, BOXP2(rg), const double *px, const double *py, BOXP2(g),
                                          const double *pxx, const double *pxy, const double *pyy, BOXP2(g2),
                                          const double *b_, BOXP2(fb)) {
  orc_arr R = V2(rhs, rg), X = V2(px, g), Y = V2(py, g), XX = V2(pxx, g2), XY = V2(pxy, g2), YY = V2(pyy, g2);
  double bb = *b_;
  obox b = B2(fb);
  FOR2(b) {
    double x = A2(X, i, j), y = A2(Y, i, j);
    double g = x * x + y * y, curv;
    if (g < ORC_ZERO_TOL) {
      curv = 0.0;
    } else {
      curv = A2(XX, i, j) * y * y + A2(YY, i, j) * x * x - 2.0 * A2(XY, i, j) * x * y;
      curv = curv / g;
    }
    A2(R, i, j) = A2(R, i, j) + bb * curv;
  }
}
/* :345-393 -- the loops run j = jlo_rhs_fb..jhi_rhs_gb, i = ilo_rhs_gb..ihi_rhs_gb */
void oracle_lsm2daddconstprecomputedcurvtermtolserhs(double *rhs, BOXP2(rg), const double *kappa, BOXP2(kg),
                                                     const double *gmag, BOXP2(gg), const double *b_, BOXP2(fb)) {
  orc_arr R = V2(rhs, rg), K = V2(kappa, kg), G = V2(gmag, gg);
  double bb = *b_;
  for (int j = *jlo_fb; j <= *jhi_rg; j++)
    for (int i = *ilo_rg; i <= *ihi_rg; i++) A2(R, i, j) = A2(R, i, j) + bb * A2(K, i, j) * A2(G, i, j);
}
/* Per-axis Godunov selection of :464-503 for H1 = u + V phi_pm. */
static double ext_pick(double u, double vn, double pp, double pm) {
  double hp = u + vn * pp, hm = u + vn * pm;
  if (hp > 0.0) {
    if (hm > 0.0) return pm;
    return -(u / vn);
  }
  if (hm < 0.0) return pp;
  if (fabs(hp) > fabs(hm)) return pp;
  return pm;
}
/* :399-522 -- over the fill box; nothing is added where |V| < zero_tol */
void oracle_lsm2daddexternalandnormalveltermtolserhs(double *rhs, BOXP2(rg), const double *pxp, const double *pyp,
                                                     BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm),
                                                     const double *vn, const double *vx, const double *vy,
                                                     BOXP2(vg), BOXP2(fb)) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  orc_arr N = V2(vn, vg), U = V2(vx, vg), V = V2(vy, vg);
  obox b = B2(fb);
  FOR2(b) {
    double n = A2(N, i, j), u = A2(U, i, j), v = A2(V, i, j);
    double px = ext_pick(u, n, A2(XP, i, j), A2(XM, i, j));
    double py = ext_pick(v, n, A2(YP, i, j), A2(YM, i, j));
    double s = px * px + py * py;
    if (fabs(n) >= ORC_ZERO_TOL) A2(R, i, j) = A2(R, i, j) - u * px - v * py - n * sqrt(s);
  }
}

/* ---------------------------------------------------------------------------
 * TVD Runge-Kutta stages, src/toolbox/lsm_tvd_runge_kutta2d.f
 * ------------------------------------------------------------------------- */
/* lsm2dRK1Step :32-76 (TVD-RK2 stage 1, :96-142, and RK3 stage 1, :232-278, call it) */
void oracle_lsm2drk1step(double *un, BOXP2(ng), const double *uc, BOXP2(cg), const double *rhs, BOXP2(rg),
                         BOXP2(fb), const double *dt) {
  orc_arr N = V2(un, ng), C = V2(uc, cg), R = V2(rhs, rg);
  double t = *dt;
  obox b = B2(fb);
  FOR2(b) A2(N, i, j) = A2(C, i, j) + t * A2(R, i, j);
}
void oracle_lsm2dtvdrk2stage1(double *un, BOXP2(ng), const double *uc, BOXP2(cg), const double *rhs, BOXP2(rg),
                              BOXP2(fb), const double *dt) {
  oracle_lsm2drk1step(un, ilo_ng, ihi_ng, jlo_ng, jhi_ng, uc, ilo_cg, ihi_cg, jlo_cg, jhi_cg, rhs, ilo_rg, ihi_rg,
                      jlo_rg, jhi_rg, ilo_fb, ihi_fb, jlo_fb, jhi_fb, dt);
}
void oracle_lsm2dtvdrk3stage1(double *un, BOXP2(ng), const double *uc, BOXP2(cg), const double *rhs, BOXP2(rg),
                              BOXP2(fb), const double *dt) {
  oracle_lsm2drk1step(un, ilo_ng, ihi_ng, jlo_ng, jhi_ng, uc, ilo_cg, ihi_cg, jlo_cg, jhi_cg, rhs, ilo_rg, ihi_rg,
                      jlo_rg, jhi_rg, ilo_fb, ihi_fb, jlo_fb, jhi_fb, dt);
}
/* lsm2dTVDRK2Stage2 :160-212: 0.5 (u + u1 + dt L) */
void oracle_lsm2dtvdrk2stage2(double *un, BOXP2(ng), const double *u1, BOXP2(g1), const double *uc, BOXP2(cg),
                              const double *rhs, BOXP2(rg), BOXP2(fb), const double *dt) {
  orc_arr N = V2(un, ng), S1 = V2(u1, g1), C = V2(uc, cg), R = V2(rhs, rg);
  double t = *dt;
  obox b = B2(fb);
  FOR2(b) A2(N, i, j) = 0.5 * (A2(C, i, j) + A2(S1, i, j) + t * A2(R, i, j));
}
/* lsm2dTVDRK3Stage2 :296-348: 0.75 u + 0.25 (u1 + dt L) */
void oracle_lsm2dtvdrk3stage2(double *u2, BOXP2(g2), const double *u1, BOXP2(g1), const double *uc, BOXP2(cg),
                              const double *rhs, BOXP2(rg), BOXP2(fb), const double *dt) {
  orc_arr N = V2(u2, g2), S1 = V2(u1, g1), C = V2(uc, cg), R = V2(rhs, rg);
  double t = *dt;
  obox b = B2(fb);
  FOR2(b) A2(N, i, j) = 0.75 * A2(C, i, j) + 0.25 * (A2(S1, i, j) + t * A2(R, i, j));
}
/* lsm2dTVDRK3Stage3 :366-421: 1/3 u + 2/3 (u2 + dt L) */
void oracle_lsm2dtvdrk3stage3(double *un, BOXP2(ng), const double *u2, BOXP2(g2), const double *uc, BOXP2(cg),
                              const double *rhs, BOXP2(rg), BOXP2(fb), const double *dt) {
  orc_arr N = V2(un, ng), S2 = V2(u2, g2), C = V2(uc, cg), R = V2(rhs, rg);
  double t = *dt;
  obox b = B2(fb);
  FOR2(b) A2(N, i, j) = (1.0 / 3.0) * A2(C, i, j) + (2.0 / 3.0) * (A2(S2, i, j) + t * A2(R, i, j));
}

/* ---------------------------------------------------------------------------
 * Reinitialization and orthogonalization, src/reinitialization/lsm_reinitialization2d.f.in
 * ------------------------------------------------------------------------- */
/* :48-227 -- Godunov |grad phi| with the smoothed sign of phi (|grad phi|^2
 * max(dx,dy)^2 in the smoothing) or of phi0 (max(dx,dy)^2 only); 0 where
 * |sign source| <= zero_tol (strict: the 3D routine tests >=). */
void oracle_lsm2dcomputereinitializationeqnrhs(double *rr, BOXP2(rg), const double *phi, BOXP2(pg),
                                               const double *phi0, BOXP2(p0g), const double *pxp, const double *pyp,
                                               BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm),
                                               BOXP2(fb), const double *dx, const double *dy,
                                               const int *use_phi0_for_sgn) {
  orc_arr R = V2(rr, rg), F = V2(phi, pg), F0 = V2(phi0, p0g);
  orc_arr XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  double h = orc_max(*dx, *dy), dx_sq = h * h;
  int use0 = (*use_phi0_for_sgn == 1);
  obox b = B2(fb);
  FOR2(b) {
    double s = use0 ? A2(F0, i, j) : A2(F, i, j);
    double gp[2] = {A2(XP, i, j), A2(YP, i, j)}, gm[2] = {A2(XM, i, j), A2(YM, i, j)}, star[2];
    for (int d = 0; d < 2; d++) {
      double p, m;
      if (s > 0.0) {
        p = orc_max(-gp[d], 0.0);
        m = orc_max(gm[d], 0.0);
      } else {
        p = orc_max(gp[d], 0.0);
        m = orc_max(-gm[d], 0.0);
      }
      star[d] = orc_max(p, m);
    }
    if (fabs(s) > ORC_ZERO_TOL) {
      double n2 = star[0] * star[0] + star[1] * star[1];
      double sgn = use0 ? s / sqrt(s * s + dx_sq) : s / sqrt(s * s + n2 * dx_sq);
      A2(R, i, j) = sgn * (1.0 - sqrt(n2));
    } else {
      A2(R, i, j) = 0.0;
    }
  }
}
/* :255-386 -- the smoothed sign of psi is psi/sqrt(psi^2 + max(dx,dy)^2)
 * (the 3D routine also scales max(dx,dy,dz)^2 by |grad psi|^2). */
void oracle_lsm2dcomputeorthogonalizationeqnrhs(double *rhs, BOXP2(rg), const double *pxp, const double *pyp,
                                                BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm),
                                                const double *psi, BOXP2(sg), const double *qxp, const double *qyp,
                                                BOXP2(qp), const double *qxm, const double *qym, BOXP2(qm),
                                                BOXP2(fb), const double *dx, const double *dy) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm), S = V2(psi, sg);
  orc_arr QXP = V2(qxp, qp), QYP = V2(qyp, qp), QXM = V2(qxm, qm), QYM = V2(qym, qm);
  double h = orc_max(*dx, *dy), dx_sq = h * h;
  obox b = B2(fb);
  FOR2(b) {
    double g1 = 0.5 * (A2(QXP, i, j) + A2(QXM, i, j));
    double g2 = 0.5 * (A2(QYP, i, j) + A2(QYM, i, j));
    double ng = sqrt(g1 * g1 + g2 * g2);
    double s = A2(S, i, j);
    if (fabs(s) > ORC_ZERO_TOL && ng > ORC_ZERO_TOL) {
      double sgn = s / sqrt(s * s + dx_sq);
      double a1 = sgn * g1, a2 = sgn * g2;
      A2(R, i, j) = -1.0 / ng *
                    (orc_max(a1, 0.0) * A2(XM, i, j) + orc_min(a1, 0.0) * A2(XP, i, j) +
                     orc_max(a2, 0.0) * A2(YM, i, j) + orc_min(a2, 0.0) * A2(YP, i, j));
    } else {
      A2(R, i, j) = 0.0;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Curvature, src/geometry/lsm_curvature2d.f.in and lsm_geometry2d.f.in
 * ------------------------------------------------------------------------- */
/* lsm2dComputeMeanCurvatureOrder2 (curvature2d :15-92): numerator over |grad phi|^3 */
void oracle_lsm2dcomputemeancurvatureorder2(double *kappa, BOXP2(kg), const double *phi, BOXP2(pg), const double *px,
                                            const double *py, const double *gmag, BOXP2(gg), BOXP2(fb),
                                            const double *dx, const double *dy) {
  orc_arr K = V2(kappa, kg), F = V2(phi, pg), X = V2(px, gg), Y = V2(py, gg), G = V2(gmag, gg);
  double fxx = 1.0 / *dx / *dx, fyy = 1.0 / *dy / *dy, fxy = 0.25 / *dx / *dy;
  obox b = B2(fb);
  FOR2(b) {
    double c = A2(F, i, j);
    double xx = (A2(F, i + 1, j) - 2.0 * c + A2(F, i - 1, j)) * fxx;
    double yy = (A2(F, i, j + 1) - 2.0 * c + A2(F, i, j - 1)) * fyy;
    double xy = (A2(F, i + 1, j + 1) - A2(F, i + 1, j - 1) - A2(F, i - 1, j + 1) + A2(F, i - 1, j - 1)) * fxy;
    double g = A2(G, i, j), den = g * g * g;
    if (den < ORC_ZERO_TOL) {
      A2(K, i, j) = 0.0;
    } else {
      double x = A2(X, i, j), y = A2(Y, i, j);
      A2(K, i, j) = (xx * y * y + yy * x * x - 2.0 * xy * x * y) / den;
    }
  }
}
/* lsm2dComputeMeanCurvatureOrder4 (curvature2d :252-337): fourth-order second
 * differences of phi, phi_xy as the fourth-order difference of phi_x along y */
void oracle_lsm2dcomputemeancurvatureorder4(double *kappa, BOXP2(kg), const double *phi, BOXP2(pg), const double *px,
                                            const double *py, const double *gmag, BOXP2(gg), BOXP2(fb),
                                            const double *dx, const double *dy) {
  orc_arr K = V2(kappa, kg), F = V2(phi, pg), X = V2(px, gg), Y = V2(py, gg), G = V2(gmag, gg);
  const double c12 = 0.0833333333333333333333;
  double fxx = c12 / *dx / *dx, fyy = c12 / *dy / *dy, fy = c12 / *dy;
  obox b = B2(fb);
  FOR2(b) {
    double c = A2(F, i, j);
    double xx = (-A2(F, i + 2, j) + 16.0 * A2(F, i + 1, j) - 30.0 * c - A2(F, i - 2, j) + 16.0 * A2(F, i - 1, j)) * fxx;
    double yy = (-A2(F, i, j + 2) + 16.0 * A2(F, i, j + 1) - 30.0 * c - A2(F, i, j - 2) + 16.0 * A2(F, i, j - 1)) * fyy;
    double xy = (-A2(X, i, j + 2) + 8.0 * A2(X, i, j + 1) + A2(X, i, j - 2) - 8.0 * A2(X, i, j - 1)) * fy;
    double g = A2(G, i, j), den = g * g * g;
    if (den < ORC_ZERO_TOL) {
      A2(K, i, j) = 0.0;
    } else {
      double x = A2(X, i, j), y = A2(Y, i, j);
      A2(K, i, j) = (xx * y * y + yy * x * x - 2.0 * xy * x * y) / den;
    }
  }
}
/* lsm2dComputeMeanCurvature (geometry2d :218-286): from a given Hessian, over
 * pow(|grad phi|^2, 1.5) */
void oracle_lsm2dcomputemeancurvature(double *kappa, BOXP2(kg), const double *px, const double *py, BOXP2(gg),
                                      const double *pxx, const double *pxy, const double *pyy, BOXP2(hg), BOXP2(fb)) {
  orc_arr K = V2(kappa, kg), X = V2(px, gg), Y = V2(py, gg), XX = V2(pxx, hg), XY = V2(pxy, hg), YY = V2(pyy, hg);
  obox b = B2(fb);
  FOR2(b) {
    double x = A2(X, i, j), y = A2(Y, i, j), ngs = x * x + y * y;
    if (ngs < ORC_ZERO_TOL)
      A2(K, i, j) = 0.0;
    else
      A2(K, i, j) = (A2(XX, i, j) * y * y - 2.0 * A2(XY, i, j) * x * y + A2(YY, i, j) * x * x) / pow(ngs, 1.5);
  }
}
/* lsm2dComputeUnitNormal (geometry2d :22-83) */
void oracle_lsm2dcomputeunitnormal(double *nx, double *ny, BOXP2(ng), const double *px, const double *py, BOXP2(gg),
                                   BOXP2(fb)) {
  orc_arr NX = V2(nx, ng), NY = V2(ny, ng), X = V2(px, gg), Y = V2(py, gg);
  obox b = B2(fb);
  FOR2(b) {
    double x = A2(X, i, j), y = A2(Y, i, j), norm = sqrt(x * x + y * y);
    if (norm >= ORC_ZERO_TOL) {
      double inv = 1.0 / norm;
      A2(NX, i, j) = x * inv;
      A2(NY, i, j) = y * inv;
    } else {
      A2(NX, i, j) = 1.0;
      A2(NY, i, j) = 0.0;
    }
  }
}
/* lsm2dComputeSignedUnitNormal (geometry2d :103-195) */
void oracle_lsm2dcomputesignedunitnormal(double *nx, double *ny, BOXP2(ng), const double *px, const double *py,
                                         BOXP2(gg), const double *phi, BOXP2(pg), BOXP2(fb), const double *dx,
                                         const double *dy) {
  orc_arr NX = V2(nx, ng), NY = V2(ny, ng), X = V2(px, gg), Y = V2(py, gg), P = V2(phi, pg);
  double h = orc_max(*dx, *dy), dx_sq = h * h;
  obox b = B2(fb);
  FOR2(b) {
    double c = A2(P, i, j);
    if (fabs(c) > ORC_ZERO_TOL) {
      double x = A2(X, i, j), y = A2(Y, i, j), ngs = x * x + y * y;
      if (ngs >= ORC_ZERO_TOL) {
        double sgn = c / sqrt(c * c + ngs * dx_sq), inv = 1.0 / sqrt(ngs);
        A2(NX, i, j) = sgn * x * inv;
        A2(NY, i, j) = sgn * y * inv;
      } else {
        A2(NX, i, j) = 1.0;
        A2(NY, i, j) = 0.0;
      }
    } else {
      A2(NX, i, j) = 0.0;
      A2(NY, i, j) = 0.0;
    }
  }
}

/* lsm2dComputeFieldExtensionEqnRHS, src/field_extension/lsm_field_extension2d.f:31-125 */
void oracle_lsm2dcomputefieldextensioneqnrhs(double *rhs, BOXP2(rg), const double *S_, BOXP2(sg), const double *phi,
                                             BOXP2(pg), const double *sx, const double *sy, BOXP2(ug),
                                             const double *nx, const double *ny, BOXP2(ng), BOXP2(fb),
                                             const double *dx, const double *dy) {
  orc_arr R = V2(rhs, rg), P = V2(phi, pg), SX = V2(sx, ug), SY = V2(sy, ug), NX = V2(nx, ng), NY = V2(ny, ng);
  double cutoff = 3.0 * orc_max(*dx, *dy);
  obox b = B2(fb);
  (void)S_;
  FOR2(b) {
    double c = A2(P, i, j);
    int away = fabs(c) > cutoff || (c * A2(P, i - 1, j) > 0.0 && c * A2(P, i + 1, j) > 0.0 &&
                                    c * A2(P, i, j - 1) > 0.0 && c * A2(P, i, j + 1) > 0.0);
    A2(R, i, j) = away ? -(A2(NX, i, j) * A2(SX, i, j) + A2(NY, i, j) * A2(SY, i, j)) : 0.0;
  }
}

/* ---------------------------------------------------------------------------
 * Reductions, src/toolbox/lsm_math_utils2d.f.in and src/geometry/lsm_geometry2d.f.in
 * ------------------------------------------------------------------------- */
/* lsm2dMaxNormDiff :15-64 */
void oracle_lsm2dmaxnormdiff(double *out, const double *f1, BOXP2(g1), const double *f2, BOXP2(g2), BOXP2(ib)) {
  orc_arr A = V2(f1, g1), B = V2(f2, g2);
  obox b = B2(ib);
  double m = fabs(A2(A, b.ilo, b.jlo) - A2(B, b.ilo, b.jlo));
  FOR2(b) {
    double d = fabs(A2(A, i, j) - A2(B, i, j));
    if (d > m) m = d;
  }
  *out = m;
}
/* lsm2dComputeStableAdvectionDt :218-280 */
void oracle_lsm2dcomputestableadvectiondt(double *dt, const double *vx, const double *vy, BOXP2(vg), BOXP2(ib),
                                          const double *dx, const double *dy, const double *cfl) {
  orc_arr U = V2(vx, vg), V = V2(vy, vg);
  double m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy;
  obox b = B2(ib);
  FOR2(b) {
    double c = fabs(A2(U, i, j)) * ix + fabs(A2(V, i, j)) * iy;
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}
static double normal_h2(double absv, double xp, double yp, double xm, double ym, double mdx2, double ix, double iy) {
  double x = orc_max(fabs(xp), fabs(xm)), y = orc_max(fabs(yp), fabs(ym));
  return absv / sqrt(x * x + y * y + mdx2) * (x * ix + y * iy);
}
/* lsm2dComputeStableNormalVelDt :285-376 */
void oracle_lsm2dcomputestablenormalveldt(double *dt, const double *vn, BOXP2(vg), const double *pxp,
                                          const double *pyp, BOXP2(gp), const double *pxm, const double *pym,
                                          BOXP2(gm), BOXP2(ib), const double *dx, const double *dy,
                                          const double *cfl) {
  orc_arr N = V2(vn, vg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  double h = orc_max(*dx, *dy), mdx2 = h * h, m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy;
  obox b = B2(ib);
  FOR2(b) {
    double c = normal_h2(fabs(A2(N, i, j)), A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j), mdx2, ix, iy);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}
/* lsm2dComputeStableConstNormalVelDt :382-468 */
void oracle_lsm2dcomputestableconstnormalveldt(double *dt, const double *vn, const double *pxp, const double *pyp,
                                               BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm), BOXP2(ib),
                                               const double *dx, const double *dy, const double *cfl) {
  orc_arr XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  double h = orc_max(*dx, *dy), mdx2 = h * h, m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy, av = fabs(*vn);
  obox b = B2(ib);
  FOR2(b) {
    double c = normal_h2(av, A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j), mdx2, ix, iy);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

#define ORC_PI 3.14159265358979323846
#define ORC_ONE_OVER_PI 0.31830988618379

/* lsm2dVolumeIntegralPhiLessThanZero (math_utils2d :473-544), serial sum */
void oracle_lsm2dvolumeintegralphilessthanzero(double *intF, const double *F, BOXP2(fg), const double *phi,
                                               BOXP2(pg), BOXP2(ib), const double *dx, const double *dy,
                                               const double *eps) {
  orc_arr FF = V2(F, fg), P = V2(phi, pg);
  double dV = *dx * *dy, e = *eps, s = 0.0;
  obox b = B2(ib);
  FOR2(b) {
    double p = A2(P, i, j), pe = p / e;
    if (p < -e) s = s + A2(FF, i, j) * dV;
    else if (p < e) s = s + 0.5 * (1.0 - pe - ORC_ONE_OVER_PI * sin(ORC_PI * pe)) * A2(FF, i, j) * dV;
  }
  *intF = s;
}
/* lsm2dVolumeIntegralPhiGreaterThanZero (math_utils2d :549-620) */
void oracle_lsm2dvolumeintegralphigreaterthanzero(double *intF, const double *F, BOXP2(fg), const double *phi,
                                                  BOXP2(pg), BOXP2(ib), const double *dx, const double *dy,
                                                  const double *eps) {
  orc_arr FF = V2(F, fg), P = V2(phi, pg);
  double dV = *dx * *dy, e = *eps, s = 0.0;
  obox b = B2(ib);
  FOR2(b) {
    double p = A2(P, i, j), pe = p / e;
    if (p > e) s = s + A2(FF, i, j) * dV;
    else if (p > -e) s = s + 0.5 * (1.0 + pe + ORC_ONE_OVER_PI * sin(ORC_PI * pe)) * A2(FF, i, j) * dV;
  }
  *intF = s;
}
/* lsm2dSurfaceIntegral (math_utils2d :625-708) */
void oracle_lsm2dsurfaceintegral(double *intF, const double *F, BOXP2(fg), const double *phi, BOXP2(pg),
                                 const double *px, const double *py, BOXP2(gg), BOXP2(ib), const double *dx,
                                 const double *dy, const double *eps) {
  orc_arr FF = V2(F, fg), P = V2(phi, pg), X = V2(px, gg), Y = V2(py, gg);
  double dV = *dx * *dy, ie = 1.0 / *eps, s = 0.0;
  obox b = B2(ib);
  FOR2(b) {
    double p = A2(P, i, j);
    if (fabs(p) < *eps) {
      double delta = 0.5 * ie * (1.0 + cos(ORC_PI * p * ie));
      double ng = sqrt(A2(X, i, j) * A2(X, i, j) + A2(Y, i, j) * A2(Y, i, j));
      s = s + delta * ng * A2(FF, i, j) * dV;
    }
  }
  *intF = s;
}
/* lsm2dAreaRegionPhiLessThanZero (geometry2d :304-368) */
void oracle_lsm2darearegionphilessthanzero(double *area, const double *phi, BOXP2(pg), BOXP2(ib), const double *dx,
                                           const double *dy, const double *eps) {
  orc_arr P = V2(phi, pg);
  double dA = *dx * *dy, e = *eps, s = 0.0;
  obox b = B2(ib);
  FOR2(b) {
    double p = A2(P, i, j), pe = p / e;
    if (p < -e) s = s + dA;
    else if (p < e) s = s + 0.5 * (1.0 - pe - ORC_ONE_OVER_PI * sin(ORC_PI * pe)) * dA;
  }
  *area = s;
}
/* lsm2dAreaRegionPhiGreaterThanZero (geometry2d :386-449) */
void oracle_lsm2darearegionphigreaterthanzero(double *area, const double *phi, BOXP2(pg), BOXP2(ib),
                                              const double *dx, const double *dy, const double *eps) {
  orc_arr P = V2(phi, pg);
  double dA = *dx * *dy, e = *eps, s = 0.0;
  obox b = B2(ib);
  FOR2(b) {
    double p = A2(P, i, j), pe = p / e;
    if (p > e) s = s + dA;
    else if (p > -e) s = s + 0.5 * (1.0 + pe + ORC_ONE_OVER_PI * sin(ORC_PI * pe)) * dA;
  }
  *area = s;
}
/* lsm2dPerimeterZeroLevelSet (geometry2d :469-545) */
void oracle_lsm2dperimeterzerolevelset(double *perim, const double *phi, BOXP2(pg), const double *px,
                                       const double *py, BOXP2(gg), BOXP2(ib), const double *dx, const double *dy,
                                       const double *eps) {
  orc_arr P = V2(phi, pg), X = V2(px, gg), Y = V2(py, gg);
  double dA = *dx * *dy, ie = 1.0 / *eps, s = 0.0;
  obox b = B2(ib);
  FOR2(b) {
    double p = A2(P, i, j);
    if (fabs(p) < *eps) {
      double delta = 0.5 * ie * (1.0 + cos(ORC_PI * p * ie));
      double ng = sqrt(A2(X, i, j) * A2(X, i, j) + A2(Y, i, j) * A2(Y, i, j));
      s = s + delta * ng * dA;
    }
  }
  *perim = s;
}

/* ---------------------------------------------------------------------------
 * Ghost-cell fills, src/boundary_conditions/lsm_boundary_conditions2d.f
 * (faces 0..3: x_lo, x_hi, y_lo, y_hi; each sweeps the other axis' full ghost range)
 * ------------------------------------------------------------------------- */
static void face2(double *phi, BOXP2(gb), BOXP2(fb), int idx, int kind) {
  if (idx >= 0 && idx <= 3) orc_bc_face(V2(phi, gb), B2(gb), B2(fb), idx, kind);
}
/* :44-126 */
void oracle_lsm2dlinearextrapolation(double *phi, BOXP2(gb), BOXP2(fb), const int *idx) {
  face2(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, *idx, BC_LINEAR);
}
/* :149-242 */
void oracle_lsm2dsignedlinearextrapolation(double *phi, BOXP2(gb), BOXP2(fb), const int *idx) {
  face2(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, *idx, BC_SIGNED);
}
/* :264-337; the four homogeneous Neumann fills (:360-510) are copy extrapolation */
void oracle_lsm2dcopyextrapolation(double *phi, BOXP2(gb), BOXP2(fb), const int *idx) {
  face2(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, *idx, BC_COPY);
}
void oracle_lsm2dhomogeneousneumanneno1(double *phi, BOXP2(gb), BOXP2(fb), const int *idx) {
  face2(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, *idx, BC_COPY);
}
void oracle_lsm2dhomogeneousneumanneno2(double *phi, BOXP2(gb), BOXP2(fb), const int *idx) {
  face2(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, *idx, BC_COPY);
}
void oracle_lsm2dhomogeneousneumanneno3(double *phi, BOXP2(gb), BOXP2(fb), const int *idx) {
  face2(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, *idx, BC_COPY);
}
void oracle_lsm2dhomogeneousneumannweno5(double *phi, BOXP2(gb), BOXP2(fb), const int *idx) {
  face2(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, ilo_fb, ihi_fb, jlo_fb, jhi_fb, *idx, BC_COPY);
}
