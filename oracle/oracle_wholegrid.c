/*
 * oracle_wholegrid.c -- TEST INFRASTRUCTURE ONLY (see lsm_oracle.h).
 *
 * Whole-grid kernels: undivided differences, HJ-ENO1/2/3, HJ-WENO5 (plain and
 * velocity-upwinded), central gradients, Laplacian, gradient magnitude,
 * level-set equation RHS terms, reinitialization RHS, mean curvature, TVD
 * Runge-Kutta stages, CFL time steps, max-norm difference, smoothed volume and
 * surface integrals, ghost-cell extrapolation and the smoothed Heaviside/delta.
 * Each function cites the reference routine whose operation order it keeps.
 */
#include "lsm_oracle.h"

static obox mkbox(const int *ilo, const int *ihi, const int *jlo, const int *jhi,
                  const int *klo, const int *khi) {
  obox b = {*ilo, *ihi, *jlo, *jhi, *klo, *khi};
  return b;
}
#define BOX(n) mkbox(ilo_##n, ihi_##n, jlo_##n, jhi_##n, klo_##n, khi_##n)

static obox widen(obox b, int dir, int w) {
  if (dir == 0) { b.ilo -= w; b.ihi += w; }
  if (dir == 1) { b.jlo -= w; b.jhi += w; }
  if (dir == 2) { b.klo -= w; b.khi += w; }
  return b;
}

/* lsm3dComputeDn, src/toolbox/lsm_spatial_derivatives3d.f.in:49-175.
 * Odd n: Dn(i) = Dm1(i) - Dm1(i-1) on fb plus one extra face in dir;
 * even n: Dn(i) = -1*(Dm1(i) - Dm1(i+1)) on fb.  Every other entry of Dn's
 * ghost box is set to 1e10. */
static void computeDn(orc_arr Dn, obox gb, orc_arr Dm1, obox fb, int n, int dir) {
  int off[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
  double sgn;
  if (n % 2 == 1) { off[dir] = 1; sgn = 1.0; sh[dir] = 1; }
  else { off[dir] = -1; sgn = -1.0; }
  for (int k = fb.klo; k <= fb.khi + sh[2]; k++)
    for (int j = fb.jlo; j <= fb.jhi + sh[1]; j++)
      for (int i = fb.ilo; i <= fb.ihi + sh[0]; i++)
        AT(Dn, i, j, k) = sgn * (AT(Dm1, i, j, k) - AT(Dm1, i - off[0], j - off[1], k - off[2]));
  for (int k = gb.klo; k <= gb.khi; k++)
    for (int j = gb.jlo; j <= gb.jhi; j++) {
      for (int i = gb.ilo; i <= fb.ilo - 1; i++) AT(Dn, i, j, k) = ORC_BIG;
      for (int i = fb.ihi + sh[0] + 1; i <= gb.ihi; i++) AT(Dn, i, j, k) = ORC_BIG;
    }
  for (int k = gb.klo; k <= gb.khi; k++) {
    for (int j = gb.jlo; j <= fb.jlo - 1; j++)
      for (int i = gb.ilo; i <= gb.ihi; i++) AT(Dn, i, j, k) = ORC_BIG;
    for (int j = fb.jhi + sh[1] + 1; j <= gb.jhi; j++)
      for (int i = gb.ilo; i <= gb.ihi; i++) AT(Dn, i, j, k) = ORC_BIG;
  }
  for (int k = gb.klo; k <= fb.klo - 1; k++)
    for (int j = gb.jlo; j <= gb.jhi; j++)
      for (int i = gb.ilo; i <= gb.ihi; i++) AT(Dn, i, j, k) = ORC_BIG;
  for (int k = fb.khi + sh[2] + 1; k <= gb.khi; k++)
    for (int j = gb.jlo; j <= gb.jhi; j++)
      for (int i = gb.ilo; i <= gb.ihi; i++) AT(Dn, i, j, k) = ORC_BIG;
}

/* Exported for direct comparison with the reference's internal routine. */
void oracle_lsm3dcomputedn(double *Dn, BOXP(Dn_gb), const double *Dm1, BOXP(Dm1_gb),
                           BOXP(fb), const int *n, const int *dir) {
  computeDn(VIEW(Dn, Dn_gb), BOX(Dn_gb), VIEW(Dm1, Dm1_gb), BOX(fb), *n, *dir - 1);
}

/* Offsets of the stencil neighbour s (in units of the axis step) along dir. */
#define NB(a, dir, i, j, k, s) \
  AT(a, (i) + ((dir) == 0 ? (s) : 0), (j) + ((dir) == 1 ? (s) : 0), (k) + ((dir) == 2 ? (s) : 0))

/* One HJ-WENO5 reconstruction from five scaled differences; operation order of
 * src/toolbox/lsm_spatial_derivatives3d.f.in:1296-1323. */
static double weno5(double v1, double v2, double v3, double v4, double v5) {
  const double one_third = 1.0 / 3.0, seven_sixths = 7.0 / 6.0, eleven_sixths = 11.0 / 6.0;
  const double one_sixth = 1.0 / 6.0, five_sixths = 5.0 / 6.0;
  const double thirteen_twelfths = 13.0 / 12.0, one_fourth = 0.25;
  double m = v1 * v1;
  if (v2 * v2 > m) m = v2 * v2;
  if (v3 * v3 > m) m = v3 * v3;
  if (v4 * v4 > m) m = v4 * v4;
  if (v5 * v5 > m) m = v5 * v5;
  double eps = 1e-6 * m + ORC_TINY;
  double p1 = one_third * v1 - seven_sixths * v2 + eleven_sixths * v3;
  double p2 = -(one_sixth * v2) + five_sixths * v3 + one_third * v4;
  double p3 = one_third * v3 + five_sixths * v4 - one_sixth * v5;
  double S1 = thirteen_twelfths * orc_sq(v1 - 2.0 * v2 + v3) + one_fourth * orc_sq(v1 - 4.0 * v2 + 3.0 * v3);
  double S2 = thirteen_twelfths * orc_sq(v2 - 2.0 * v3 + v4) + one_fourth * orc_sq(v2 - v4);
  double S3 = thirteen_twelfths * orc_sq(v3 - 2.0 * v4 + v5) + one_fourth * orc_sq(3.0 * v3 - 4.0 * v4 + v5);
  double a1 = 0.1 / orc_sq(S1 + eps);
  double a2 = 0.6 / orc_sq(S2 + eps);
  double a3 = 0.3 / orc_sq(S3 + eps);
  double inv = 1.0 / (a1 + a2 + a3);
  a1 = a1 * inv;
  a2 = a2 * inv;
  a3 = a3 * inv;
  return a1 * p1 + a2 * p2 + a3 * p3;
}

/* ENO2 plus/minus selections, src/toolbox/lsm_spatial_derivatives3d.f.in:515-540. */
static double eno2_plus(orc_arr D1, orc_arr D2, int d, int i, int j, int k) {
  if (fabs(NB(D2, d, i, j, k, 0)) < fabs(NB(D2, d, i, j, k, 1)))
    return NB(D1, d, i, j, k, 1) - 0.5 * NB(D2, d, i, j, k, 0);
  return NB(D1, d, i, j, k, 1) - 0.5 * NB(D2, d, i, j, k, 1);
}
static double eno2_minus(orc_arr D1, orc_arr D2, int d, int i, int j, int k) {
  if (fabs(NB(D2, d, i, j, k, -1)) < fabs(NB(D2, d, i, j, k, 0)))
    return NB(D1, d, i, j, k, 0) + 0.5 * NB(D2, d, i, j, k, -1);
  return NB(D1, d, i, j, k, 0) + 0.5 * NB(D2, d, i, j, k, 0);
}

/* ENO3 plus/minus (undivided), src/toolbox/lsm_spatial_derivatives3d.f.in:837-903.
 * The reference accumulates into the output array, i.e. rounds after each add. */
static double eno3_plus(orc_arr D1, orc_arr D2, orc_arr D3, int d, int i, int j, int k) {
  const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
  double r = NB(D1, d, i, j, k, 1);
  if (fabs(NB(D2, d, i, j, k, 0)) < fabs(NB(D2, d, i, j, k, 1))) {
    r = r - 0.5 * NB(D2, d, i, j, k, 0);
    if (fabs(NB(D3, d, i, j, k, 0)) < fabs(NB(D3, d, i, j, k, 1))) r = r - sixth * NB(D3, d, i, j, k, 0);
    else r = r - sixth * NB(D3, d, i, j, k, 1);
  } else {
    r = r - 0.5 * NB(D2, d, i, j, k, 1);
    if (fabs(NB(D3, d, i, j, k, 1)) < fabs(NB(D3, d, i, j, k, 2))) r = r + third * NB(D3, d, i, j, k, 1);
    else r = r + third * NB(D3, d, i, j, k, 2);
  }
  return r;
}
static double eno3_minus(orc_arr D1, orc_arr D2, orc_arr D3, int d, int i, int j, int k) {
  const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
  double r = NB(D1, d, i, j, k, 0);
  if (fabs(NB(D2, d, i, j, k, -1)) < fabs(NB(D2, d, i, j, k, 0))) {
    r = r + 0.5 * NB(D2, d, i, j, k, -1);
    if (fabs(NB(D3, d, i, j, k, -1)) < fabs(NB(D3, d, i, j, k, 0))) r = r + third * NB(D3, d, i, j, k, -1);
    else r = r + third * NB(D3, d, i, j, k, 0);
  } else {
    r = r + 0.5 * NB(D2, d, i, j, k, 0);
    if (fabs(NB(D3, d, i, j, k, 0)) < fabs(NB(D3, d, i, j, k, 1))) r = r - sixth * NB(D3, d, i, j, k, 0);
    else r = r - sixth * NB(D3, d, i, j, k, 1);
  }
  return r;
}

static double weno5_plus(orc_arr D1, int d, int i, int j, int k, double inv) {
  return weno5(NB(D1, d, i, j, k, 3) * inv, NB(D1, d, i, j, k, 2) * inv, NB(D1, d, i, j, k, 1) * inv,
               NB(D1, d, i, j, k, 0) * inv, NB(D1, d, i, j, k, -1) * inv);
}
static double weno5_minus(orc_arr D1, int d, int i, int j, int k, double inv) {
  return weno5(NB(D1, d, i, j, k, -2) * inv, NB(D1, d, i, j, k, -1) * inv, NB(D1, d, i, j, k, 0) * inv,
               NB(D1, d, i, j, k, 1) * inv, NB(D1, d, i, j, k, 2) * inv);
}

#define FOR_FB(fb)                                 \
  for (int k = (fb).klo; k <= (fb).khi; k++)       \
    for (int j = (fb).jlo; j <= (fb).jhi; j++)     \
      for (int i = (fb).ilo; i <= (fb).ihi; i++)

/* Shared driver of the plus/minus HJ-ENO/WENO routines
 * (src/toolbox/lsm_spatial_derivatives3d.f.in:199-1611): for each axis compute
 * the needed differences over the widened fill box, then the two one-sided
 * derivatives on the fill box. */
void orc_hj_plus_minus(int nd, int order, orc_arr *plus, orc_arr *minus, orc_arr phi, orc_arr D1, obox gb1,
                          orc_arr D2, obox gb2, orc_arr D3, obox gb3, obox fb, const double inv[3]) {
  for (int d = 0; d < nd; d++) {
    if (order == 1) {
      computeDn(D1, gb1, phi, fb, 1, d);
    } else if (order == 2) {
      computeDn(D1, gb1, phi, widen(fb, d, 1), 1, d);
      computeDn(D2, gb2, D1, widen(fb, d, 1), 2, d);
    } else if (order == 3) {
      computeDn(D1, gb1, phi, widen(fb, d, 2), 1, d);
      computeDn(D2, gb2, D1, widen(fb, d, 2), 2, d);
      computeDn(D3, gb3, D2, widen(fb, d, 1), 3, d);
    } else {
      computeDn(D1, gb1, phi, widen(fb, d, 2), 1, d);
    }
    FOR_FB(fb) {
      double p, m;
      if (order == 1) {
        p = NB(D1, d, i, j, k, 1) * inv[d];
        m = NB(D1, d, i, j, k, 0) * inv[d];
      } else if (order == 2) {
        p = eno2_plus(D1, D2, d, i, j, k) * inv[d];
        m = eno2_minus(D1, D2, d, i, j, k) * inv[d];
      } else if (order == 3) {
        p = eno3_plus(D1, D2, D3, d, i, j, k) * inv[d];
        m = eno3_minus(D1, D2, D3, d, i, j, k) * inv[d];
      } else {
        p = weno5_plus(D1, d, i, j, k, inv[d]);
        m = weno5_minus(D1, d, i, j, k, inv[d]);
      }
      AT(plus[d], i, j, k) = p;
      AT(minus[d], i, j, k) = m;
    }
  }
}

#define INV3(inv) double inv[3] = {1.0 / *dx, 1.0 / *dy, 1.0 / *dz}

/* lsm3dHJENO1, src/toolbox/lsm_spatial_derivatives3d.f.in:199-368 */
void oracle_lsm3dhjeno1(double *pxp, double *pyp, double *pzp, BOXP(gp), double *pxm, double *pym,
                        double *pzm, BOXP(gm), const double *phi, BOXP(phi_gb), double *D1, BOXP(d1),
                        BOXP(fb), const double *dx, const double *dy, const double *dz) {
  INV3(inv);
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  orc_arr none = {0};
  obox nb = {0};
  orc_hj_plus_minus(3, 1, P, M, VIEW(phi, phi_gb), VIEW(D1, d1), BOX(d1), none, nb, none, nb, BOX(fb), inv);
}

/* lsm3dHJENO2, src/toolbox/lsm_spatial_derivatives3d.f.in:393-663 */
void oracle_lsm3dhjeno2(double *pxp, double *pyp, double *pzp, BOXP(gp), double *pxm, double *pym,
                        double *pzm, BOXP(gm), const double *phi, BOXP(phi_gb), double *D1, BOXP(d1),
                        double *D2, BOXP(d2), BOXP(fb), const double *dx, const double *dy,
                        const double *dz) {
  INV3(inv);
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  orc_arr none = {0};
  obox nb = {0};
  orc_hj_plus_minus(3, 2, P, M, VIEW(phi, phi_gb), VIEW(D1, d1), BOX(d1), VIEW(D2, d2), BOX(d2), none, nb,
                BOX(fb), inv);
}

/* lsm3dHJENO3, src/toolbox/lsm_spatial_derivatives3d.f.in:689-1137 */
void oracle_lsm3dhjeno3(double *pxp, double *pyp, double *pzp, BOXP(gp), double *pxm, double *pym,
                        double *pzm, BOXP(gm), const double *phi, BOXP(phi_gb), double *D1, BOXP(d1),
                        double *D2, BOXP(d2), double *D3, BOXP(d3), BOXP(fb), const double *dx,
                        const double *dy, const double *dz) {
  INV3(inv);
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  orc_hj_plus_minus(3, 3, P, M, VIEW(phi, phi_gb), VIEW(D1, d1), BOX(d1), VIEW(D2, d2), BOX(d2), VIEW(D3, d3),
                BOX(d3), BOX(fb), inv);
}

/* lsm3dHJWENO5, src/toolbox/lsm_spatial_derivatives3d.f.in:1161-1611 */
void oracle_lsm3dhjweno5(double *pxp, double *pyp, double *pzp, BOXP(gp), double *pxm, double *pym,
                         double *pzm, BOXP(gm), const double *phi, BOXP(phi_gb), double *D1, BOXP(d1),
                         BOXP(fb), const double *dx, const double *dy, const double *dz) {
  INV3(inv);
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  orc_arr none = {0};
  obox nb = {0};
  orc_hj_plus_minus(3, 5, P, M, VIEW(phi, phi_gb), VIEW(D1, d1), BOX(d1), none, nb, none, nb, BOX(fb), inv);
}

/* Shared driver of the velocity-upwinded routines
 * (src/toolbox/lsm_spatial_derivatives3d.f.in:1630-2999): |v| < zero_tol gives
 * 0, v > 0 the minus (backward) derivative, otherwise the plus one. */
void orc_hj_upwind(int nd, int order, orc_arr *out, orc_arr phi, orc_arr *vel, orc_arr D1, obox gb1, orc_arr D2,
                      obox gb2, orc_arr D3, obox gb3, obox fb, const double inv[3]) {
  for (int d = 0; d < nd; d++) {
    if (order == 1) {
      computeDn(D1, gb1, phi, fb, 1, d);
    } else if (order == 2) {
      computeDn(D1, gb1, phi, widen(fb, d, 1), 1, d);
      computeDn(D2, gb2, D1, widen(fb, d, 1), 2, d);
    } else if (order == 3) {
      computeDn(D1, gb1, phi, widen(fb, d, 2), 1, d);
      computeDn(D2, gb2, D1, widen(fb, d, 2), 2, d);
      computeDn(D3, gb3, D2, widen(fb, d, 1), 3, d);
    } else {
      computeDn(D1, gb1, phi, widen(fb, d, 2), 1, d);
    }
    FOR_FB(fb) {
      double v = AT(vel[d], i, j, k), r;
      if (fabs(v) < ORC_ZERO_TOL) {
        r = 0.0;
      } else if (v > 0) {
        if (order == 1) r = NB(D1, d, i, j, k, 0) * inv[d];
        else if (order == 2) r = eno2_minus(D1, D2, d, i, j, k) * inv[d];
        else if (order == 3) r = eno3_minus(D1, D2, D3, d, i, j, k) * inv[d];
        else r = weno5_minus(D1, d, i, j, k, inv[d]);
      } else {
        if (order == 1) r = NB(D1, d, i, j, k, 1) * inv[d];
        else if (order == 2) r = eno2_plus(D1, D2, d, i, j, k) * inv[d];
        else if (order == 3) r = eno3_plus(D1, D2, D3, d, i, j, k) * inv[d];
        else r = weno5_plus(D1, d, i, j, k, inv[d]);
      }
      AT(out[d], i, j, k) = r;
    }
  }
}

/* lsm3dUpwindHJENO1, src/toolbox/lsm_spatial_derivatives3d.f.in:1630-1836 */
void oracle_lsm3dupwindhjeno1(double *px, double *py, double *pz, BOXP(g), const double *phi, BOXP(phi_gb),
                              const double *vx, const double *vy, const double *vz, BOXP(vel_gb), double *D1,
                              BOXP(d1), BOXP(fb), const double *dx, const double *dy, const double *dz) {
  INV3(inv);
  orc_arr O[3] = {VIEW(px, g), VIEW(py, g), VIEW(pz, g)};
  orc_arr V[3] = {VIEW(vx, vel_gb), VIEW(vy, vel_gb), VIEW(vz, vel_gb)};
  orc_arr none = {0};
  obox nb = {0};
  orc_hj_upwind(3, 1, O, VIEW(phi, phi_gb), V, VIEW(D1, d1), BOX(d1), none, nb, none, nb, BOX(fb), inv);
}

/* lsm3dUpwindHJENO2, src/toolbox/lsm_spatial_derivatives3d.f.in:1854-2156 */
void oracle_lsm3dupwindhjeno2(double *px, double *py, double *pz, BOXP(g), const double *phi, BOXP(phi_gb),
                              const double *vx, const double *vy, const double *vz, BOXP(vel_gb), double *D1,
                              BOXP(d1), double *D2, BOXP(d2), BOXP(fb), const double *dx, const double *dy,
                              const double *dz) {
  INV3(inv);
  orc_arr O[3] = {VIEW(px, g), VIEW(py, g), VIEW(pz, g)};
  orc_arr V[3] = {VIEW(vx, vel_gb), VIEW(vy, vel_gb), VIEW(vz, vel_gb)};
  orc_arr none = {0};
  obox nb = {0};
  orc_hj_upwind(3, 2, O, VIEW(phi, phi_gb), V, VIEW(D1, d1), BOX(d1), VIEW(D2, d2), BOX(d2), none, nb, BOX(fb), inv);
}

/* lsm3dUpwindHJENO3, src/toolbox/lsm_spatial_derivatives3d.f.in:2174-2609 */
void oracle_lsm3dupwindhjeno3(double *px, double *py, double *pz, BOXP(g), const double *phi, BOXP(phi_gb),
                              const double *vx, const double *vy, const double *vz, BOXP(vel_gb), double *D1,
                              BOXP(d1), double *D2, BOXP(d2), double *D3, BOXP(d3), BOXP(fb),
                              const double *dx, const double *dy, const double *dz) {
  INV3(inv);
  orc_arr O[3] = {VIEW(px, g), VIEW(py, g), VIEW(pz, g)};
  orc_arr V[3] = {VIEW(vx, vel_gb), VIEW(vy, vel_gb), VIEW(vz, vel_gb)};
  orc_hj_upwind(3, 3, O, VIEW(phi, phi_gb), V, VIEW(D1, d1), BOX(d1), VIEW(D2, d2), BOX(d2), VIEW(D3, d3), BOX(d3),
            BOX(fb), inv);
}

/* lsm3dUpwindHJWENO5, src/toolbox/lsm_spatial_derivatives3d.f.in:2628-2999 */
void oracle_lsm3dupwindhjweno5(double *px, double *py, double *pz, BOXP(g), const double *phi,
                               BOXP(phi_gb), const double *vx, const double *vy, const double *vz,
                               BOXP(vel_gb), double *D1, BOXP(d1), BOXP(fb), const double *dx,
                               const double *dy, const double *dz) {
  INV3(inv);
  orc_arr O[3] = {VIEW(px, g), VIEW(py, g), VIEW(pz, g)};
  orc_arr V[3] = {VIEW(vx, vel_gb), VIEW(vy, vel_gb), VIEW(vz, vel_gb)};
  orc_arr none = {0};
  obox nb = {0};
  orc_hj_upwind(3, 5, O, VIEW(phi, phi_gb), V, VIEW(D1, d1), BOX(d1), none, nb, none, nb, BOX(fb), inv);
}

/* lsm3dCentralGradOrder2, src/toolbox/lsm_spatial_derivatives3d.f.in:3016-3077 */
void oracle_lsm3dcentralgradorder2(double *px, double *py, double *pz, BOXP(g), const double *phi,
                                   BOXP(phi_gb), BOXP(fb), const double *dx, const double *dy,
                                   const double *dz) {
  orc_arr X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g), F = VIEW(phi, phi_gb);
  double fx = 0.5 / *dx, fy = 0.5 / *dy, fz = 0.5 / *dz;
  obox b = BOX(fb);
  FOR_FB(b) {
    AT(X, i, j, k) = (AT(F, i + 1, j, k) - AT(F, i - 1, j, k)) * fx;
    AT(Y, i, j, k) = (AT(F, i, j + 1, k) - AT(F, i, j - 1, k)) * fy;
    AT(Z, i, j, k) = (AT(F, i, j, k + 1) - AT(F, i, j, k - 1)) * fz;
  }
}

/* lsm3dCentralGradOrder4, src/toolbox/lsm_spatial_derivatives3d.f.in:3094-3162 */
void oracle_lsm3dcentralgradorder4(double *px, double *py, double *pz, BOXP(g), const double *phi,
                                   BOXP(phi_gb), BOXP(fb), const double *dx, const double *dy,
                                   const double *dz) {
  orc_arr X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g), F = VIEW(phi, phi_gb);
  double fx = 0.0833333333333333333333 / *dx, fy = 0.0833333333333333333333 / *dy;
  double fz = 0.0833333333333333333333 / *dz;
  obox b = BOX(fb);
  FOR_FB(b) {
    AT(X, i, j, k) = (-AT(F, i + 2, j, k) + 8.0 * AT(F, i + 1, j, k) + AT(F, i - 2, j, k) -
                      8.0 * AT(F, i - 1, j, k)) * fx;
    AT(Y, i, j, k) = (-AT(F, i, j + 2, k) + 8.0 * AT(F, i, j + 1, k) + AT(F, i, j - 2, k) -
                      8.0 * AT(F, i, j - 1, k)) * fy;
    AT(Z, i, j, k) = (-AT(F, i, j, k + 2) + 8.0 * AT(F, i, j, k + 1) + AT(F, i, j, k - 2) -
                      8.0 * AT(F, i, j, k - 1)) * fz;
  }
}

/* lsm3dLaplacianOrder2, src/toolbox/lsm_spatial_derivatives3d.f.in:3180-3252 */
void oracle_lsm3dlaplacianorder2(double *lap, BOXP(g), const double *phi, BOXP(phi_gb), BOXP(fb),
                                 const double *dx, const double *dy, const double *dz) {
  orc_arr L = VIEW(lap, g), F = VIEW(phi, phi_gb);
  double ix = 1.0 / *dx / *dx, iy = 1.0 / *dy / *dy, iz = 1.0 / *dz / *dz;
  obox b = BOX(fb);
  FOR_FB(b) {
    double c = AT(F, i, j, k);
    AT(L, i, j, k) = ix * (AT(F, i + 1, j, k) - 2.0 * c + AT(F, i - 1, j, k)) +
                     iy * (AT(F, i, j + 1, k) - 2.0 * c + AT(F, i, j - 1, k)) +
                     iz * (AT(F, i, j, k + 1) - 2.0 * c + AT(F, i, j, k - 1));
  }
}

/* lsm3dAverageGradPhi, src/toolbox/lsm_spatial_derivatives3d.f.in:3486-3568 */
void oracle_lsm3daveragegradphi(double *px, double *py, double *pz, BOXP(g), const double *pxp,
                                const double *pyp, const double *pzp, BOXP(gp), const double *pxm,
                                const double *pym, const double *pzm, BOXP(gm), BOXP(fb)) {
  orc_arr O[3] = {VIEW(px, g), VIEW(py, g), VIEW(pz, g)};
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  obox b = BOX(fb);
  FOR_FB(b) for (int d = 0; d < 3; d++) AT(O[d], i, j, k) = 0.5 * (AT(P[d], i, j, k) + AT(M[d], i, j, k));
}

/* lsm3dGradientMagnitude, src/toolbox/lsm_spatial_derivatives3d.f.in:3586-3659
 * (tolerance 1e-8 on each squared component and on the result). */
void oracle_lsm3dgradientmagnitude(const double *px, const double *py, const double *pz, double *mag,
                                   BOXP(g), BOXP(fb)) {
  orc_arr X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g), G = VIEW(mag, g);
  const double tol = 1.e-8;
  obox b = BOX(fb);
  FOR_FB(b) {
    double sx = AT(X, i, j, k) * AT(X, i, j, k);
    double sy = AT(Y, i, j, k) * AT(Y, i, j, k);
    double sz = AT(Z, i, j, k) * AT(Z, i, j, k);
    if (sx < tol) sx = 0.0;
    if (sy < tol) sy = 0.0;
    if (sz < tol) sz = 0.0;
    double t = sqrt(sx + sy + sz);
    AT(G, i, j, k) = t < tol ? 0.0 : t;
  }
}

/* ---------------------------------------------------------------------------
 * Level-set equation RHS terms, src/toolbox/lsm_level_set_evolution3d.f.in
 * ------------------------------------------------------------------------- */

/* :16-46 -- zeroes the whole ghost box */
void oracle_lsm3dzerooutlevelseteqnrhs(double *rhs, BOXP(g)) {
  orc_arr R = VIEW(rhs, g);
  obox b = BOX(g);
  FOR_FB(b) AT(R, i, j, k) = 0.0;
}

/* :53-124 */
void oracle_lsm3daddadvectiontermtolserhs(double *rhs, BOXP(rg), const double *px, const double *py,
                                          const double *pz, BOXP(g), const double *vx, const double *vy,
                                          const double *vz, BOXP(vg), BOXP(fb)) {
  orc_arr R = VIEW(rhs, rg), X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g);
  orc_arr U = VIEW(vx, vg), V = VIEW(vy, vg), W = VIEW(vz, vg);
  obox b = BOX(fb);
  FOR_FB(b) {
    AT(R, i, j, k) = AT(R, i, j, k) - (AT(U, i, j, k) * AT(X, i, j, k) + AT(V, i, j, k) * AT(Y, i, j, k) +
                                       AT(W, i, j, k) * AT(Z, i, j, k));
  }
}

/* Godunov |grad phi|^2 for normal motion, :210-224 */
static double godunov_sq(double vn, double xp, double yp, double zp, double xm, double ym, double zm) {
  if (vn > 0.0)
    return orc_max(orc_sq(orc_max(xm, 0.0)), orc_sq(orc_min(xp, 0.0))) +
           orc_max(orc_sq(orc_max(ym, 0.0)), orc_sq(orc_min(yp, 0.0))) +
           orc_max(orc_sq(orc_max(zm, 0.0)), orc_sq(orc_min(zp, 0.0)));
  return orc_max(orc_sq(orc_min(xm, 0.0)), orc_sq(orc_max(xp, 0.0))) +
         orc_max(orc_sq(orc_min(ym, 0.0)), orc_sq(orc_max(yp, 0.0))) +
         orc_max(orc_sq(orc_min(zm, 0.0)), orc_sq(orc_max(zp, 0.0)));
}

/* :131-241 */
void oracle_lsm3daddnormalveltermtolserhs(double *rhs, BOXP(rg), const double *pxp, const double *pyp,
                                          const double *pzp, BOXP(gp), const double *pxm, const double *pym,
                                          const double *pzm, BOXP(gm), const double *vn, BOXP(vg), BOXP(fb)) {
  orc_arr R = VIEW(rhs, rg), VN = VIEW(vn, vg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  obox b = BOX(fb);
  FOR_FB(b) {
    double v = AT(VN, i, j, k);
    if (fabs(v) >= ORC_ZERO_TOL) {
      double s = godunov_sq(v, AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k), AT(XM, i, j, k),
                            AT(YM, i, j, k), AT(ZM, i, j, k));
      AT(R, i, j, k) = AT(R, i, j, k) - v * sqrt(s);
    }
  }
}

/* :248-349 */
void oracle_lsm3daddconstnormalveltermtolserhs(double *rhs, BOXP(rg), const double *pxp, const double *pyp,
                                               const double *pzp, BOXP(gp), const double *pxm,
                                               const double *pym, const double *pzm, BOXP(gm),
                                               const double *vel_n, BOXP(fb)) {
  orc_arr R = VIEW(rhs, rg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  double v = *vel_n;
  if (!(fabs(v) >= ORC_ZERO_TOL)) return;
  obox b = BOX(fb);
  FOR_FB(b) {
    double s = godunov_sq(v, AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k), AT(XM, i, j, k),
                          AT(YM, i, j, k), AT(ZM, i, j, k));
    AT(R, i, j, k) = AT(R, i, j, k) - v * sqrt(s);
  }
}

/* Curvature numerator of :433-451 (also lsm_curvature3d.f.in:103-120). */
static double curv_num(double xx, double xy, double xz, double yy, double yz, double zz, double x, double y,
                       double z) {
  return xx * y * y + yy * x * x - 2.0 * xy * x * y + xx * z * z + zz * x * x - 2.0 * xz * x * z +
         yy * z * z + zz * y * y - 2.0 * yz * y * z;
}

/* :356-459 */
void oracle_lsm3daddconstcurvtermtolserhs(double *rhs, BOXP(rg), const double *px, const double *py,
                                          const double *pz, BOXP(g), const double *pxx, const double *pxy,
                                          const double *pxz, const double *pyy, const double *pyz,
                                          const double *pzz, BOXP(g2), const double *b_, BOXP(fb)) {
  orc_arr R = VIEW(rhs, rg), X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g);
  orc_arr XX = VIEW(pxx, g2), XY = VIEW(pxy, g2), XZ = VIEW(pxz, g2), YY = VIEW(pyy, g2), YZ = VIEW(pyz, g2),
          ZZ = VIEW(pzz, g2);
  double bb = *b_;
  obox b = BOX(fb);
  FOR_FB(b) {
    double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
    double g2v = x * x + y * y + z * z, curv;
    if (g2v < ORC_ZERO_TOL) {
      curv = 0.0;
    } else {
      curv = curv_num(AT(XX, i, j, k), AT(XY, i, j, k), AT(XZ, i, j, k), AT(YY, i, j, k), AT(YZ, i, j, k),
                      AT(ZZ, i, j, k), x, y, z);
      curv = curv / g2v;
    }
    AT(R, i, j, k) = AT(R, i, j, k) + bb * curv;
  }
}

/* :465-525 -- the loops run k = klo_fb..khi_gb, j = jlo_fb..jhi_gb, i = ilo_gb..ihi_gb (rhs ghost box). */
void oracle_lsm3daddconstprecomputedcurvtermtolserhs(double *rhs, BOXP(rg), const double *kappa, BOXP(kg),
                                                     const double *gmag, BOXP(gg), const double *b_,
                                                     BOXP(fb)) {
  orc_arr R = VIEW(rhs, rg), K = VIEW(kappa, kg), G = VIEW(gmag, gg);
  double bb = *b_;
  for (int k = *klo_fb; k <= *khi_rg; k++)
    for (int j = *jlo_fb; j <= *jhi_rg; j++)
      for (int i = *ilo_rg; i <= *ihi_rg; i++)
        AT(R, i, j, k) = AT(R, i, j, k) + bb * AT(K, i, j, k) * AT(G, i, j, k);
}

/* Per-axis Godunov selection of :629-693 for H1 = u + V*phi_pm. */
static double ext_normal_pick(double u, double vn, double pp, double pm, int paren) {
  double hp = u + vn * pp, hm = u + vn * pm;
  if (hp > 0.0) {
    if (hm > 0.0) return pm;
    return paren ? -(u / vn) : -(u / vn);
  }
  if (hm < 0.0) return pp;
  if (fabs(hp) > fabs(hm)) return pp;
  return pm;
}

/* :532-718 -- same quirky loop bounds as the precomputed-curvature term. */
void oracle_lsm3daddexternalandnormalveltermtolserhs(double *rhs, BOXP(rg), const double *pxp,
                                                     const double *pyp, const double *pzp, BOXP(gp),
                                                     const double *pxm, const double *pym, const double *pzm,
                                                     BOXP(gm), const double *vn, const double *vx,
                                                     const double *vy, const double *vz, BOXP(vg), BOXP(fb)) {
  orc_arr R = VIEW(rhs, rg), VN = VIEW(vn, vg), U = VIEW(vx, vg), V = VIEW(vy, vg), W = VIEW(vz, vg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  for (int k = *klo_fb; k <= *khi_rg; k++)
    for (int j = *jlo_fb; j <= *jhi_rg; j++)
      for (int i = *ilo_rg; i <= *ihi_rg; i++) {
        double n = AT(VN, i, j, k), u = AT(U, i, j, k), v = AT(V, i, j, k), w = AT(W, i, j, k);
        double px = ext_normal_pick(u, n, AT(XP, i, j, k), AT(XM, i, j, k), 0);
        double py = ext_normal_pick(v, n, AT(YP, i, j, k), AT(YM, i, j, k), 1);
        double pz = ext_normal_pick(w, n, AT(ZP, i, j, k), AT(ZM, i, j, k), 1);
        double s = px * px + py * py + pz * pz;
        if (fabs(n) >= ORC_ZERO_TOL && s >= ORC_ZERO_TOL)
          AT(R, i, j, k) = AT(R, i, j, k) - u * px - v * py - w * pz - n * sqrt(s);
        else
          AT(R, i, j, k) = AT(R, i, j, k) - u * px - v * py - w * pz;
      }
}

/* ---------------------------------------------------------------------------
 * Reinitialization RHS, src/reinitialization/lsm_reinitialization3d.f.in:48-267
 * ------------------------------------------------------------------------- */
static double reinit_cell(double phi_cur, double sgn_src, int use_phi0, const double gp[3], const double gm[3],
                          double dx_sq) {
  double star[3];
  double s = use_phi0 ? sgn_src : phi_cur;
  for (int d = 0; d < 3; d++) {
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
  if (fabs(s) >= ORC_ZERO_TOL) {
    double n2 = star[0] * star[0] + star[1] * star[1] + star[2] * star[2];
    double sgn = use_phi0 ? s / sqrt(s * s + dx_sq) : s / sqrt(s * s + n2 * dx_sq);
    return sgn * (1.0 - sqrt(n2));
  }
  return 0.0;
}

void oracle_lsm3dcomputereinitializationeqnrhs(double *rr, BOXP(rg), const double *phi, BOXP(pg),
                                               const double *phi0, BOXP(p0g), const double *pxp,
                                               const double *pyp, const double *pzp, BOXP(gp),
                                               const double *pxm, const double *pym, const double *pzm,
                                               BOXP(gm), BOXP(fb), const double *dx, const double *dy,
                                               const double *dz, const int *use_phi0_for_sgn) {
  orc_arr R = VIEW(rr, rg), F = VIEW(phi, pg), F0 = VIEW(phi0, p0g);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  double h = orc_max(orc_max(*dx, *dy), *dz);
  double dx_sq = h * h;
  int use0 = (*use_phi0_for_sgn == 1);
  obox b = BOX(fb);
  FOR_FB(b) {
    double gpv[3] = {AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k)};
    double gmv[3] = {AT(XM, i, j, k), AT(YM, i, j, k), AT(ZM, i, j, k)};
    AT(R, i, j, k) = reinit_cell(AT(F, i, j, k), use0 ? AT(F0, i, j, k) : 0.0, use0, gpv, gmv, dx_sq);
  }
}

/* ---------------------------------------------------------------------------
 * Mean curvature, src/geometry/lsm_curvature3d.f.in:15-128
 * ------------------------------------------------------------------------- */
void oracle_lsm3dcomputemeancurvatureorder2(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                            const double *px, const double *py, const double *pz,
                                            const double *gmag, BOXP(gg), BOXP(fb), const double *dx,
                                            const double *dy, const double *dz) {
  orc_arr K = VIEW(kappa, kg), F = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg),
          G = VIEW(gmag, gg);
  double fxx = 1.0 / *dx / *dx, fyy = 1.0 / *dy / *dy, fzz = 1.0 / *dz / *dz;
  double fxy = 0.25 / *dx / *dy, fxz = 0.25 / *dx / *dz, fyz = 0.25 / *dy / *dz;
  obox b = BOX(fb);
  FOR_FB(b) {
    double c = AT(F, i, j, k);
    double xx = (AT(F, i + 1, j, k) - 2.0 * c + AT(F, i - 1, j, k)) * fxx;
    double yy = (AT(F, i, j + 1, k) - 2.0 * c + AT(F, i, j - 1, k)) * fyy;
    double zz = (AT(F, i, j, k + 1) - 2.0 * c + AT(F, i, j, k - 1)) * fzz;
    double xy = (AT(F, i + 1, j + 1, k) - AT(F, i + 1, j - 1, k) - AT(F, i - 1, j + 1, k) +
                 AT(F, i - 1, j - 1, k)) * fxy;
    double xz = (AT(F, i + 1, j, k + 1) - AT(F, i + 1, j, k - 1) - AT(F, i - 1, j, k + 1) +
                 AT(F, i - 1, j, k - 1)) * fxz;
    double yz = (AT(F, i, j + 1, k + 1) - AT(F, i, j + 1, k - 1) - AT(F, i, j - 1, k + 1) +
                 AT(F, i, j - 1, k - 1)) * fyz;
    double g = AT(G, i, j, k);
    double den = g * g * g;
    if (den < ORC_ZERO_TOL) {
      AT(K, i, j, k) = 0.0;
    } else {
      double num = curv_num(xx, xy, xz, yy, yz, zz, AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k));
      AT(K, i, j, k) = num / den;
    }
  }
}

/* ---------------------------------------------------------------------------
 * TVD Runge-Kutta stages, src/toolbox/lsm_tvd_runge_kutta3d.f
 * ------------------------------------------------------------------------- */
/* :32-86 (RK1 / TVD-RK2 stage 1 / TVD-RK3 stage 1) */
void oracle_lsm3drk1step(double *un, BOXP(ng), const double *uc, BOXP(cg), const double *rhs, BOXP(rg),
                         BOXP(fb), const double *dt) {
  orc_arr N = VIEW(un, ng), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  double t = *dt;
  obox b = BOX(fb);
  FOR_FB(b) AT(N, i, j, k) = AT(C, i, j, k) + t * AT(R, i, j, k);
}
void oracle_lsm3dtvdrk2stage1(double *un, BOXP(ng), const double *uc, BOXP(cg), const double *rhs,
                              BOXP(rg), BOXP(fb), const double *dt) {
  oracle_lsm3drk1step(un, ilo_ng, ihi_ng, jlo_ng, jhi_ng, klo_ng, khi_ng, uc, ilo_cg, ihi_cg, jlo_cg, jhi_cg,
                      klo_cg, khi_cg, rhs, ilo_rg, ihi_rg, jlo_rg, jhi_rg, klo_rg, khi_rg, ilo_fb, ihi_fb,
                      jlo_fb, jhi_fb, klo_fb, khi_fb, dt);
}
void oracle_lsm3dtvdrk3stage1(double *un, BOXP(ng), const double *uc, BOXP(cg), const double *rhs,
                              BOXP(rg), BOXP(fb), const double *dt) {
  oracle_lsm3drk1step(un, ilo_ng, ihi_ng, jlo_ng, jhi_ng, klo_ng, khi_ng, uc, ilo_cg, ihi_cg, jlo_cg, jhi_cg,
                      klo_cg, khi_cg, rhs, ilo_rg, ihi_rg, jlo_rg, jhi_rg, klo_rg, khi_rg, ilo_fb, ihi_fb,
                      jlo_fb, jhi_fb, klo_fb, khi_fb, dt);
}
/* :185-248 */
void oracle_lsm3dtvdrk2stage2(double *un, BOXP(ng), const double *u1, BOXP(g1), const double *uc, BOXP(cg),
                              const double *rhs, BOXP(rg), BOXP(fb), const double *dt) {
  orc_arr N = VIEW(un, ng), S1 = VIEW(u1, g1), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  double t = *dt;
  obox b = BOX(fb);
  FOR_FB(b) AT(N, i, j, k) = 0.5 * (AT(C, i, j, k) + AT(S1, i, j, k) + t * AT(R, i, j, k));
}
/* :347-412 */
void oracle_lsm3dtvdrk3stage2(double *u2, BOXP(g2), const double *u1, BOXP(g1), const double *uc, BOXP(cg),
                              const double *rhs, BOXP(rg), BOXP(fb), const double *dt) {
  orc_arr S2 = VIEW(u2, g2), S1 = VIEW(u1, g1), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  double t = *dt;
  obox b = BOX(fb);
  FOR_FB(b) AT(S2, i, j, k) = 0.75 * AT(C, i, j, k) + 0.25 * (AT(S1, i, j, k) + t * AT(R, i, j, k));
}
/* :431-500 */
void oracle_lsm3dtvdrk3stage3(double *un, BOXP(ng), const double *u2, BOXP(g2), const double *uc, BOXP(cg),
                              const double *rhs, BOXP(rg), BOXP(fb), const double *dt) {
  orc_arr N = VIEW(un, ng), S2 = VIEW(u2, g2), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  double t = *dt;
  const double one_third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
  obox b = BOX(fb);
  FOR_FB(b) AT(N, i, j, k) = one_third * AT(C, i, j, k) + two_thirds * (AT(S2, i, j, k) + t * AT(R, i, j, k));
}

/* ---------------------------------------------------------------------------
 * Reductions, src/toolbox/lsm_math_utils3d.f.in and src/geometry/lsm_geometry3d_fort.f.in
 * ------------------------------------------------------------------------- */
/* lsm_math_utils3d.f.in:16-75 -- starts from the first include-box cell */
void oracle_lsm3dmaxnormdiff(double *out, const double *f1, BOXP(g1), const double *f2, BOXP(g2), BOXP(ib)) {
  orc_arr A = VIEW(f1, g1), B = VIEW(f2, g2);
  double m = fabs(AT(A, *ilo_ib, *jlo_ib, *klo_ib) - AT(B, *ilo_ib, *jlo_ib, *klo_ib));
  obox b = BOX(ib);
  FOR_FB(b) {
    double d = fabs(AT(A, i, j, k) - AT(B, i, j, k));
    if (d > m) m = d;
  }
  *out = m;
}

/* lsm_math_utils3d.f.in:253-326 */
void oracle_lsm3dcomputestableadvectiondt(double *dt, const double *vx, const double *vy, const double *vz,
                                          BOXP(vg), BOXP(ib), const double *dx, const double *dy,
                                          const double *dz, const double *cfl) {
  orc_arr U = VIEW(vx, vg), V = VIEW(vy, vg), W = VIEW(vz, vg);
  double m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  obox b = BOX(ib);
  FOR_FB(b) {
    double c = fabs(AT(U, i, j, k)) * ix + fabs(AT(V, i, j, k)) * iy + fabs(AT(W, i, j, k)) * iz;
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

/* H/dX of the normal-motion CFL bound, lsm_math_utils3d.f.in:413-429 */
static double normal_h_over_dx(double absv, double xp, double yp, double zp, double xm, double ym, double zm,
                               double max_dx_sq, double ix, double iy, double iz) {
  double x = orc_max(fabs(xp), fabs(xm)), y = orc_max(fabs(yp), fabs(ym)), z = orc_max(fabs(zp), fabs(zm));
  double n = sqrt(x * x + y * y + z * z + max_dx_sq);
  return absv / n * (x * ix + y * iy + z * iz);
}

/* lsm_math_utils3d.f.in:332-449 */
void oracle_lsm3dcomputestablenormalveldt(double *dt, const double *vn, BOXP(vg), const double *pxp,
                                          const double *pyp, const double *pzp, BOXP(gp), const double *pxm,
                                          const double *pym, const double *pzm, BOXP(gm), BOXP(ib),
                                          const double *dx, const double *dy, const double *dz,
                                          const double *cfl) {
  orc_arr VN = VIEW(vn, vg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  double h = orc_max(orc_max(*dx, *dy), *dz), mdx2 = h * h;
  double m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  obox b = BOX(ib);
  FOR_FB(b) {
    double c = normal_h_over_dx(fabs(AT(VN, i, j, k)), AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k),
                                AT(XM, i, j, k), AT(YM, i, j, k), AT(ZM, i, j, k), mdx2, ix, iy, iz);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

/* lsm_math_utils3d.f.in:455-567 */
void oracle_lsm3dcomputestableconstnormalveldt(double *dt, const double *vel_n, const double *pxp,
                                               const double *pyp, const double *pzp, BOXP(gp),
                                               const double *pxm, const double *pym, const double *pzm,
                                               BOXP(gm), BOXP(ib), const double *dx, const double *dy,
                                               const double *dz, const double *cfl) {
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  double h = orc_max(orc_max(*dx, *dy), *dz), mdx2 = h * h;
  double m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz, av = fabs(*vel_n);
  obox b = BOX(ib);
  FOR_FB(b) {
    double c = normal_h_over_dx(av, AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k), AT(XM, i, j, k),
                                AT(YM, i, j, k), AT(ZM, i, j, k), mdx2, ix, iy, iz);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

#define ORC_PI 3.14159265358979323846
#define ORC_ONE_OVER_PI 0.31830988618379

/* lsm_math_utils3d.f.in:573-653 (naive serial sum, kept as the reference sums) */
void oracle_lsm3dvolumeintegralphilessthanzero(double *intF, const double *F, BOXP(fg), const double *phi,
                                               BOXP(pg), BOXP(ib), const double *dx, const double *dy,
                                               const double *dz, const double *epsilon) {
  orc_arr FF = VIEW(F, fg), P = VIEW(phi, pg);
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double p = AT(P, i, j, k), q = p / e;
    if (p < -e) s = s + AT(FF, i, j, k) * dV;
    else if (p < e) s = s + 0.5 * (1.0 - q - ORC_ONE_OVER_PI * sin(ORC_PI * q)) * AT(FF, i, j, k) * dV;
  }
  *intF = s;
}

/* lsm_math_utils3d.f.in:659-739 */
void oracle_lsm3dvolumeintegralphigreaterthanzero(double *intF, const double *F, BOXP(fg), const double *phi,
                                                  BOXP(pg), BOXP(ib), const double *dx, const double *dy,
                                                  const double *dz, const double *epsilon) {
  orc_arr FF = VIEW(F, fg), P = VIEW(phi, pg);
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double p = AT(P, i, j, k), q = p / e;
    if (p > e) s = s + AT(FF, i, j, k) * dV;
    else if (p > -e) s = s + 0.5 * (1.0 + q + ORC_ONE_OVER_PI * sin(ORC_PI * q)) * AT(FF, i, j, k) * dV;
  }
  *intF = s;
}

/* lsm_math_utils3d.f.in:745-845 */
void oracle_lsm3dsurfaceintegral(double *intF, const double *F, BOXP(fg), const double *phi, BOXP(pg),
                                 const double *px, const double *py, const double *pz, BOXP(gg), BOXP(ib),
                                 const double *dx, const double *dy, const double *dz, const double *epsilon) {
  orc_arr FF = VIEW(F, fg), P = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg);
  double dV = *dx * *dy * *dz, e = *epsilon, ie = 1.0 / e, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double p = AT(P, i, j, k);
    if (fabs(p) < e) {
      double delta = 0.5 * ie * (1.0 + cos(ORC_PI * p * ie));
      double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
      double n = sqrt(x * x + y * y + z * z);
      s = s + delta * n * AT(FF, i, j, k) * dV;
    }
  }
  *intF = s;
}

/* Per-cell addends of the smoothed volume and area sums below, in the
 * reference's operations: each sum adds exactly these values in loop order,
 * so the *_terms exports (one addend per cell of ib, 0 where the reference
 * adds nothing) let a test form the exactly rounded sum of the same terms. */
static inline double vol_lt0_term(double p, double e, double dV) {
  double q = p / e;
  if (p < -e) return dV;
  if (p < e) return 0.5 * (1.0 - q - ORC_ONE_OVER_PI * sin(ORC_PI * q)) * dV;
  return 0.0;
}
static inline double vol_gt0_term(double p, double e, double dV) {
  double q = p / e;
  if (p > e) return dV;
  if (p > -e) return 0.5 * (1.0 + q + ORC_ONE_OVER_PI * sin(ORC_PI * q)) * dV;
  return 0.0;
}
static inline double area_term(double p, double x, double y, double z, double e, double ie, double dV) {
  if (!(fabs(p) < e)) return 0.0;
  double delta = 0.5 * ie * (1.0 + cos(ORC_PI * p * ie));
  return delta * sqrt(x * x + y * y + z * z) * dV;
}

/* lsm_geometry3d_fort.f.in:395-465 */
void oracle_lsm3dvolumeregionphilessthanzero(double *vol, const double *phi, BOXP(pg), BOXP(ib),
                                             const double *dx, const double *dy, const double *dz,
                                             const double *epsilon) {
  orc_arr P = VIEW(phi, pg);
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double p = AT(P, i, j, k);
    if (p < e) s = s + vol_lt0_term(p, e, dV);
  }
  *vol = s;
}

/* lsm_geometry3d_fort.f.in:484-556 */
void oracle_lsm3dvolumeregionphigreaterthanzero(double *vol, const double *phi, BOXP(pg), BOXP(ib),
                                                const double *dx, const double *dy, const double *dz,
                                                const double *epsilon) {
  orc_arr P = VIEW(phi, pg);
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double p = AT(P, i, j, k);
    if (p > -e) s = s + vol_gt0_term(p, e, dV);
  }
  *vol = s;
}

/* lsm_geometry3d_fort.f.in:574-664 */
void oracle_lsm3dsurfaceareazerolevelset(double *area, const double *phi, BOXP(pg), const double *px,
                                         const double *py, const double *pz, BOXP(gg), BOXP(ib),
                                         const double *dx, const double *dy, const double *dz,
                                         const double *epsilon) {
  orc_arr P = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg);
  double dV = *dx * *dy * *dz, e = *epsilon, ie = 1.0 / e, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double p = AT(P, i, j, k);
    if (fabs(p) < e) s = s + area_term(p, AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k), e, ie, dV);
  }
  *area = s;
}

/* The addends of the three sums above, one per cell of ib in (k, j, i) loop
 * order (which: 0 volume of phi < 0, 1 volume of phi > 0, 2 area; px, py, pz
 * are read for the area only). */
void oracle_integral_terms(double *terms, int which, const double *phi, BOXP(pg), const double *px,
                           const double *py, const double *pz, BOXP(gg), BOXP(ib), const double *dx,
                           const double *dy, const double *dz, const double *epsilon) {
  orc_arr P = VIEW(phi, pg);
  double dV = *dx * *dy * *dz, e = *epsilon, ie = 1.0 / e;
  obox b = BOX(ib);
  long t = 0;
  FOR_FB(b) {
    double p = AT(P, i, j, k);
    if (which == 0) terms[t] = vol_lt0_term(p, e, dV);
    else if (which == 1) terms[t] = vol_gt0_term(p, e, dV);
    else {
      orc_arr X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg);
      terms[t] = area_term(p, AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k), e, ie, dV);
    }
    t++;
  }
}

/* ---------------------------------------------------------------------------
 * Ghost-cell fills, src/boundary_conditions/lsm_boundary_conditions3d.f:46-638
 * Each face sweeps the full ghost box of the other two axes.
 * ------------------------------------------------------------------------- */

void orc_bc_face(orc_arr P, obox gb, obox fb, int idx, int kind) {
  int d = idx / 2, hi = idx % 2;
  int lo_g[3] = {gb.ilo, gb.jlo, gb.klo}, hi_g[3] = {gb.ihi, gb.jhi, gb.khi};
  int lo_f[3] = {fb.ilo, fb.jlo, fb.klo}, hi_f[3] = {fb.ihi, fb.jhi, fb.khi};
  int a = (d + 1) % 3, c = (d + 2) % 3; /* the two transverse axes */
  int edge = hi ? hi_f[d] : lo_f[d];
  int inner = hi ? hi_f[d] - 1 : lo_f[d] + 1;
  int g0 = hi ? hi_f[d] + 1 : lo_g[d], g1 = hi ? hi_g[d] : lo_f[d] - 1;
  for (int ta = lo_g[a]; ta <= hi_g[a]; ta++)
    for (int tc = lo_g[c]; tc <= hi_g[c]; tc++) {
      int e[3], n[3];
      e[d] = edge; e[a] = ta; e[c] = tc;
      n[d] = inner; n[a] = ta; n[c] = tc;
      double pb = AT(P, e[0], e[1], e[2]);
      double slope = 0.0;
      if (kind == BC_LINEAR) slope = pb - AT(P, n[0], n[1], n[2]);
      if (kind == BC_SIGNED) slope = copysign(1.0, pb) * fabs(pb - AT(P, n[0], n[1], n[2]));
      for (int g = g0; g <= g1; g++) {
        int q[3];
        q[d] = g; q[a] = ta; q[c] = tc;
        if (kind == BC_COPY) {
          AT(P, q[0], q[1], q[2]) = pb;
        } else {
          double dist = (double)(hi ? g - edge : edge - g);
          AT(P, q[0], q[1], q[2]) = pb + slope * dist;
        }
      }
    }
}

void oracle_lsm3dlinearextrapolation(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  if (*idx >= 0 && *idx <= 5) orc_bc_face(VIEW(phi, gb), BOX(gb), BOX(fb), *idx, BC_LINEAR);
}
void oracle_lsm3dsignedlinearextrapolation(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  if (*idx >= 0 && *idx <= 5) orc_bc_face(VIEW(phi, gb), BOX(gb), BOX(fb), *idx, BC_SIGNED);
}
void oracle_lsm3dcopyextrapolation(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  if (*idx >= 0 && *idx <= 5) orc_bc_face(VIEW(phi, gb), BOX(gb), BOX(fb), *idx, BC_COPY);
}
/* lsm_boundary_conditions3d.f:486-638 -- all four are copy extrapolation */
void oracle_lsm3dhomogeneousneumanneno1(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  oracle_lsm3dcopyextrapolation(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb, ilo_fb, ihi_fb, jlo_fb,
                                jhi_fb, klo_fb, khi_fb, idx);
}
void oracle_lsm3dhomogeneousneumanneno2(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  oracle_lsm3dcopyextrapolation(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb, ilo_fb, ihi_fb, jlo_fb,
                                jhi_fb, klo_fb, khi_fb, idx);
}
void oracle_lsm3dhomogeneousneumanneno3(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  oracle_lsm3dcopyextrapolation(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb, ilo_fb, ihi_fb, jlo_fb,
                                jhi_fb, klo_fb, khi_fb, idx);
}
void oracle_lsm3dhomogeneousneumannweno5(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  oracle_lsm3dcopyextrapolation(phi, ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb, ilo_fb, ihi_fb, jlo_fb,
                                jhi_fb, klo_fb, khi_fb, idx);
}

/* Periodic ghost fill (no reference counterpart: the serial package has no
 * periodic boundary condition).  Face idx copies the interior layers that lie
 * one period away; like the reference faces it sweeps the full ghost box of
 * the two transverse axes, so faces applied in the order x, y, z fill edges
 * and corners periodically as well. */
void oracle_lsmhip_periodic_face(double *phi, BOXP(gb), BOXP(fb), const int *idx) {
  orc_arr P = VIEW(phi, gb);
  obox g = BOX(gb), f = BOX(fb);
  int d = *idx / 2, hi = *idx % 2;
  int lo_g[3] = {g.ilo, g.jlo, g.klo}, hi_g[3] = {g.ihi, g.jhi, g.khi};
  int lo_f[3] = {f.ilo, f.jlo, f.klo}, hi_f[3] = {f.ihi, f.jhi, f.khi};
  int period = hi_f[d] - lo_f[d] + 1;
  int a = (d + 1) % 3, c = (d + 2) % 3;
  int g0 = hi ? hi_f[d] + 1 : lo_g[d], g1 = hi ? hi_g[d] : lo_f[d] - 1;
  for (int ta = lo_g[a]; ta <= hi_g[a]; ta++)
    for (int tc = lo_g[c]; tc <= hi_g[c]; tc++)
      for (int gi = g0; gi <= g1; gi++) {
        int q[3], s[3];
        q[d] = gi; q[a] = ta; q[c] = tc;
        s[d] = hi ? gi - period : gi + period; s[a] = ta; s[c] = tc;
        AT(P, q[0], q[1], q[2]) = AT(P, s[0], s[1], s[2]);
      }
}

/* ---------------------------------------------------------------------------
 * Smoothed Heaviside / delta, src/toolbox/lsm_calculus_toolbox.f:25-92
 * ------------------------------------------------------------------------- */
void oracle_lsmheaviside(double *H, const double *x, const double *epsilon) {
  double e = *epsilon, v = *x;
  if (v < -e) *H = 0.0;
  else if (v > e) *H = 1.0;
  else {
    double ie = 1.0 / e;
    *H = 0.5 * (1.0 + v * ie + ORC_ONE_OVER_PI * sin(ORC_PI * v * ie));
  }
}
void oracle_lsmdeltafunction(double *delta, const double *x, const double *epsilon) {
  double e = *epsilon, v = *x;
  if (fabs(v) > e) *delta = 0.0;
  else {
    double ie = 1.0 / e;
    *delta = 0.5 * ie * (1.0 + cos(ORC_PI * v * ie));
  }
}

/* ---------------------------------------------------------------------------
 * Averages, voxel counts and control-volume forms, src/toolbox/lsm_math_utils3d.f.in
 * and src/geometry/lsm_geometry3d_fort.f.in.  A control-volume routine visits
 * the cells of ib whose control_vol is > 0 (control_vol_sgn > 0) or < 0
 * (otherwise) and does there what its whole-box twin does.
 * ------------------------------------------------------------------------- */
#define CV_ARGS const double *cv, BOXP(cg), const int *cvs
#define CV_VIEW orc_arr CV = VIEW(cv, cg)
#define CV_IN (*cvs > 0 ? AT(CV, i, j, k) > 0.0 : AT(CV, i, j, k) < 0.0)

/* lsm_math_utils3d.f.in:81-148: serial sum of |f1 - f2| and count, then their
 * quotient (0 for an empty box) */
void oracle_lsm3daveabsdiff(double *out, const double *f1, BOXP(g1), const double *f2, BOXP(g2), BOXP(ib)) {
  orc_arr A = VIEW(f1, g1), B = VIEW(f2, g2);
  double s = 0.0, n = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    s = s + fabs(AT(A, i, j, k) - AT(B, i, j, k));
    n = n + 1.0;
  }
  *out = n > 0.0 ? s / n : 0.0;
}

/* :154-198 and :203-247 */
void oracle_lsm3dvoxelcountgreaterthanzero(int *count, const double *phi, BOXP(pg), BOXP(fb)) {
  orc_arr P = VIEW(phi, pg);
  int c = 0;
  obox b = BOX(fb);
  FOR_FB(b) if (AT(P, i, j, k) > 0.0) c++;
  *count = c;
}
void oracle_lsm3dvoxelcountlessthanzero(int *count, const double *phi, BOXP(pg), BOXP(fb)) {
  orc_arr P = VIEW(phi, pg);
  int c = 0;
  obox b = BOX(fb);
  FOR_FB(b) if (AT(P, i, j, k) < 0.0) c++;
  *count = c;
}

/* :852-955: starts from the first include-box cell whatever its control volume */
void oracle_lsm3dmaxnormdiffcontrolvolume(double *out, const double *f1, BOXP(g1), const double *f2, BOXP(g2),
                                          CV_ARGS, BOXP(ib)) {
  orc_arr A = VIEW(f1, g1), B = VIEW(f2, g2);
  CV_VIEW;
  double m = fabs(AT(A, *ilo_ib, *jlo_ib, *klo_ib) - AT(B, *ilo_ib, *jlo_ib, *klo_ib));
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double d = fabs(AT(A, i, j, k) - AT(B, i, j, k));
    if (d > m) m = d;
  }
  *out = m;
}

/* :960-1075 */
void oracle_lsm3dcomputestableadvectiondtcontrolvolume(double *dt, const double *vx, const double *vy,
                                                       const double *vz, BOXP(vg), CV_ARGS, BOXP(ib),
                                                       const double *dx, const double *dy, const double *dz,
                                                       const double *cfl) {
  orc_arr U = VIEW(vx, vg), V = VIEW(vy, vg), W = VIEW(vz, vg);
  CV_VIEW;
  double m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double c = fabs(AT(U, i, j, k)) * ix + fabs(AT(V, i, j, k)) * iy + fabs(AT(W, i, j, k)) * iz;
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

/* :1080-1250 (vel_n a field) and :1255-1415 (a constant) */
static void normal_dt_cv(double *dt, const double *vn, orc_arr VN, double vconst, const double *pxp,
                         const double *pyp, const double *pzp, BOXP(gp), const double *pxm, const double *pym,
                         const double *pzm, BOXP(gm), CV_ARGS, BOXP(ib), const double *dx, const double *dy,
                         const double *dz, const double *cfl) {
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  CV_VIEW;
  double h = orc_max(orc_max(*dx, *dy), *dz), mdx2 = h * h;
  double m = -1.0, ix = 1.0 / *dx, iy = 1.0 / *dy, iz = 1.0 / *dz;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double av = vn ? fabs(AT(VN, i, j, k)) : vconst;
    double c = normal_h_over_dx(av, AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k), AT(XM, i, j, k),
                                AT(YM, i, j, k), AT(ZM, i, j, k), mdx2, ix, iy, iz);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}
void oracle_lsm3dcomputestablenormalveldtcontrolvolume(double *dt, const double *vn, BOXP(vg), const double *pxp,
                                                       const double *pyp, const double *pzp, BOXP(gp),
                                                       const double *pxm, const double *pym, const double *pzm,
                                                       BOXP(gm), CV_ARGS, BOXP(ib), const double *dx,
                                                       const double *dy, const double *dz, const double *cfl) {
  normal_dt_cv(dt, vn, VIEW(vn, vg), 0.0, pxp, pyp, pzp, ilo_gp, ihi_gp, jlo_gp, jhi_gp, klo_gp, khi_gp, pxm, pym,
               pzm, ilo_gm, ihi_gm, jlo_gm, jhi_gm, klo_gm, khi_gm, cv, ilo_cg, ihi_cg, jlo_cg, jhi_cg, klo_cg,
               khi_cg, cvs, ilo_ib, ihi_ib, jlo_ib, jhi_ib, klo_ib, khi_ib, dx, dy, dz, cfl);
}
void oracle_lsm3dcomputestableconstnormalveldtcontrolvolume(double *dt, const double *vel_n, const double *pxp,
                                                            const double *pyp, const double *pzp, BOXP(gp),
                                                            const double *pxm, const double *pym,
                                                            const double *pzm, BOXP(gm), CV_ARGS, BOXP(ib),
                                                            const double *dx, const double *dy, const double *dz,
                                                            const double *cfl) {
  orc_arr none = {0};
  normal_dt_cv(dt, (const double *)0, none, fabs(*vel_n), pxp, pyp, pzp, ilo_gp, ihi_gp, jlo_gp, jhi_gp, klo_gp, khi_gp, pxm,
               pym, pzm, ilo_gm, ihi_gm, jlo_gm, jhi_gm, klo_gm, khi_gm, cv, ilo_cg, ihi_cg, jlo_cg, jhi_cg, klo_cg,
               khi_cg, cvs, ilo_ib, ihi_ib, jlo_ib, jhi_ib, klo_ib, khi_ib, dx, dy, dz, cfl);
}

/* :1420-1551 and :1556-1686 (naive serial sums, as the reference's) */
void oracle_lsm3dvolumeintegralphilessthanzerocontrolvolume(double *intF, const double *F, BOXP(fg),
                                                            const double *phi, BOXP(pg), CV_ARGS, BOXP(ib),
                                                            const double *dx, const double *dy, const double *dz,
                                                            const double *epsilon) {
  orc_arr FF = VIEW(F, fg), P = VIEW(phi, pg);
  CV_VIEW;
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double p = AT(P, i, j, k), q = p / e;
    if (p < -e) s = s + AT(FF, i, j, k) * dV;
    else if (p < e) s = s + 0.5 * (1.0 - q - ORC_ONE_OVER_PI * sin(ORC_PI * q)) * AT(FF, i, j, k) * dV;
  }
  *intF = s;
}
void oracle_lsm3dvolumeintegralphigreaterthanzerocontrolvolume(double *intF, const double *F, BOXP(fg),
                                                               const double *phi, BOXP(pg), CV_ARGS, BOXP(ib),
                                                               const double *dx, const double *dy,
                                                               const double *dz, const double *epsilon) {
  orc_arr FF = VIEW(F, fg), P = VIEW(phi, pg);
  CV_VIEW;
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double p = AT(P, i, j, k), q = p / e;
    if (p > e) s = s + AT(FF, i, j, k) * dV;
    else if (p > -e) s = s + 0.5 * (1.0 + q + ORC_ONE_OVER_PI * sin(ORC_PI * q)) * AT(FF, i, j, k) * dV;
  }
  *intF = s;
}

/* :1691-1844 */
void oracle_lsm3dsurfaceintegralcontrolvolume(double *intF, const double *F, BOXP(fg), const double *phi,
                                              BOXP(pg), const double *px, const double *py, const double *pz,
                                              BOXP(gg), CV_ARGS, BOXP(ib), const double *dx, const double *dy,
                                              const double *dz, const double *epsilon) {
  orc_arr FF = VIEW(F, fg), P = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg);
  CV_VIEW;
  double dV = *dx * *dy * *dz, e = *epsilon, ie = 1.0 / e, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double p = AT(P, i, j, k);
    if (fabs(p) < e) {
      double delta = 0.5 * ie * (1.0 + cos(ORC_PI * p * ie));
      double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
      double n = sqrt(x * x + y * y + z * z);
      s = s + delta * n * AT(FF, i, j, k) * dV;
    }
  }
  *intF = s;
}

/* :1849-1925 and :1931-2007: these two take control_vol >= 0 (sgn > 0) and
 * <= 0 (otherwise) */
void oracle_lsm3dvoxelcountgreaterthanzerocontrolvolume(int *count, const double *phi, BOXP(pg), CV_ARGS,
                                                        BOXP(fb)) {
  orc_arr P = VIEW(phi, pg);
  CV_VIEW;
  int c = 0;
  obox b = BOX(fb);
  FOR_FB(b) {
    double v = AT(CV, i, j, k);
    if (AT(P, i, j, k) > 0.0 && (*cvs > 0 ? v >= 0.0 : v <= 0.0)) c++;
  }
  *count = c;
}
void oracle_lsm3dvoxelcountlessthanzerocontrolvolume(int *count, const double *phi, BOXP(pg), CV_ARGS,
                                                     BOXP(fb)) {
  orc_arr P = VIEW(phi, pg);
  CV_VIEW;
  int c = 0;
  obox b = BOX(fb);
  FOR_FB(b) {
    double v = AT(CV, i, j, k);
    if (AT(P, i, j, k) < 0.0 && (*cvs > 0 ? v >= 0.0 : v <= 0.0)) c++;
  }
  *count = c;
}

/* lsm_geometry3d_fort.f.in:669-745: delta_phi |grad phi| dV over the cells
 * where delta_phi > 0 */
void oracle_lsm3dsurfaceareazerolevelsetdelta(double *area, const double *dphi, BOXP(pg), const double *px,
                                              const double *py, const double *pz, BOXP(gg), BOXP(ib),
                                              const double *dx, const double *dy, const double *dz) {
  orc_arr D = VIEW(dphi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg);
  double dV = *dx * *dy * *dz, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double d = AT(D, i, j, k);
    if (d > 0.0) {
      double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
      s = s + d * sqrt(x * x + y * y + z * z) * dV;
    }
  }
  *area = s;
}

/* :770-892 and :916-1035 */
void oracle_lsm3dvolumeregionphilessthanzerocontrolvolume(double *vol, const double *phi, BOXP(pg), CV_ARGS,
                                                          BOXP(ib), const double *dx, const double *dy,
                                                          const double *dz, const double *epsilon) {
  orc_arr P = VIEW(phi, pg);
  CV_VIEW;
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double p = AT(P, i, j, k);
    if (p < e) s = s + vol_lt0_term(p, e, dV);
  }
  *vol = s;
}
void oracle_lsm3dvolumeregionphigreaterthanzerocontrolvolume(double *vol, const double *phi, BOXP(pg), CV_ARGS,
                                                             BOXP(ib), const double *dx, const double *dy,
                                                             const double *dz, const double *epsilon) {
  orc_arr P = VIEW(phi, pg);
  CV_VIEW;
  double dV = *dx * *dy * *dz, e = *epsilon, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double p = AT(P, i, j, k);
    if (p > -e) s = s + vol_gt0_term(p, e, dV);
  }
  *vol = s;
}

/* :1061-1204 */
void oracle_lsm3dsurfaceareazerolevelsetcontrolvolume(double *area, const double *phi, BOXP(pg),
                                                      const double *px, const double *py, const double *pz,
                                                      BOXP(gg), CV_ARGS, BOXP(ib), const double *dx,
                                                      const double *dy, const double *dz, const double *epsilon) {
  orc_arr P = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg);
  CV_VIEW;
  double dV = *dx * *dy * *dz, e = *epsilon, ie = 1.0 / e, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    if (!CV_IN) continue;
    double p = AT(P, i, j, k);
    if (fabs(p) < e) s = s + area_term(p, AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k), e, ie, dV);
  }
  *area = s;
}

/* :1209-1329 */
void oracle_lsm3dsurfaceareazerolevelsetdeltacontrolvolume(double *area, const double *dphi, BOXP(pg),
                                                           const double *px, const double *py, const double *pz,
                                                           BOXP(gg), CV_ARGS, BOXP(ib), const double *dx,
                                                           const double *dy, const double *dz) {
  orc_arr D = VIEW(dphi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg);
  CV_VIEW;
  double dV = *dx * *dy * *dz, s = 0.0;
  obox b = BOX(ib);
  FOR_FB(b) {
    double d = AT(D, i, j, k);
    if (CV_IN && d > 0.0) {
      double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
      s = s + d * sqrt(x * x + y * y + z * z) * dV;
    }
  }
  *area = s;
}

/* ---------------------------------------------------------------------------
 * Hessian, phi-upwind gradient, first-order delta function
 * ------------------------------------------------------------------------- */
/* lsm3dCentralHessian, src/toolbox/lsm_spatial_derivatives3d.f.in:3658-3715 */
void oracle_lsm3dcentralhessian(double *pxx, double *pxy, double *pxz, double *pyy, double *pyz, double *pzz,
                                BOXP(hg), const double *phi, BOXP(pg), BOXP(fb), const double *dx,
                                const double *dy, const double *dz) {
  orc_arr XX = VIEW(pxx, hg), XY = VIEW(pxy, hg), XZ = VIEW(pxz, hg), YY = VIEW(pyy, hg), YZ = VIEW(pyz, hg),
          ZZ = VIEW(pzz, hg), P = VIEW(phi, pg);
  obox b = BOX(fb);
  FOR_FB(b) {
    double c = AT(P, i, j, k);
    AT(XX, i, j, k) = (AT(P, i + 1, j, k) - 2.0 * c + AT(P, i - 1, j, k)) / (*dx * *dx);
    AT(XY, i, j, k) = 0.25 * (AT(P, i + 1, j + 1, k) - AT(P, i - 1, j + 1, k) - AT(P, i + 1, j - 1, k) +
                              AT(P, i - 1, j - 1, k)) / (*dx * *dy);
    AT(XZ, i, j, k) = 0.25 * (AT(P, i + 1, j, k + 1) - AT(P, i - 1, j, k + 1) - AT(P, i + 1, j, k - 1) +
                              AT(P, i - 1, j, k - 1)) / (*dx * *dz);
    AT(YY, i, j, k) = (AT(P, i, j + 1, k) - 2.0 * c + AT(P, i, j - 1, k)) / (*dy * *dy);
    AT(YZ, i, j, k) = 0.25 * (AT(P, i, j + 1, k + 1) - AT(P, i, j - 1, k + 1) - AT(P, i, j + 1, k - 1) +
                              AT(P, i, j - 1, k - 1)) / (*dy * *dz);
    AT(ZZ, i, j, k) = (AT(P, i, j, k + 1) - 2.0 * c + AT(P, i, j, k - 1)) / (*dz * *dz);
  }
}

/* One axis of lsm3dPhiUpwindGradF (:3270-3458): the side whose phi is nearer
 * zero (smaller for phi > 0, larger for phi < 0), the plus side on a tie of
 * the two neighbours only where the minus side does not qualify. */
static double phi_upwind_pick(double c, double nm, double np, double fp, double fm) {
  if (c > 0.0) {
    if (nm <= c) return np < nm ? fp : fm;
    if (np <= c) return fp;
    return 0.0;
  }
  if (nm >= c) return np > nm ? fp : fm;
  if (np >= c) return fp;
  return 0.0;
}
void oracle_lsm3dphiupwindgradf(double *Fx, double *Fy, double *Fz, BOXP(g), const double *Fxp,
                                const double *Fyp, const double *Fzp, BOXP(gp), const double *Fxm,
                                const double *Fym, const double *Fzm, BOXP(gm), const double *phi, BOXP(pg),
                                BOXP(fb)) {
  orc_arr X = VIEW(Fx, g), Y = VIEW(Fy, g), Z = VIEW(Fz, g), XP = VIEW(Fxp, gp), YP = VIEW(Fyp, gp),
          ZP = VIEW(Fzp, gp), XM = VIEW(Fxm, gm), YM = VIEW(Fym, gm), ZM = VIEW(Fzm, gm), P = VIEW(phi, pg);
  obox b = BOX(fb);
  FOR_FB(b) {
    double c = AT(P, i, j, k);
    if (!(c > 0.0) && !(c < 0.0)) continue; /* phi == 0 (or NaN): left as is */
    AT(X, i, j, k) = phi_upwind_pick(c, AT(P, i - 1, j, k), AT(P, i + 1, j, k), AT(XP, i, j, k), AT(XM, i, j, k));
    AT(Y, i, j, k) = phi_upwind_pick(c, AT(P, i, j - 1, k), AT(P, i, j + 1, k), AT(YP, i, j, k), AT(YM, i, j, k));
    AT(Z, i, j, k) = phi_upwind_pick(c, AT(P, i, j, k - 1), AT(P, i, j, k + 1), AT(ZP, i, j, k), AT(ZM, i, j, k));
  }
}

/* lsm3dDeltaFunctionOrder1, src/toolbox/lsm_calculus_toolbox3d.f.in:15-141: the
 * sum of |phi_n N / dx / (phi_n - phi)| over the six neighbours n where phi
 * changes sign (phi_n phi <= 0); the z terms use norm_phi_x and 1/dx. */
static double delta_side(double pn, double c, double nrm, double inv) {
  if (!(pn * c <= 0.0)) return 0.0;
  double t = pn * nrm * inv;
  return t / (pn - c);
}
void oracle_lsm3ddeltafunctionorder1(const double *phi, double *delta, BOXP(pg), const double *nx,
                                     const double *ny, const double *nz, BOXP(gg), BOXP(fb), const double *dx,
                                     const double *dy, const double *dz) {
  orc_arr P = VIEW(phi, pg), D = VIEW(delta, pg), NX = VIEW(nx, gg), NY = VIEW(ny, gg);
  double ix = 1.0 / *dx, iy = 1.0 / *dy;
  obox b = BOX(fb);
  FOR_FB(b) {
    double c = AT(P, i, j, k), x = AT(NX, i, j, k), y = AT(NY, i, j, k);
    double xp = delta_side(AT(P, i + 1, j, k), c, x, ix), xm = delta_side(AT(P, i - 1, j, k), c, x, ix);
    double yp = delta_side(AT(P, i, j + 1, k), c, y, iy), ym = delta_side(AT(P, i, j - 1, k), c, y, iy);
    double zp = delta_side(AT(P, i, j, k + 1), c, x, ix), zm = delta_side(AT(P, i, j, k - 1), c, x, ix);
    AT(D, i, j, k) = fabs(xp) + fabs(xm) + fabs(yp) + fabs(ym) + fabs(zp) + fabs(zm);
  }
}
