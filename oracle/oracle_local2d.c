/*
 * oracle_local2d.c -- TEST INFRASTRUCTURE ONLY (see lsm_oracle.h).
 *
 * The 2D narrow-band ("local") kernels and 2D band construction.  Index lists
 * are addressed as the reference's Fortran declarations address them (an
 * array declared index_x(nlo:nhi) is passed as a pointer to element nlo);
 * marks are integer*1 values compared as signed bytes.  Each function cites
 * the reference routine whose operation order it keeps.
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
static orc_mask m2(const unsigned char *p, const int *ilo, const int *ihi, const int *jlo, const int *jhi) {
  orc_mask a;
  a.p = (unsigned char *)p;
  a.ilo = *ilo;
  a.jlo = *jlo;
  a.klo = 0;
  a.nx = (long)(*ihi - *ilo + 1);
  a.nxy = a.nx * (long)(*jhi - *jlo + 1);
  return a;
}
#define V2(p, n) v2((p), ilo_##n, ihi_##n, jlo_##n, jhi_##n)
#define M2(p, n) m2((p), ilo_##n, ihi_##n, jlo_##n, jhi_##n)
#define A2(a, i, j) AT(a, i, j, 0)
#define MK2(a, i, j) ((signed char)AT(a, i, j, 0))
/* Points l = lo..hi of a list whose first element is index lo, mark-filtered. */
#define FOR_BAND2(lo, hi, nbv, mark)                                             \
  for (int l = (lo); l <= (hi); l++)                                             \
    for (int i = LIDX(ix, (lo), l), j = LIDX(iy, (lo), l), _once = 1;            \
         _once && MK2(nbv, i, j) <= (signed char)(mark); _once = 0)
/* Offset of the stencil neighbour s along axis d. */
#define NB2(a, d, i, j, s) A2(a, (i) + ((d) == 0 ? (s) : 0), (j) + ((d) == 1 ? (s) : 0))

/* ---------------------------------------------------------------------------
 * Spatial derivatives, src/toolbox/lsm_spatial_derivatives2d_local.f.in
 * ------------------------------------------------------------------------- */
/* lsm2dComputeDnLOCAL :43-127 -- at listed points with mark <= markD,
 * Dn = sgn (Dm1(i) - Dm1(i - off)); elsewhere on the list 1e10. */
static void dn2(orc_arr Dn, orc_arr Dm1, int n, int d, const int *ix, const int *iy, int lo, int hi, int base,
                orc_mask nb, signed char markD) {
  int off = n % 2 == 1 ? 1 : -1;
  double sgn = n % 2 == 1 ? 1.0 : -1.0;
  for (int l = lo; l <= hi; l++) {
    int i = LIDX(ix, base, l), j = LIDX(iy, base, l);
    if (MK2(nb, i, j) <= markD)
      A2(Dn, i, j) = sgn * (A2(Dm1, i, j) - A2(Dm1, i - (d == 0 ? off : 0), j - (d == 1 ? off : 0)));
    else
      A2(Dn, i, j) = ORC_BIG;
  }
}

/* lsm2dHJENO1LOCAL :133-274 */
void oracle_lsm2dhjeno1local(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                             const double *phi, BOXP2(pg), double *D1, BOXP2(d1), const double *dx, const double *dy,
                             const int *ix, const int *iy, const int *nlo0, const int *nhi0, const int *nlo1,
                             const int *nhi1, const unsigned char *nb, BOXP2(nbg), const unsigned char *mark_fb,
                             const unsigned char *mark_D1) {
  double inv[2] = {1.0 / *dx, 1.0 / *dy};
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)}, F = V2(phi, pg), D = V2(D1, d1);
  orc_mask N = M2(nb, nbg);
  int base = *nlo0;
  (void)nlo1;
  for (int d = 0; d < 2; d++) {
    dn2(D, F, 1, d, ix, iy, *nlo0, *nhi1, base, N, (signed char)*mark_D1);
    FOR_BAND2(*nlo0, *nhi0, N, *mark_fb) {
      A2(P[d], i, j) = NB2(D, d, i, j, 1) * inv[d];
      A2(M[d], i, j) = NB2(D, d, i, j, 0) * inv[d];
    }
  }
}

/* ENO2 plus/minus selections of :375-384 (the 3D order of the whole-grid ENO2) */
static double eno2p(orc_arr A, orc_arr B, int d, int i, int j) {
  if (fabs(NB2(B, d, i, j, 0)) < fabs(NB2(B, d, i, j, 1))) return NB2(A, d, i, j, 1) - 0.5 * NB2(B, d, i, j, 0);
  return NB2(A, d, i, j, 1) - 0.5 * NB2(B, d, i, j, 1);
}
static double eno2m(orc_arr A, orc_arr B, int d, int i, int j) {
  if (fabs(NB2(B, d, i, j, -1)) < fabs(NB2(B, d, i, j, 0))) return NB2(A, d, i, j, 0) + 0.5 * NB2(B, d, i, j, -1);
  return NB2(A, d, i, j, 0) + 0.5 * NB2(B, d, i, j, 0);
}

/* lsm2dHJENO2LOCAL :280-493: D1 over levels 0..2, D2 over 0..1 */
void oracle_lsm2dhjeno2local(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                             const double *phi, BOXP2(pg), double *D1, BOXP2(d1), double *D2, BOXP2(d2),
                             const double *dx, const double *dy, const int *ix, const int *iy, const int *nlo0,
                             const int *nhi0, const int *nlo1, const int *nhi1, const int *nlo2, const int *nhi2,
                             const unsigned char *nb, BOXP2(nbg), const unsigned char *mark_fb,
                             const unsigned char *mark_D1, const unsigned char *mark_D2) {
  double inv[2] = {1.0 / *dx, 1.0 / *dy};
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)}, F = V2(phi, pg), A = V2(D1, d1),
          B = V2(D2, d2);
  orc_mask N = M2(nb, nbg);
  int base = *nlo0;
  (void)nlo1;
  (void)nlo2;
  for (int d = 0; d < 2; d++) {
    dn2(A, F, 1, d, ix, iy, *nlo0, *nhi2, base, N, (signed char)*mark_D1);
    dn2(B, A, 2, d, ix, iy, *nlo0, *nhi1, base, N, (signed char)*mark_D2);
    FOR_BAND2(*nlo0, *nhi0, N, *mark_fb) {
      A2(P[d], i, j) = eno2p(A, B, d, i, j) * inv[d];
      A2(M[d], i, j) = eno2m(A, B, d, i, j) * inv[d];
    }
  }
}

/* lsm2dHJENO3LOCAL :498-811: D1 over levels 0..3, D2 over 0..2, D3 over 0..1;
 * the output accumulates term by term. */
void oracle_lsm2dhjeno3local(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                             const double *phi, BOXP2(pg), double *D1, BOXP2(d1), double *D2, BOXP2(d2),
                             double *D3, BOXP2(d3), const double *dx, const double *dy, const int *ix, const int *iy,
                             const int *nlo0, const int *nhi0, const int *nlo1, const int *nhi1, const int *nlo2,
                             const int *nhi2, const int *nlo3, const int *nhi3, const unsigned char *nb, BOXP2(nbg),
                             const unsigned char *mark_fb, const unsigned char *mark_D1,
                             const unsigned char *mark_D2, const unsigned char *mark_D3) {
  const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
  double inv[2] = {1.0 / *dx, 1.0 / *dy};
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)}, F = V2(phi, pg), A = V2(D1, d1),
          B = V2(D2, d2), C = V2(D3, d3);
  orc_mask N = M2(nb, nbg);
  int base = *nlo0;
  (void)nlo1;
  (void)nlo2;
  (void)nlo3;
  for (int d = 0; d < 2; d++) {
    dn2(A, F, 1, d, ix, iy, *nlo0, *nhi3, base, N, (signed char)*mark_D1);
    dn2(B, A, 2, d, ix, iy, *nlo0, *nhi2, base, N, (signed char)*mark_D2);
    dn2(C, B, 3, d, ix, iy, *nlo0, *nhi1, base, N, (signed char)*mark_D3);
    FOR_BAND2(*nlo0, *nhi0, N, *mark_fb) {
      double r = NB2(A, d, i, j, 1);
      if (fabs(NB2(B, d, i, j, 0)) < fabs(NB2(B, d, i, j, 1))) {
        r = r - 0.5 * NB2(B, d, i, j, 0);
        if (fabs(NB2(C, d, i, j, 0)) < fabs(NB2(C, d, i, j, 1))) r = r - sixth * NB2(C, d, i, j, 0);
        else r = r - sixth * NB2(C, d, i, j, 1);
      } else {
        r = r - 0.5 * NB2(B, d, i, j, 1);
        if (fabs(NB2(C, d, i, j, 1)) < fabs(NB2(C, d, i, j, 2))) r = r + third * NB2(C, d, i, j, 1);
        else r = r + third * NB2(C, d, i, j, 2);
      }
      A2(P[d], i, j) = r * inv[d];
      r = NB2(A, d, i, j, 0);
      if (fabs(NB2(B, d, i, j, -1)) < fabs(NB2(B, d, i, j, 0))) {
        r = r + 0.5 * NB2(B, d, i, j, -1);
        if (fabs(NB2(C, d, i, j, -1)) < fabs(NB2(C, d, i, j, 0))) r = r + third * NB2(C, d, i, j, -1);
        else r = r + third * NB2(C, d, i, j, 0);
      } else {
        r = r + 0.5 * NB2(B, d, i, j, 0);
        if (fabs(NB2(C, d, i, j, 0)) < fabs(NB2(C, d, i, j, 1))) r = r - sixth * NB2(C, d, i, j, 0);
        else r = r - sixth * NB2(C, d, i, j, 1);
      }
      A2(M[d], i, j) = r * inv[d];
    }
  }
}

/* HJ-WENO5 with the 2D local routine's literal tiny number 1e-35 (:832). */
static double weno5_35(double v1, double v2, double v3, double v4, double v5) {
  const double one_third = 1.0 / 3.0, seven_sixths = 7.0 / 6.0, eleven_sixths = 11.0 / 6.0;
  const double one_sixth = 1.0 / 6.0, five_sixths = 5.0 / 6.0;
  const double thirteen_twelfths = 13.0 / 12.0, one_fourth = 0.25;
  double m = v1 * v1;
  if (v2 * v2 > m) m = v2 * v2;
  if (v3 * v3 > m) m = v3 * v3;
  if (v4 * v4 > m) m = v4 * v4;
  if (v5 * v5 > m) m = v5 * v5;
  double eps = 1e-6 * m + 1.e-35;
  double p1 = one_third * v1 - seven_sixths * v2 + eleven_sixths * v3;
  double p2 = -(one_sixth * v2) + five_sixths * v3 + one_third * v4;
  double p3 = one_third * v3 + five_sixths * v4 - one_sixth * v5;
  double S1 = thirteen_twelfths * orc_sq(v1 - 2.0 * v2 + v3) + one_fourth * orc_sq(v1 - 4.0 * v2 + 3.0 * v3);
  double S2 = thirteen_twelfths * orc_sq(v2 - 2.0 * v3 + v4) + one_fourth * orc_sq(v2 - v4);
  double S3 = thirteen_twelfths * orc_sq(v3 - 2.0 * v4 + v5) + one_fourth * orc_sq(3.0 * v3 - 4.0 * v4 + v5);
  double a1 = 0.1 / orc_sq(S1 + eps), a2 = 0.6 / orc_sq(S2 + eps), a3 = 0.3 / orc_sq(S3 + eps);
  double inv = 1.0 / (a1 + a2 + a3);
  a1 = a1 * inv;
  a2 = a2 * inv;
  a3 = a3 * inv;
  return a1 * p1 + a2 * p2 + a3 * p3;
}

/* lsm2dHJWENO5LOCAL :816-1154: D1 over levels 0..3, WENO5 on level 0 */
void oracle_lsm2dhjweno5local(double *pxp, double *pyp, BOXP2(gp), double *pxm, double *pym, BOXP2(gm),
                              const double *phi, BOXP2(pg), double *D1, BOXP2(d1), const double *dx, const double *dy,
                              const int *ix, const int *iy, const int *nlo0, const int *nhi0, const int *nlo1,
                              const int *nhi1, const int *nlo2, const int *nhi2, const int *nlo3, const int *nhi3,
                              const unsigned char *nb, BOXP2(nbg), const unsigned char *mark_fb,
                              const unsigned char *mark_D1) {
  double inv[2] = {1.0 / *dx, 1.0 / *dy};
  orc_arr P[2] = {V2(pxp, gp), V2(pyp, gp)}, M[2] = {V2(pxm, gm), V2(pym, gm)}, F = V2(phi, pg), A = V2(D1, d1);
  orc_mask N = M2(nb, nbg);
  int base = *nlo0;
  (void)nlo1; (void)nhi1; (void)nlo2; (void)nhi2; (void)nlo3;
  for (int d = 0; d < 2; d++) {
    dn2(A, F, 1, d, ix, iy, *nlo0, *nhi3, base, N, (signed char)*mark_D1);
    FOR_BAND2(*nlo0, *nhi0, N, *mark_fb) {
      double s = inv[d];
      A2(P[d], i, j) = weno5_35(NB2(A, d, i, j, 3) * s, NB2(A, d, i, j, 2) * s, NB2(A, d, i, j, 1) * s,
                                NB2(A, d, i, j, 0) * s, NB2(A, d, i, j, -1) * s);
      A2(M[d], i, j) = weno5_35(NB2(A, d, i, j, -2) * s, NB2(A, d, i, j, -1) * s, NB2(A, d, i, j, 0) * s,
                                NB2(A, d, i, j, 1) * s, NB2(A, d, i, j, 2) * s);
    }
  }
}

/* lsm2dUpwindHJENO2LOCAL :1160-1386: D1 over levels 0..2, D2 over 0..1 */
void oracle_lsm2dupwindhjeno2local(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg), const double *vx,
                                   const double *vy, BOXP2(vg), double *D1, BOXP2(d1), double *D2, BOXP2(d2),
                                   const double *dx, const double *dy, const int *ix, const int *iy, const int *nlo0,
                                   const int *nhi0, const int *nlo1, const int *nhi1, const int *nlo2,
                                   const int *nhi2, const unsigned char *nb, BOXP2(nbg),
                                   const unsigned char *mark_fb, const unsigned char *mark_D1,
                                   const unsigned char *mark_D2) {
  double inv[2] = {1.0 / *dx, 1.0 / *dy};
  orc_arr O[2] = {V2(px, g), V2(py, g)}, V[2] = {V2(vx, vg), V2(vy, vg)}, F = V2(phi, pg), A = V2(D1, d1),
          B = V2(D2, d2);
  orc_mask N = M2(nb, nbg);
  int base = *nlo0;
  (void)nlo1;
  (void)nlo2;
  for (int d = 0; d < 2; d++) {
    dn2(A, F, 1, d, ix, iy, *nlo0, *nhi2, base, N, (signed char)*mark_D1);
    dn2(B, A, 2, d, ix, iy, *nlo0, *nhi1, base, N, (signed char)*mark_D2);
    FOR_BAND2(*nlo0, *nhi0, N, *mark_fb) {
      double v = A2(V[d], i, j);
      if (fabs(v) < ORC_ZERO_TOL) A2(O[d], i, j) = 0.0;
      else if (v > 0) A2(O[d], i, j) = eno2m(A, B, d, i, j) * inv[d];
      else A2(O[d], i, j) = eno2p(A, B, d, i, j) * inv[d];
    }
  }
}

#define LOCAL2_TAIL \
  const int *ix, const int *iy, const int *nlo, const int *nhi, const unsigned char *nb, BOXP2(nbg), \
      const unsigned char *mark_fb

/* lsm2dCentralGradOrder2LOCAL :1392-1460 */
void oracle_lsm2dcentralgradorder2local(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg),
                                        const double *dx, const double *dy, LOCAL2_TAIL) {
  orc_arr X = V2(px, g), Y = V2(py, g), F = V2(phi, pg);
  orc_mask N = M2(nb, nbg);
  double fx = 0.5 / *dx, fy = 0.5 / *dy;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    A2(X, i, j) = (A2(F, i + 1, j) - A2(F, i - 1, j)) * fx;
    A2(Y, i, j) = (A2(F, i, j + 1) - A2(F, i, j - 1)) * fy;
  }
}
/* lsm2dCentralGradOrder4LOCAL :1465-1537 (its 1/12 literal rounds to the same double) */
void oracle_lsm2dcentralgradorder4local(double *px, double *py, BOXP2(g), const double *phi, BOXP2(pg),
                                        const double *dx, const double *dy, LOCAL2_TAIL) {
  orc_arr X = V2(px, g), Y = V2(py, g), F = V2(phi, pg);
  orc_mask N = M2(nb, nbg);
  double fx = 0.0833333333333333333332 / *dx, fy = 0.0833333333333333333332 / *dy;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    A2(X, i, j) = (-A2(F, i + 2, j) + 8.0 * A2(F, i + 1, j) + A2(F, i - 2, j) - 8.0 * A2(F, i - 1, j)) * fx;
    A2(Y, i, j) = (-A2(F, i, j + 2) + 8.0 * A2(F, i, j + 1) + A2(F, i, j - 2) - 8.0 * A2(F, i, j - 1)) * fy;
  }
}
/* lsm2dLaplacianOrder2LOCAL :1542-1610 */
void oracle_lsm2dlaplacianorder2local(double *lap, BOXP2(lg), const double *phi, BOXP2(pg), const double *dx,
                                      const double *dy, LOCAL2_TAIL) {
  orc_arr L = V2(lap, lg), F = V2(phi, pg);
  orc_mask N = M2(nb, nbg);
  double ax = 1.0 / *dx / *dx, ay = 1.0 / *dy / *dy;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double c = A2(F, i, j);
    A2(L, i, j) = ax * (A2(F, i + 1, j) - 2.0 * c + A2(F, i - 1, j)) + ay * (A2(F, i, j + 1) - 2.0 * c + A2(F, i, j - 1));
  }
}
/* lsm2dComputeAveGradPhiLOCAL :1616-1683: mean of the central |grad phi| over the band */
void oracle_lsm2dcomputeavegradphilocal(double *ave, const double *phi, BOXP2(pg), const double *dx,
                                        const double *dy, LOCAL2_TAIL) {
  orc_arr F = V2(phi, pg);
  orc_mask N = M2(nb, nbg);
  double fx = 0.5 / *dx, fy = 0.5 / *dy, s = 0.0;
  int count = 0;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double x = (A2(F, i + 1, j) - A2(F, i - 1, j)) * fx, y = (A2(F, i, j + 1) - A2(F, i, j - 1)) * fy;
    s = s + sqrt(x * x + y * y);
    count++;
  }
  if (count > 0) s = s / (double)count;
  *ave = s;
}
/* lsm2dGradientMagnitudeLocal :1689-1763 (both cut-offs zero_tol) */
void oracle_lsm2dgradientmagnitudelocal(const double *px, const double *py, double *mag, BOXP2(g), LOCAL2_TAIL) {
  orc_arr X = V2(px, g), Y = V2(py, g), Mg = V2(mag, g);
  orc_mask N = M2(nb, nbg);
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double xs = A2(X, i, j) * A2(X, i, j), ys = A2(Y, i, j) * A2(Y, i, j);
    if (xs < ORC_ZERO_TOL) xs = 0.0;
    if (ys < ORC_ZERO_TOL) ys = 0.0;
    double t = sqrt(xs + ys);
    A2(Mg, i, j) = t < ORC_ZERO_TOL ? 0.0 : t;
  }
}

/* ---------------------------------------------------------------------------
 * RHS terms, src/toolbox/lsm_level_set_evolution2d_local.f.in
 * ------------------------------------------------------------------------- */
/* :16-49 -- every listed point, no mark filter */
void oracle_lsm2dzerooutlevelseteqnrhslocal(double *rhs, BOXP2(rg), const int *ix, const int *iy, const int *nlo,
                                            const int *nhi) {
  orc_arr R = V2(rhs, rg);
  for (int l = *nlo; l <= *nhi; l++) A2(R, LIDX(ix, *nlo, l), LIDX(iy, *nlo, l)) = 0.0;
}
/* :55-120 */
void oracle_lsm2daddadvectiontermtolserhslocal(double *rhs, BOXP2(rg), const double *px, const double *py, BOXP2(g),
                                               const double *vx, const double *vy, BOXP2(vg), LOCAL2_TAIL) {
  orc_arr R = V2(rhs, rg), X = V2(px, g), Y = V2(py, g), U = V2(vx, vg), V = V2(vy, vg);
  orc_mask N = M2(nb, nbg);
  FOR_BAND2(*nlo, *nhi, N, *mark_fb)
  A2(R, i, j) = A2(R, i, j) - (A2(U, i, j) * A2(X, i, j) + A2(V, i, j) * A2(Y, i, j));
}
static double godunov2(int vpos, double xp, double yp, double xm, double ym) {
  if (vpos)
    return orc_max(orc_sq(orc_max(xm, 0.0)), orc_sq(orc_min(xp, 0.0))) +
           orc_max(orc_sq(orc_max(ym, 0.0)), orc_sq(orc_min(yp, 0.0)));
  return orc_max(orc_sq(orc_min(xm, 0.0)), orc_sq(orc_max(xp, 0.0))) +
         orc_max(orc_sq(orc_min(ym, 0.0)), orc_sq(orc_max(yp, 0.0)));
}
/* :126-221 */
void oracle_lsm2daddnormalveltermtolserhslocal(double *rhs, BOXP2(rg), const double *pxp, const double *pyp,
                                               BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm),
                                               const double *vn, BOXP2(vg), LOCAL2_TAIL) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm), VN = V2(vn, vg);
  orc_mask N = M2(nb, nbg);
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double v = A2(VN, i, j);
    double s = godunov2(v > 0.0, A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j));
    if (fabs(v) >= ORC_ZERO_TOL) A2(R, i, j) = A2(R, i, j) - v * sqrt(s);
  }
}
/* :227-314 */
void oracle_lsm2daddconstnormalveltermtolserhslocal(double *rhs, BOXP2(rg), const double *pxp, const double *pyp,
                                                    BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm),
                                                    const double *vn, LOCAL2_TAIL) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  orc_mask N = M2(nb, nbg);
  double v = *vn;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double s = godunov2(v > 0.0, A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j));
    if (fabs(v) >= ORC_ZERO_TOL) A2(R, i, j) = A2(R, i, j) - v * sqrt(s);
  }
}
/* :320-402 */
void oracle_lsm2daddconstcurvtermtolserhslocal(double *rhs, BOXP2(rg), const double *px, const double *py,
                                               BOXP2(g), const double *pxx, const double *pxy, const double *pyy,
                                               BOXP2(g2), const double *b_, LOCAL2_TAIL) {
  orc_arr R = V2(rhs, rg), X = V2(px, g), Y = V2(py, g), XX = V2(pxx, g2), XY = V2(pxy, g2), YY = V2(pyy, g2);
  orc_mask N = M2(nb, nbg);
  double bb = *b_;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double x = A2(X, i, j), y = A2(Y, i, j), g = x * x + y * y, curv;
    if (g < ORC_ZERO_TOL) {
      curv = 0.0;
    } else {
      curv = A2(XX, i, j) * y * y + A2(YY, i, j) * x * x - 2.0 * A2(XY, i, j) * x * y;
      curv = curv / g;
    }
    A2(R, i, j) = A2(R, i, j) + bb * curv;
  }
}
/* :408-471 */
void oracle_lsm2daddconstprecomputedcurvtermtolserhslocal(double *rhs, BOXP2(rg), const double *kappa, BOXP2(kg),
                                                          const double *gmag, BOXP2(gg), const double *b_,
                                                          LOCAL2_TAIL) {
  orc_arr R = V2(rhs, rg), K = V2(kappa, kg), G = V2(gmag, gg);
  orc_mask N = M2(nb, nbg);
  double bb = *b_;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) A2(R, i, j) = A2(R, i, j) + bb * A2(K, i, j) * A2(G, i, j);
}
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
/* :474-615 -- the advection part is added everywhere on the band; the normal
 * part where |V| >= zero_tol and the picked |grad phi|^2 >= zero_tol */
void oracle_lsm2daddexternalandnormalveltermtolserhslocal(double *rhs, BOXP2(rg), const double *pxp,
                                                          const double *pyp, BOXP2(gp), const double *pxm,
                                                          const double *pym, BOXP2(gm), const double *vn,
                                                          const double *vx, const double *vy, BOXP2(vg),
                                                          LOCAL2_TAIL) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  orc_arr VN = V2(vn, vg), U = V2(vx, vg), V = V2(vy, vg);
  orc_mask N = M2(nb, nbg);
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double n = A2(VN, i, j), u = A2(U, i, j), v = A2(V, i, j);
    double px = ext_pick(u, n, A2(XP, i, j), A2(XM, i, j));
    double py = ext_pick(v, n, A2(YP, i, j), A2(YM, i, j));
    double s = px * px + py * py;
    if (fabs(n) >= ORC_ZERO_TOL && s >= ORC_ZERO_TOL)
      A2(R, i, j) = A2(R, i, j) - u * px - v * py - n * sqrt(s);
    else
      A2(R, i, j) = A2(R, i, j) - u * px - v * py;
  }
}

/* ---------------------------------------------------------------------------
 * TVD Runge-Kutta, src/toolbox/lsm_tvd_runge_kutta2d_local.f
 * ------------------------------------------------------------------------- */
/* lsm2dRK1StepLOCAL :33-96 (TVD-RK2/RK3 stage 1, :123-188 / :308-373, call it) */
void oracle_lsm2drk1steplocal(double *un, BOXP2(ng), const double *uc, BOXP2(cg), const double *rhs, BOXP2(rg),
                              const double *dt, LOCAL2_TAIL) {
  orc_arr U = V2(un, ng), C = V2(uc, cg), R = V2(rhs, rg);
  orc_mask N = M2(nb, nbg);
  double t = *dt;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) A2(U, i, j) = A2(C, i, j) + t * A2(R, i, j);
}
void oracle_lsm2dtvdrk2stage1local(double *un, BOXP2(ng), const double *uc, BOXP2(cg), const double *rhs,
                                   BOXP2(rg), const double *dt, LOCAL2_TAIL) {
  oracle_lsm2drk1steplocal(un, ilo_ng, ihi_ng, jlo_ng, jhi_ng, uc, ilo_cg, ihi_cg, jlo_cg, jhi_cg, rhs, ilo_rg,
                           ihi_rg, jlo_rg, jhi_rg, dt, ix, iy, nlo, nhi, nb, ilo_nbg, ihi_nbg, jlo_nbg, jhi_nbg,
                           mark_fb);
}
void oracle_lsm2dtvdrk3stage1local(double *un, BOXP2(ng), const double *uc, BOXP2(cg), const double *rhs,
                                   BOXP2(rg), const double *dt, LOCAL2_TAIL) {
  oracle_lsm2drk1steplocal(un, ilo_ng, ihi_ng, jlo_ng, jhi_ng, uc, ilo_cg, ihi_cg, jlo_cg, jhi_cg, rhs, ilo_rg,
                           ihi_rg, jlo_rg, jhi_rg, dt, ix, iy, nlo, nhi, nb, ilo_nbg, ihi_nbg, jlo_nbg, jhi_nbg,
                           mark_fb);
}
/* lsm2dTVDRK2Stage2LOCAL :212-283 */
void oracle_lsm2dtvdrk2stage2local(double *un, BOXP2(ng), const double *u1, BOXP2(g1), const double *uc, BOXP2(cg),
                                   const double *rhs, BOXP2(rg), const double *dt, LOCAL2_TAIL) {
  orc_arr U = V2(un, ng), S = V2(u1, g1), C = V2(uc, cg), R = V2(rhs, rg);
  orc_mask N = M2(nb, nbg);
  double t = *dt;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) A2(U, i, j) = 0.5 * (A2(C, i, j) + A2(S, i, j) + t * A2(R, i, j));
}
/* lsm2dTVDRK3Stage2LOCAL :396-468 */
void oracle_lsm2dtvdrk3stage2local(double *u2, BOXP2(g2), const double *u1, BOXP2(g1), const double *uc, BOXP2(cg),
                                   const double *rhs, BOXP2(rg), const double *dt, LOCAL2_TAIL) {
  orc_arr U = V2(u2, g2), S = V2(u1, g1), C = V2(uc, cg), R = V2(rhs, rg);
  orc_mask N = M2(nb, nbg);
  double t = *dt;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) A2(U, i, j) = 0.75 * A2(C, i, j) + 0.25 * (A2(S, i, j) + t * A2(R, i, j));
}
/* lsm2dTVDRK3Stage3LOCAL :491-564 */
void oracle_lsm2dtvdrk3stage3local(double *un, BOXP2(ng), const double *u2, BOXP2(g2), const double *uc, BOXP2(cg),
                                   const double *rhs, BOXP2(rg), const double *dt, LOCAL2_TAIL) {
  orc_arr U = V2(un, ng), S = V2(u2, g2), C = V2(uc, cg), R = V2(rhs, rg);
  orc_mask N = M2(nb, nbg);
  double t = *dt;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb)
  A2(U, i, j) = (1.0 / 3.0) * A2(C, i, j) + (2.0 / 3.0) * (A2(S, i, j) + t * A2(R, i, j));
}

/* ---------------------------------------------------------------------------
 * Reductions, src/toolbox/lsm_math_utils2d_local.f.in
 * ------------------------------------------------------------------------- */
/* :16-81 */
void oracle_lsm2dmaxnormdifflocal(double *out, const double *f1, BOXP2(g1), const double *f2, BOXP2(g2),
                                  LOCAL2_TAIL) {
  orc_arr A = V2(f1, g1), B = V2(f2, g2);
  orc_mask N = M2(nb, nbg);
  double m = 0.0;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double d = fabs(A2(A, i, j) - A2(B, i, j));
    if (d > m) m = d;
  }
  *out = m;
}
/* :165-241 */
void oracle_lsm2dcomputestableadvectiondtlocal(double *dt, const double *vx, const double *vy, BOXP2(vg),
                                               const double *dx, const double *dy, const double *cfl,
                                               LOCAL2_TAIL) {
  orc_arr U = V2(vx, vg), V = V2(vy, vg);
  orc_mask N = M2(nb, nbg);
  double m = -1.0, ix_ = 1.0 / *dx, iy_ = 1.0 / *dy;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double c = fabs(A2(U, i, j)) * ix_ + fabs(A2(V, i, j)) * iy_;
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}
static double h2(double absv, double xp, double yp, double xm, double ym, double mdx2, double ix_, double iy_) {
  double x = orc_max(fabs(xp), fabs(xm)), y = orc_max(fabs(yp), fabs(ym));
  return absv / sqrt(x * x + y * y + mdx2) * (x * ix_ + y * iy_);
}
/* :246-354 */
void oracle_lsm2dcomputestablenormalveldtlocal(double *dt, const double *vn, BOXP2(vg), const double *pxp,
                                               const double *pyp, BOXP2(gp), const double *pxm, const double *pym,
                                               BOXP2(gm), const double *dx, const double *dy, const double *cfl,
                                               LOCAL2_TAIL) {
  orc_arr VN = V2(vn, vg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  orc_mask N = M2(nb, nbg);
  double h = orc_max(*dx, *dy), mdx2 = h * h, m = -1.0, ix_ = 1.0 / *dx, iy_ = 1.0 / *dy;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double c = h2(fabs(A2(VN, i, j)), A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j), mdx2, ix_, iy_);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}
/* :360-462 */
void oracle_lsm2dcomputestableconstnormalveldtlocal(double *dt, const double *vn, const double *pxp,
                                                    const double *pyp, BOXP2(gp), const double *pxm,
                                                    const double *pym, BOXP2(gm), const double *dx, const double *dy,
                                                    const double *cfl, LOCAL2_TAIL) {
  orc_arr XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  orc_mask N = M2(nb, nbg);
  double h = orc_max(*dx, *dy), mdx2 = h * h, m = -1.0, ix_ = 1.0 / *dx, iy_ = 1.0 / *dy, av = fabs(*vn);
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double c = h2(av, A2(XP, i, j), A2(YP, i, j), A2(XM, i, j), A2(YM, i, j), mdx2, ix_, iy_);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

/* ---------------------------------------------------------------------------
 * Reinitialization / orthogonalization, src/reinitialization/lsm_reinitialization2d_local.f.in
 * ------------------------------------------------------------------------- */
/* :49-251 -- the local routine tests |sign source| >= zero_tol (the 2D
 * whole-grid routine tests >) */
void oracle_lsm2dcomputereinitializationeqnrhslocal(double *rr, BOXP2(rg), const double *phi, BOXP2(pg),
                                                    const double *phi0, BOXP2(p0g), const double *pxp,
                                                    const double *pyp, BOXP2(gp), const double *pxm,
                                                    const double *pym, BOXP2(gm), const double *dx, const double *dy,
                                                    const int *use_phi0_for_sgn, LOCAL2_TAIL) {
  orc_arr R = V2(rr, rg), F = V2(phi, pg), F0 = V2(phi0, p0g);
  orc_arr XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm);
  orc_mask N = M2(nb, nbg);
  double h = orc_max(*dx, *dy), dx_sq = h * h;
  int use0 = (*use_phi0_for_sgn == 1);
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double s = use0 ? A2(F0, i, j) : A2(F, i, j);
    double gp_[2] = {A2(XP, i, j), A2(YP, i, j)}, gm_[2] = {A2(XM, i, j), A2(YM, i, j)}, star[2];
    for (int d = 0; d < 2; d++) {
      double p, m;
      if (s > 0.0) {
        p = orc_max(-gp_[d], 0.0);
        m = orc_max(gm_[d], 0.0);
      } else {
        p = orc_max(gp_[d], 0.0);
        m = orc_max(-gm_[d], 0.0);
      }
      star[d] = orc_max(p, m);
    }
    if (fabs(s) >= ORC_ZERO_TOL) {
      double n2 = star[0] * star[0] + star[1] * star[1];
      double sgn = use0 ? s / sqrt(s * s + dx_sq) : s / sqrt(s * s + n2 * dx_sq);
      A2(R, i, j) = sgn * (1.0 - sqrt(n2));
    } else {
      A2(R, i, j) = 0.0;
    }
  }
}
/* :284-432 */
void oracle_lsm2dcomputeorthogonalizationeqnrhslocal(double *rhs, BOXP2(rg), const double *pxp, const double *pyp,
                                                     BOXP2(gp), const double *pxm, const double *pym, BOXP2(gm),
                                                     const double *psi, BOXP2(sg), const double *qxp,
                                                     const double *qyp, BOXP2(qp), const double *qxm,
                                                     const double *qym, BOXP2(qm), const double *dx,
                                                     const double *dy, LOCAL2_TAIL) {
  orc_arr R = V2(rhs, rg), XP = V2(pxp, gp), YP = V2(pyp, gp), XM = V2(pxm, gm), YM = V2(pym, gm), S = V2(psi, sg);
  orc_arr QXP = V2(qxp, qp), QYP = V2(qyp, qp), QXM = V2(qxm, qm), QYM = V2(qym, qm);
  orc_mask N = M2(nb, nbg);
  double h = orc_max(*dx, *dy), dx_sq = h * h;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double g1 = 0.5 * (A2(QXP, i, j) + A2(QXM, i, j)), g2 = 0.5 * (A2(QYP, i, j) + A2(QYM, i, j));
    double ng = sqrt(g1 * g1 + g2 * g2), s = A2(S, i, j);
    if (fabs(s) > ORC_ZERO_TOL && ng > ORC_ZERO_TOL) {
      double sgn = s / sqrt(s * s + dx_sq), a1 = sgn * g1, a2 = sgn * g2;
      A2(R, i, j) = -1.0 / ng *
                    (orc_max(a1, 0.0) * A2(XM, i, j) + orc_min(a1, 0.0) * A2(XP, i, j) +
                     orc_max(a2, 0.0) * A2(YM, i, j) + orc_min(a2, 0.0) * A2(YP, i, j));
    } else {
      A2(R, i, j) = 0.0;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Curvature and normals, src/geometry/lsm_curvature2d_local.f.in, lsm_geometry2d_local.f.in
 * ------------------------------------------------------------------------- */
/* curvature2d_local :15-108 (ORDER 2) and :299-402 (ORDER 4, phi_xy = d/dy phi_x) */
static void mean_curv_local2(int order, orc_arr K, orc_arr F, orc_arr X, orc_arr Y, orc_arr G, double dx,
                             double dy, const int *ix, const int *iy, int lo, int hi, orc_mask N,
                             signed char mark) {
  const double c12 = 0.0833333333333333333333;
  double fxx = order == 2 ? 1.0 / dx / dx : c12 / dx / dx, fyy = order == 2 ? 1.0 / dy / dy : c12 / dy / dy;
  double fxy = order == 2 ? 0.25 / dx / dy : c12 / dy;
  FOR_BAND2(lo, hi, N, mark) {
    double c = A2(F, i, j), xx, yy, xy;
    if (order == 2) {
      xx = (A2(F, i + 1, j) - 2.0 * c + A2(F, i - 1, j)) * fxx;
      yy = (A2(F, i, j + 1) - 2.0 * c + A2(F, i, j - 1)) * fyy;
      xy = (A2(F, i + 1, j + 1) - A2(F, i + 1, j - 1) - A2(F, i - 1, j + 1) + A2(F, i - 1, j - 1)) * fxy;
    } else {
      xx = (-A2(F, i + 2, j) + 16.0 * A2(F, i + 1, j) - 30.0 * c - A2(F, i - 2, j) + 16.0 * A2(F, i - 1, j)) * fxx;
      yy = (-A2(F, i, j + 2) + 16.0 * A2(F, i, j + 1) - 30.0 * c - A2(F, i, j - 2) + 16.0 * A2(F, i, j - 1)) * fyy;
      xy = (-A2(X, i, j + 2) + 8.0 * A2(X, i, j + 1) + A2(X, i, j - 2) - 8.0 * A2(X, i, j - 1)) * fxy;
    }
    double g = A2(G, i, j), den = g * g * g;
    if (den < ORC_ZERO_TOL) {
      A2(K, i, j) = 0.0;
    } else {
      double x = A2(X, i, j), y = A2(Y, i, j);
      A2(K, i, j) = (xx * y * y + yy * x * x - 2.0 * xy * x * y) / den;
    }
  }
}
void oracle_lsm2dcomputemeancurvatureorder2local(double *kappa, BOXP2(kg), const double *phi, BOXP2(pg),
                                                 const double *px, const double *py, const double *gmag, BOXP2(gg),
                                                 const double *dx, const double *dy, LOCAL2_TAIL) {
  mean_curv_local2(2, V2(kappa, kg), V2(phi, pg), V2(px, gg), V2(py, gg), V2(gmag, gg), *dx, *dy, ix, iy, *nlo, *nhi,
                   M2(nb, nbg), (signed char)*mark_fb);
}
void oracle_lsm2dcomputemeancurvatureorder4local(double *kappa, BOXP2(kg), const double *phi, BOXP2(pg),
                                                 const double *px, const double *py, const double *gmag, BOXP2(gg),
                                                 const double *dx, const double *dy, LOCAL2_TAIL) {
  mean_curv_local2(4, V2(kappa, kg), V2(phi, pg), V2(px, gg), V2(py, gg), V2(gmag, gg), *dx, *dy, ix, iy, *nlo, *nhi,
                   M2(nb, nbg), (signed char)*mark_fb);
}
/* geometry2d_local :34-113 (dx, dy unused) */
void oracle_lsm2dcomputeunitnormallocal(double *nx, double *ny, BOXP2(ng), const double *px, const double *py,
                                        BOXP2(gg), const double *dx, const double *dy, LOCAL2_TAIL) {
  orc_arr NX = V2(nx, ng), NY = V2(ny, ng), X = V2(px, gg), Y = V2(py, gg);
  orc_mask N = M2(nb, nbg);
  (void)dx;
  (void)dy;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
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
/* geometry2d_local :140-248 */
void oracle_lsm2dcomputesignedunitnormallocal(double *nx, double *ny, BOXP2(ng), const double *px, const double *py,
                                              BOXP2(gg), const double *phi, BOXP2(pg), const double *dx,
                                              const double *dy, LOCAL2_TAIL) {
  orc_arr NX = V2(nx, ng), NY = V2(ny, ng), X = V2(px, gg), Y = V2(py, gg), P = V2(phi, pg);
  orc_mask N = M2(nb, nbg);
  double h = orc_max(*dx, *dy), dx_sq = h * h;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
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

/* ---------------------------------------------------------------------------
 * Band construction, src/toolbox/lsm_localization2d.f
 * ------------------------------------------------------------------------- */
/* :48-133 -- first-claim growth of levels 1..level in parent order and
 * neighbour order +x, -x, +y, -y */
void oracle_lsm2dmarknarrowbandneighbors(unsigned char *nb, BOXP2(g), int *ix, int *iy, const int *nlo,
                                         const int *nhi, int *n_lo, int *n_hi, const int *level) {
  orc_mask N = M2(nb, g);
  int base = *nlo, count = n_hi[0] + 1;
  int lo[2] = {*ilo_g, *jlo_g}, hi[2] = {*ihi_g, *jhi_g};
  (void)nhi;
  for (int L = 1; L <= *level; L++) {
    unsigned char mark = (unsigned char)(L + 1);
    for (int m = n_lo[L - 1]; m <= n_hi[L - 1]; m++) {
      int p[2] = {LIDX(ix, base, m), LIDX(iy, base, m)};
      for (int d = 0; d < 4; d++) {
        int ax = d / 2, up = (d % 2 == 0);
        if (up ? !(p[ax] < hi[ax]) : !(p[ax] > lo[ax])) continue;
        int q0 = p[0] + (ax == 0 ? (up ? 1 : -1) : 0), q1 = p[1] + (ax == 1 ? (up ? 1 : -1) : 0);
        if (AT(N, q0, q1, 0) == 0) {
          LIDX(ix, base, count) = q0;
          LIDX(iy, base, count) = q1;
          AT(N, q0, q1, 0) = mark;
          count++;
        }
      }
    }
    if (count > n_hi[L - 1] + 1) {
      n_lo[L] = n_hi[L - 1] + 1;
      n_hi[L] = count - 1;
    } else {
      n_lo[L] = -1;
      n_hi[L] = -1;
    }
  }
}

/* :139-243 -- level 0 is the (j,i)-ordered set |phi| < width over phi's ghost
 * box; outer points (|phi| >= width_inner) go to index_outer, phi <= 0 from the
 * front and phi > 0 from the back.  Unlike the 3D routine the counts are set
 * unconditionally, and the band grows only when level > 0 and level 0 holds
 * at least two points. */
void oracle_lsm2ddeterminenarrowband(const double *phi, BOXP2(g), unsigned char *nb, BOXP2(nbg), int *ix, int *iy,
                                     const int *nlo, const int *nhi, int *n_lo, int *n_hi, int *index_outer,
                                     const int *nlo_outer, const int *nhi_outer, int *nlo_outer_plus,
                                     int *nhi_outer_plus, int *nlo_outer_minus, int *nhi_outer_minus,
                                     const double *width, const double *width_inner, const int *level) {
  orc_arr F = V2(phi, g);
  orc_mask N = M2(nb, nbg);
  int base = *nlo, obase = *nlo_outer, count = *nlo;
  n_lo[0] = *nlo;
  int cminus = *nlo_outer, cplus = *nhi_outer;
  *nlo_outer_minus = *nlo_outer;
  *nhi_outer_plus = *nhi_outer;
  for (int j = *jlo_g; j <= *jhi_g; j++)
    for (int i = *ilo_g; i <= *ihi_g; i++) {
      double v = A2(F, i, j), a = fabs(v);
      if (a < *width) {
        LIDX(ix, base, count) = i;
        LIDX(iy, base, count) = j;
        AT(N, i, j, 0) = 1;
        if (a >= *width_inner) {
          if (v <= 0.0) { LIDX(index_outer, obase, cminus) = count; cminus++; }
          else { LIDX(index_outer, obase, cplus) = count; cplus--; }
        }
        count++;
      } else {
        AT(N, i, j, 0) = 0;
      }
    }
  n_hi[0] = count - 1;
  *nhi_outer_minus = cminus - 1;
  *nlo_outer_plus = cplus + 1;
  if (*level > 0 && n_hi[0] > n_lo[0])
    oracle_lsm2dmarknarrowbandneighbors(nb, ilo_nbg, ihi_nbg, jlo_nbg, jhi_nbg, ix, iy, nlo, nhi, n_lo, n_hi, level);
}

/* :486-533 -- the four fill-box lines, each spanning the ghost box */
void oracle_lsm2dmarknarrowbandboundarylayer(unsigned char *nb, BOXP2(g), BOXP2(fb), const unsigned char *mark) {
  orc_mask N = M2(nb, g);
  for (int j = *jlo_g; j <= *jhi_g; j++) {
    AT(N, *ilo_fb, j, 0) = *mark;
    AT(N, *ihi_fb, j, 0) = *mark;
  }
  for (int i = *ilo_g; i <= *ihi_g; i++) {
    AT(N, i, *jlo_fb, 0) = *mark;
    AT(N, i, *jhi_fb, 0) = *mark;
  }
}

/* :539-615 -- level 0 is {mask >= 0} (use_mask_sign > 0) or {mask < 0}
 * (strict, unlike the 3D routine's <= 0); the band then always grows. */
void oracle_lsm2ddeterminenarrowbandfrommask(const double *mask, BOXP2(mg), unsigned char *nb, BOXP2(nbg), int *ix,
                                             int *iy, const int *nlo, const int *nhi, int *n_lo, int *n_hi,
                                             const int *level, const int *use_mask_sign) {
  orc_arr Mk = V2(mask, mg);
  orc_mask N = M2(nb, nbg);
  int base = *nlo, count = *nlo;
  n_lo[0] = *nlo;
  for (int j = *jlo_mg; j <= *jhi_mg; j++)
    for (int i = *ilo_mg; i <= *ihi_mg; i++) {
      double v = A2(Mk, i, j);
      int in = (*use_mask_sign > 0) ? (v >= 0) : (v < 0);
      if (in) {
        LIDX(ix, base, count) = i;
        LIDX(iy, base, count) = j;
        AT(N, i, j, 0) = 1;
        count++;
      } else {
        AT(N, i, j, 0) = 0;
      }
    }
  n_hi[0] = count - 1;
  oracle_lsm2dmarknarrowbandneighbors(nb, ilo_nbg, ihi_nbg, jlo_nbg, jhi_nbg, ix, iy, nlo, nhi, n_lo, n_hi, level);
}

/* :621-688 */
void oracle_lsm2dmultiplycutofflserhslocal(const double *phi, double *rhs, BOXP2(g), const int *ix, const int *iy,
                                           const int *nlo, const int *nhi, const unsigned char *nb, BOXP2(nbg),
                                           const unsigned char *mark_fb, const double *beta, const double *gamma) {
  orc_arr F = V2(phi, g), R = V2(rhs, g);
  orc_mask N = M2(nb, nbg);
  double b = *beta, gm = *gamma, c1 = gm - 3.0 * b, c2 = gm - b;
  c2 = c2 * c2 * c2;
  FOR_BAND2(*nlo, *nhi, N, *mark_fb) {
    double a = fabs(A2(F, i, j)), cut;
    if (a <= b) cut = 1.0;
    else if (a <= gm) {
      double t = a - gm;
      cut = (t * t * (2.0 * a + c1)) / c2;
    } else cut = 0.0;
    A2(R, i, j) = cut * A2(R, i, j);
  }
}

/* :693-772 */
void oracle_lsm2dcheckouternarrowbandlayer(int *change_sign, const double *phi, BOXP2(g), const int *ix,
                                           const int *iy, const int *nlo, const int *nhi, const int *index_outer,
                                           const int *nlo_outer, const int *nhi_outer, const int *nlo_outer_plus,
                                           const int *nhi_outer_plus, const int *nlo_outer_minus,
                                           const int *nhi_outer_minus) {
  orc_arr F = V2(phi, g);
  int base = *nlo, obase = *nlo_outer;
  (void)nhi;
  (void)nhi_outer;
  *change_sign = 0;
  for (int pass = 0; pass < 2 && *change_sign == 0; pass++) {
    int lo = pass == 0 ? *nlo_outer_plus : *nlo_outer_minus, hi = pass == 0 ? *nhi_outer_plus : *nhi_outer_minus;
    int sp = 0, sm = 0;
    for (int m = lo; m <= hi; m++) {
      int l = LIDX(index_outer, obase, m);
      if (A2(F, LIDX(ix, base, l), LIDX(iy, base, l)) > 0.0) sp = 1;
      else sm = 1;
      if (sp == 1 && sm == 1) {
        *change_sign = 1;
        break;
      }
    }
  }
}

/* :778-827 */
void oracle_lsm2dimposemasklocal(double *dest, BOXP2(dg), const double *src, BOXP2(sg), const double *mask,
                                 BOXP2(mg), const int *ix, const int *iy, const int *nlo, const int *nhi) {
  orc_arr D = V2(dest, dg), S = V2(src, sg), Mk = V2(mask, mg);
  for (int l = *nlo; l <= *nhi; l++) {
    int i = LIDX(ix, *nlo, l), j = LIDX(iy, *nlo, l);
    A2(D, i, j) = orc_max(A2(Mk, i, j), A2(S, i, j));
  }
}

/* :833-875 */
void oracle_lsm2dcopydatalocal(double *dest, BOXP2(dg), const double *src, BOXP2(sg), const int *ix, const int *iy,
                               const int *nlo, const int *nhi) {
  orc_arr D = V2(dest, dg), S = V2(src, sg);
  for (int l = *nlo; l <= *nhi; l++) {
    int i = LIDX(ix, *nlo, l), j = LIDX(iy, *nlo, l);
    A2(D, i, j) = A2(S, i, j);
  }
}
