/*
 * oracle_local.c -- TEST INFRASTRUCTURE ONLY (see lsm_oracle.h).
 *
 * Narrow-band ("local") kernels and band construction.  Index lists are
 * addressed as the reference's Fortran declarations address them: an array
 * declared index_x(nlo:nhi) is passed as a pointer to its element nlo.  Marks
 * are Fortran integer*1 values compared as signed bytes.
 */
#include "lsm_oracle.h"

/* lsm3dComputeDnLOCAL, src/toolbox/lsm_spatial_derivatives3d_local.f.in:43-138 */
static void computeDnLocal(orc_arr Dn, orc_arr Dm1, int n, int dir, const int *ix, const int *iy,
                           const int *iz, int lo, int hi, int base, orc_mask nb, signed char mark) {
  int off[3] = {0, 0, 0};
  double sgn;
  if (n % 2 == 1) { off[dir] = 1; sgn = 1.0; }
  else { off[dir] = -1; sgn = -1.0; }
  for (int l = lo; l <= hi; l++) {
    int i = LIDX(ix, base, l), j = LIDX(iy, base, l), k = LIDX(iz, base, l);
    if (MARK(nb, i, j, k) <= mark)
      AT(Dn, i, j, k) = sgn * (AT(Dm1, i, j, k) - AT(Dm1, i - off[0], j - off[1], k - off[2]));
    else
      AT(Dn, i, j, k) = ORC_BIG;
  }
}

void oracle_lsm3dcomputednlocal(double *Dn, BOXP(dg), const double *Dm1, BOXP(mg), const int *n,
                                const int *dir, const int *ix, const int *iy, const int *iz, const int *nlo,
                                const int *nhi, const unsigned char *nb, BOXP(nbg), const unsigned char *mark) {
  computeDnLocal(VIEW(Dn, dg), VIEW(Dm1, mg), *n, *dir - 1, ix, iy, iz, *nlo, *nhi, *nlo, MVIEW(nb, nbg),
                 (signed char)*mark);
}

#define NB(a, dir, i, j, k, s) \
  AT(a, (i) + ((dir) == 0 ? (s) : 0), (j) + ((dir) == 1 ? (s) : 0), (k) + ((dir) == 2 ? (s) : 0))

/* lsm3dHJENO1LOCAL, src/toolbox/lsm_spatial_derivatives3d_local.f.in:170-381 */
void oracle_lsm3dhjeno1local(double *pxp, double *pyp, double *pzp, BOXP(gp), double *pxm, double *pym,
                             double *pzm, BOXP(gm), const double *phi, BOXP(pg), double *D1, BOXP(d1),
                             const double *dx, const double *dy, const double *dz, const int *ix,
                             const int *iy, const int *iz, const int *nlo0, const int *nhi0, const int *nlo1,
                             const int *nhi1, const unsigned char *nb, BOXP(nbg), const unsigned char *mark_fb,
                             const unsigned char *mark_D1) {
  double inv[3] = {1.0 / *dx, 1.0 / *dy, 1.0 / *dz};
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  orc_arr F = VIEW(phi, pg), D = VIEW(D1, d1);
  orc_mask N = MVIEW(nb, nbg);
  int base = *nlo0;
  (void)nlo1;
  for (int d = 0; d < 3; d++) {
    computeDnLocal(D, F, 1, d, ix, iy, iz, *nlo0, *nhi1, base, N, (signed char)*mark_D1);
    for (int l = *nlo0; l <= *nhi0; l++) {
      int i = LIDX(ix, base, l), j = LIDX(iy, base, l), k = LIDX(iz, base, l);
      if (MARK(N, i, j, k) <= (signed char)*mark_fb) {
        AT(P[d], i, j, k) = NB(D, d, i, j, k, 1) * inv[d];
        AT(M[d], i, j, k) = NB(D, d, i, j, k, 0) * inv[d];
      }
    }
  }
}

/* lsm3dHJENO2LOCAL, src/toolbox/lsm_spatial_derivatives3d_local.f.in:414-740 */
void oracle_lsm3dhjeno2local(double *pxp, double *pyp, double *pzp, BOXP(gp), double *pxm, double *pym,
                             double *pzm, BOXP(gm), const double *phi, BOXP(pg), double *D1, BOXP(d1),
                             double *D2, BOXP(d2), const double *dx, const double *dy, const double *dz,
                             const int *ix, const int *iy, const int *iz, const int *nlo0, const int *nhi0,
                             const int *nlo1, const int *nhi1, const int *nlo2, const int *nhi2,
                             const unsigned char *nb, BOXP(nbg), const unsigned char *mark_fb,
                             const unsigned char *mark_D1, const unsigned char *mark_D2) {
  double inv[3] = {1.0 / *dx, 1.0 / *dy, 1.0 / *dz};
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  orc_arr F = VIEW(phi, pg), A = VIEW(D1, d1), B = VIEW(D2, d2);
  orc_mask N = MVIEW(nb, nbg);
  int base = *nlo0;
  (void)nlo1;
  (void)nlo2;
  for (int d = 0; d < 3; d++) {
    computeDnLocal(A, F, 1, d, ix, iy, iz, *nlo0, *nhi2, base, N, (signed char)*mark_D1);
    computeDnLocal(B, A, 2, d, ix, iy, iz, *nlo0, *nhi1, base, N, (signed char)*mark_D2);
    for (int l = *nlo0; l <= *nhi0; l++) {
      int i = LIDX(ix, base, l), j = LIDX(iy, base, l), k = LIDX(iz, base, l);
      if (MARK(N, i, j, k) <= (signed char)*mark_fb) {
        double p, m;
        if (fabs(NB(B, d, i, j, k, 0)) < fabs(NB(B, d, i, j, k, 1))) p = NB(A, d, i, j, k, 1) - 0.5 * NB(B, d, i, j, k, 0);
        else p = NB(A, d, i, j, k, 1) - 0.5 * NB(B, d, i, j, k, 1);
        if (fabs(NB(B, d, i, j, k, -1)) < fabs(NB(B, d, i, j, k, 0))) m = NB(A, d, i, j, k, 0) + 0.5 * NB(B, d, i, j, k, -1);
        else m = NB(A, d, i, j, k, 0) + 0.5 * NB(B, d, i, j, k, 0);
        AT(P[d], i, j, k) = p * inv[d];
        AT(M[d], i, j, k) = m * inv[d];
      }
    }
  }
}

/* HJ-WENO5 reconstruction, operation order of
 * src/toolbox/lsm_spatial_derivatives3d.f.in:1296-1323 (same as the whole-grid oracle). */
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

/* 3D HJ-WENO5 on a narrow band.  The reference has no 3D routine; this one
 * follows the 2D pattern lsm2dHJWENO5LOCAL
 * (src/toolbox/lsm_spatial_derivatives2d_local.f.in:816-1154): D1 over index
 * levels 0..3 filtered by mark_D1, the WENO5 formulas of the 3D whole-grid
 * routine on level 0 filtered by mark_fb.  It uses the 3D templated
 * tiny_nonzero_number (1e-36) rather than the 2D file's literal 1e-35. */
void oracle_lsm3dhjweno5local(double *pxp, double *pyp, double *pzp, BOXP(gp), double *pxm, double *pym,
                              double *pzm, BOXP(gm), const double *phi, BOXP(pg), double *D1, BOXP(d1),
                              const double *dx, const double *dy, const double *dz, const int *ix,
                              const int *iy, const int *iz, const int *nlo0, const int *nhi0, const int *nlo1,
                              const int *nhi1, const int *nlo2, const int *nhi2, const int *nlo3,
                              const int *nhi3, const unsigned char *nb, BOXP(nbg), const unsigned char *mark_fb,
                              const unsigned char *mark_D1) {
  double inv[3] = {1.0 / *dx, 1.0 / *dy, 1.0 / *dz};
  orc_arr P[3] = {VIEW(pxp, gp), VIEW(pyp, gp), VIEW(pzp, gp)};
  orc_arr M[3] = {VIEW(pxm, gm), VIEW(pym, gm), VIEW(pzm, gm)};
  orc_arr F = VIEW(phi, pg), A = VIEW(D1, d1);
  orc_mask N = MVIEW(nb, nbg);
  int base = *nlo0;
  (void)nlo1; (void)nhi1; (void)nlo2; (void)nhi2; (void)nlo3;
  for (int d = 0; d < 3; d++) {
    computeDnLocal(A, F, 1, d, ix, iy, iz, *nlo0, *nhi3, base, N, (signed char)*mark_D1);
    for (int l = *nlo0; l <= *nhi0; l++) {
      int i = LIDX(ix, base, l), j = LIDX(iy, base, l), k = LIDX(iz, base, l);
      if (MARK(N, i, j, k) <= (signed char)*mark_fb) {
        double s = inv[d];
        AT(P[d], i, j, k) = weno5(NB(A, d, i, j, k, 3) * s, NB(A, d, i, j, k, 2) * s, NB(A, d, i, j, k, 1) * s,
                                  NB(A, d, i, j, k, 0) * s, NB(A, d, i, j, k, -1) * s);
        AT(M[d], i, j, k) = weno5(NB(A, d, i, j, k, -2) * s, NB(A, d, i, j, k, -1) * s, NB(A, d, i, j, k, 0) * s,
                                  NB(A, d, i, j, k, 1) * s, NB(A, d, i, j, k, 2) * s);
      }
    }
  }
}

/* lsm3dCentralGradOrder2LOCAL, :763-846 */
void oracle_lsm3dcentralgradorder2local(double *px, double *py, double *pz, BOXP(g), const double *phi,
                                        BOXP(pg), const double *dx, const double *dy, const double *dz,
                                        const int *ix, const int *iy, const int *iz, const int *nlo,
                                        const int *nhi, const unsigned char *nb, BOXP(nbg),
                                        const unsigned char *mark_fb) {
  orc_arr X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g), F = VIEW(phi, pg);
  orc_mask N = MVIEW(nb, nbg);
  double fx = 0.5 / *dx, fy = 0.5 / *dy, fz = 0.5 / *dz;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    AT(X, i, j, k) = (AT(F, i + 1, j, k) - AT(F, i - 1, j, k)) * fx;
    AT(Y, i, j, k) = (AT(F, i, j + 1, k) - AT(F, i, j - 1, k)) * fy;
    AT(Z, i, j, k) = (AT(F, i, j, k + 1) - AT(F, i, j, k - 1)) * fz;
  }
}

/* lsm3dCentralGradOrder4LOCAL, :868-964 */
void oracle_lsm3dcentralgradorder4local(double *px, double *py, double *pz, BOXP(g), const double *phi,
                                        BOXP(pg), const double *dx, const double *dy, const double *dz,
                                        const int *ix, const int *iy, const int *iz, const int *nlo,
                                        const int *nhi, const unsigned char *nb, BOXP(nbg),
                                        const unsigned char *mark_fb) {
  orc_arr X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g), F = VIEW(phi, pg);
  orc_mask N = MVIEW(nb, nbg);
  double fx = 0.0833333333333333333333 / *dx, fy = 0.0833333333333333333333 / *dy;
  double fz = 0.0833333333333333333333 / *dz;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    AT(X, i, j, k) = (-AT(F, i + 2, j, k) + 8.0 * AT(F, i + 1, j, k) + AT(F, i - 2, j, k) -
                      8.0 * AT(F, i - 1, j, k)) * fx;
    AT(Y, i, j, k) = (-AT(F, i, j + 2, k) + 8.0 * AT(F, i, j + 1, k) + AT(F, i, j - 2, k) -
                      8.0 * AT(F, i, j - 1, k)) * fy;
    AT(Z, i, j, k) = (-AT(F, i, j, k + 2) + 8.0 * AT(F, i, j, k + 1) + AT(F, i, j, k - 2) -
                      8.0 * AT(F, i, j, k - 1)) * fz;
  }
}

/* lsm3dLaplacianOrder2LOCAL, :982-1071 */
void oracle_lsm3dlaplacianorder2local(double *lap, BOXP(g), const double *phi, BOXP(pg), const double *dx,
                                      const double *dy, const double *dz, const int *ix, const int *iy,
                                      const int *iz, const int *nlo, const int *nhi, const unsigned char *nb,
                                      BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr L = VIEW(lap, g), F = VIEW(phi, pg);
  orc_mask N = MVIEW(nb, nbg);
  double ax = 1.0 / *dx / *dx, ay = 1.0 / *dy / *dy, az = 1.0 / *dz / *dz;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double c = AT(F, i, j, k);
    AT(L, i, j, k) = ax * (AT(F, i + 1, j, k) - 2.0 * c + AT(F, i - 1, j, k)) +
                     ay * (AT(F, i, j + 1, k) - 2.0 * c + AT(F, i, j - 1, k)) +
                     az * (AT(F, i, j, k + 1) - 2.0 * c + AT(F, i, j, k - 1));
  }
}

/* lsm3dComputeAveGradPhiLocal, :1089-1167 */
void oracle_lsm3dcomputeavegradphilocal(double *ave, const double *phi, BOXP(pg), const double *dx,
                                        const double *dy, const double *dz, const int *ix, const int *iy,
                                        const int *iz, const int *nlo, const int *nhi, const unsigned char *nb,
                                        BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr F = VIEW(phi, pg);
  orc_mask N = MVIEW(nb, nbg);
  double fx = 0.5 / *dx, fy = 0.5 / *dy, fz = 0.5 / *dz, s = 0.0;
  int count = 0;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double x = (AT(F, i + 1, j, k) - AT(F, i - 1, j, k)) * fx;
    double y = (AT(F, i, j + 1, k) - AT(F, i, j - 1, k)) * fy;
    double z = (AT(F, i, j, k + 1) - AT(F, i, j, k - 1)) * fz;
    s = s + sqrt(x * x + y * y + z * z);
    count = count + 1;
  }
  if (count > 0) s = s / (double)count;
  *ave = s;
}

/* The addends of that mean's sum, one per admitted band point in list order
 * (*count of them). */
void oracle_avegradphilocal_terms(double *terms, int *count, const double *phi, BOXP(pg), const double *dx,
                                  const double *dy, const double *dz, const int *ix, const int *iy,
                                  const int *iz, const int *nlo, const int *nhi, const unsigned char *nb,
                                  BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr F = VIEW(phi, pg);
  orc_mask N = MVIEW(nb, nbg);
  double fx = 0.5 / *dx, fy = 0.5 / *dy, fz = 0.5 / *dz;
  int c = 0;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double x = (AT(F, i + 1, j, k) - AT(F, i - 1, j, k)) * fx;
    double y = (AT(F, i, j + 1, k) - AT(F, i, j - 1, k)) * fy;
    double z = (AT(F, i, j, k + 1) - AT(F, i, j, k - 1)) * fz;
    terms[c++] = sqrt(x * x + y * y + z * z);
  }
  *count = c;
}

/* lsm3dGradientMagnitudeLocal, :1186-1282 (uses zero_tol, unlike the whole-grid 1e-8) */
void oracle_lsm3dgradientmagnitudelocal(const double *px, const double *py, const double *pz, double *mag,
                                        BOXP(g), const int *ix, const int *iy, const int *iz, const int *nlo,
                                        const int *nhi, const unsigned char *nb, BOXP(nbg),
                                        const unsigned char *mark_fb) {
  orc_arr X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g), G = VIEW(mag, g);
  orc_mask N = MVIEW(nb, nbg);
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double sx = AT(X, i, j, k) * AT(X, i, j, k), sy = AT(Y, i, j, k) * AT(Y, i, j, k);
    double sz = AT(Z, i, j, k) * AT(Z, i, j, k);
    if (sx < ORC_ZERO_TOL) sx = 0.0;
    if (sy < ORC_ZERO_TOL) sy = 0.0;
    if (sz < ORC_ZERO_TOL) sz = 0.0;
    double t = sqrt(sx + sy + sz);
    AT(G, i, j, k) = t < ORC_ZERO_TOL ? 0.0 : t;
  }
}

/* ---------------------------------------------------------------------------
 * Local RHS terms, src/toolbox/lsm_level_set_evolution3d_local.f.in
 * ------------------------------------------------------------------------- */
/* :16-54 (no mark filter) */
void oracle_lsm3dzerooutlevelseteqnrhslocal(double *rhs, BOXP(g), const int *ix, const int *iy, const int *iz,
                                            const int *nlo, const int *nhi) {
  orc_arr R = VIEW(rhs, g);
  for (int l = *nlo; l <= *nhi; l++) AT(R, LIDX(ix, *nlo, l), LIDX(iy, *nlo, l), LIDX(iz, *nlo, l)) = 0.0;
}

/* :60-146 */
void oracle_lsm3daddadvectiontermtolserhslocal(double *rhs, BOXP(rg), const double *px, const double *py,
                                               const double *pz, BOXP(g), const double *vx, const double *vy,
                                               const double *vz, BOXP(vg), const int *ix, const int *iy,
                                               const int *iz, const int *nlo, const int *nhi,
                                               const unsigned char *nb, BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr R = VIEW(rhs, rg), X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g);
  orc_arr U = VIEW(vx, vg), V = VIEW(vy, vg), W = VIEW(vz, vg);
  orc_mask N = MVIEW(nb, nbg);
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    AT(R, i, j, k) = AT(R, i, j, k) - (AT(U, i, j, k) * AT(X, i, j, k) + AT(V, i, j, k) * AT(Y, i, j, k) +
                                       AT(W, i, j, k) * AT(Z, i, j, k));
  }
}

static double godunov_sq(double vn, double xp, double yp, double zp, double xm, double ym, double zm) {
  if (vn > 0.0)
    return orc_max(orc_sq(orc_max(xm, 0.0)), orc_sq(orc_min(xp, 0.0))) +
           orc_max(orc_sq(orc_max(ym, 0.0)), orc_sq(orc_min(yp, 0.0))) +
           orc_max(orc_sq(orc_max(zm, 0.0)), orc_sq(orc_min(zp, 0.0)));
  return orc_max(orc_sq(orc_min(xm, 0.0)), orc_sq(orc_max(xp, 0.0))) +
         orc_max(orc_sq(orc_min(ym, 0.0)), orc_sq(orc_max(yp, 0.0))) +
         orc_max(orc_sq(orc_min(zm, 0.0)), orc_sq(orc_max(zp, 0.0)));
}

/* :152-276 */
void oracle_lsm3daddnormalveltermtolserhslocal(double *rhs, BOXP(rg), const double *pxp, const double *pyp,
                                               const double *pzp, BOXP(gp), const double *pxm,
                                               const double *pym, const double *pzm, BOXP(gm),
                                               const double *vn, BOXP(vg), const int *ix, const int *iy,
                                               const int *iz, const int *nlo, const int *nhi,
                                               const unsigned char *nb, BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr R = VIEW(rhs, rg), VN = VIEW(vn, vg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  orc_mask N = MVIEW(nb, nbg);
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double v = AT(VN, i, j, k);
    double s = godunov_sq(v, AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k), AT(XM, i, j, k),
                          AT(YM, i, j, k), AT(ZM, i, j, k));
    if (fabs(v) >= ORC_ZERO_TOL) AT(R, i, j, k) = AT(R, i, j, k) - v * sqrt(s);
  }
}

/* :282-398 */
void oracle_lsm3daddconstnormalveltermtolserhslocal(double *rhs, BOXP(rg), const double *pxp,
                                                    const double *pyp, const double *pzp, BOXP(gp),
                                                    const double *pxm, const double *pym, const double *pzm,
                                                    BOXP(gm), const double *vel_n, const int *ix, const int *iy,
                                                    const int *iz, const int *nlo, const int *nhi,
                                                    const unsigned char *nb, BOXP(nbg),
                                                    const unsigned char *mark_fb) {
  orc_arr R = VIEW(rhs, rg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  orc_mask N = MVIEW(nb, nbg);
  double v = *vel_n;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double s = godunov_sq(v, AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k), AT(XM, i, j, k),
                          AT(YM, i, j, k), AT(ZM, i, j, k));
    if (fabs(v) >= ORC_ZERO_TOL) AT(R, i, j, k) = AT(R, i, j, k) - v * sqrt(s);
  }
}

static double curv_num(double xx, double xy, double xz, double yy, double yz, double zz, double x, double y,
                       double z) {
  return xx * y * y + yy * x * x - 2.0 * xy * x * y + xx * z * z + zz * x * x - 2.0 * xz * x * z +
         yy * z * z + zz * y * y - 2.0 * yz * y * z;
}

/* :404-524 */
void oracle_lsm3daddconstcurvtermtolserhslocal(double *rhs, BOXP(rg), const double *px, const double *py,
                                               const double *pz, BOXP(g), const double *pxx, const double *pxy,
                                               const double *pxz, const double *pyy, const double *pyz,
                                               const double *pzz, BOXP(g2), const double *b_, const int *ix,
                                               const int *iy, const int *iz, const int *nlo, const int *nhi,
                                               const unsigned char *nb, BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr R = VIEW(rhs, rg), X = VIEW(px, g), Y = VIEW(py, g), Z = VIEW(pz, g);
  orc_arr XX = VIEW(pxx, g2), XY = VIEW(pxy, g2), XZ = VIEW(pxz, g2), YY = VIEW(pyy, g2), YZ = VIEW(pyz, g2),
          ZZ = VIEW(pzz, g2);
  orc_mask N = MVIEW(nb, nbg);
  double bb = *b_;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
    double g2v = x * x + y * y + z * z, curv;
    if (g2v < ORC_ZERO_TOL) curv = 0.0;
    else
      curv = curv_num(AT(XX, i, j, k), AT(XY, i, j, k), AT(XZ, i, j, k), AT(YY, i, j, k), AT(YZ, i, j, k),
                      AT(ZZ, i, j, k), x, y, z) / g2v;
    AT(R, i, j, k) = AT(R, i, j, k) + bb * curv;
  }
}

/* :530-606 */
void oracle_lsm3daddconstprecomputedcurvtermtolserhslocal(double *rhs, BOXP(rg), const double *kappa,
                                                          BOXP(kg), const double *gmag, BOXP(gg),
                                                          const double *b_, const int *ix, const int *iy,
                                                          const int *iz, const int *nlo, const int *nhi,
                                                          const unsigned char *nb, BOXP(nbg),
                                                          const unsigned char *mark_fb) {
  orc_arr R = VIEW(rhs, rg), K = VIEW(kappa, kg), G = VIEW(gmag, gg);
  orc_mask N = MVIEW(nb, nbg);
  double bb = *b_;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) AT(R, i, j, k) = AT(R, i, j, k) + bb * AT(K, i, j, k) * AT(G, i, j, k);
}

static double ext_normal_pick(double u, double vn, double pp, double pm) {
  double hp = u + vn * pp, hm = u + vn * pm;
  if (hp > 0.0) {
    if (hm > 0.0) return pm;
    return -(u / vn);
  }
  if (hm < 0.0) return pp;
  if (fabs(hp) > fabs(hm)) return pp;
  return pm;
}

/* :612-811 */
void oracle_lsm3daddexternalandnormalveltermtolserhslocal(double *rhs, BOXP(rg), const double *pxp,
                                                          const double *pyp, const double *pzp, BOXP(gp),
                                                          const double *pxm, const double *pym,
                                                          const double *pzm, BOXP(gm), const double *vn,
                                                          const double *vx, const double *vy, const double *vz,
                                                          BOXP(vg), const int *ix, const int *iy, const int *iz,
                                                          const int *nlo, const int *nhi,
                                                          const unsigned char *nb, BOXP(nbg),
                                                          const unsigned char *mark_fb) {
  orc_arr R = VIEW(rhs, rg), VN = VIEW(vn, vg), U = VIEW(vx, vg), V = VIEW(vy, vg), W = VIEW(vz, vg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  orc_mask N = MVIEW(nb, nbg);
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double n = AT(VN, i, j, k), u = AT(U, i, j, k), v = AT(V, i, j, k), w = AT(W, i, j, k);
    double px = ext_normal_pick(u, n, AT(XP, i, j, k), AT(XM, i, j, k));
    double py = ext_normal_pick(v, n, AT(YP, i, j, k), AT(YM, i, j, k));
    double pz = ext_normal_pick(w, n, AT(ZP, i, j, k), AT(ZM, i, j, k));
    double s = px * px + py * py + pz * pz;
    if (fabs(n) >= ORC_ZERO_TOL && s >= ORC_ZERO_TOL)
      AT(R, i, j, k) = AT(R, i, j, k) - u * px - v * py - w * pz - n * sqrt(s);
    else
      AT(R, i, j, k) = AT(R, i, j, k) - u * px - v * py - w * pz;
  }
}

/* ---------------------------------------------------------------------------
 * Local TVD Runge-Kutta, src/toolbox/lsm_tvd_runge_kutta3d_local.f:34-665
 * ------------------------------------------------------------------------- */
#define RK_LOCAL_ARGS                                                                              \
  const int *ix, const int *iy, const int *iz, const int *nlo, const int *nhi, const unsigned char *nb, \
      BOXP(nbg), const unsigned char *mark_fb

void oracle_lsm3drk1steplocal(double *un, BOXP(ng), const double *uc, BOXP(cg), const double *rhs, BOXP(rg),
                              const double *dt, RK_LOCAL_ARGS) {
  orc_arr N_ = VIEW(un, ng), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  orc_mask N = MVIEW(nb, nbg);
  double t = *dt;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) AT(N_, i, j, k) = AT(C, i, j, k) + t * AT(R, i, j, k);
}
void oracle_lsm3dtvdrk2stage1local(double *un, BOXP(ng), const double *uc, BOXP(cg), const double *rhs,
                                   BOXP(rg), const double *dt, RK_LOCAL_ARGS) {
  orc_arr N_ = VIEW(un, ng), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  orc_mask N = MVIEW(nb, nbg);
  double t = *dt;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) AT(N_, i, j, k) = AT(C, i, j, k) + t * AT(R, i, j, k);
}
void oracle_lsm3dtvdrk3stage1local(double *un, BOXP(ng), const double *uc, BOXP(cg), const double *rhs,
                                   BOXP(rg), const double *dt, RK_LOCAL_ARGS) {
  orc_arr N_ = VIEW(un, ng), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  orc_mask N = MVIEW(nb, nbg);
  double t = *dt;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) AT(N_, i, j, k) = AT(C, i, j, k) + t * AT(R, i, j, k);
}
void oracle_lsm3dtvdrk2stage2local(double *un, BOXP(ng), const double *u1, BOXP(g1), const double *uc,
                                   BOXP(cg), const double *rhs, BOXP(rg), const double *dt, RK_LOCAL_ARGS) {
  orc_arr N_ = VIEW(un, ng), S1 = VIEW(u1, g1), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  orc_mask N = MVIEW(nb, nbg);
  double t = *dt;
  FOR_BAND(*nlo, *nhi, N, *mark_fb)
  AT(N_, i, j, k) = 0.5 * (AT(C, i, j, k) + AT(S1, i, j, k) + t * AT(R, i, j, k));
}
void oracle_lsm3dtvdrk3stage2local(double *u2, BOXP(g2), const double *u1, BOXP(g1), const double *uc,
                                   BOXP(cg), const double *rhs, BOXP(rg), const double *dt, RK_LOCAL_ARGS) {
  orc_arr S2 = VIEW(u2, g2), S1 = VIEW(u1, g1), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  orc_mask N = MVIEW(nb, nbg);
  double t = *dt;
  FOR_BAND(*nlo, *nhi, N, *mark_fb)
  AT(S2, i, j, k) = 0.75 * AT(C, i, j, k) + 0.25 * (AT(S1, i, j, k) + t * AT(R, i, j, k));
}
void oracle_lsm3dtvdrk3stage3local(double *un, BOXP(ng), const double *u2, BOXP(g2), const double *uc,
                                   BOXP(cg), const double *rhs, BOXP(rg), const double *dt, RK_LOCAL_ARGS) {
  orc_arr N_ = VIEW(un, ng), S2 = VIEW(u2, g2), C = VIEW(uc, cg), R = VIEW(rhs, rg);
  orc_mask N = MVIEW(nb, nbg);
  double t = *dt;
  const double one_third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
  FOR_BAND(*nlo, *nhi, N, *mark_fb)
  AT(N_, i, j, k) = one_third * AT(C, i, j, k) + two_thirds * (AT(S2, i, j, k) + t * AT(R, i, j, k));
}

/* ---------------------------------------------------------------------------
 * Local reductions, src/toolbox/lsm_math_utils3d_local.f.in
 * ------------------------------------------------------------------------- */
/* :16-92 (starts from 0, unlike the whole-grid routine) */
void oracle_lsm3dmaxnormdifflocal(double *out, const double *f1, BOXP(g1), const double *f2, BOXP(g2),
                                  const int *ix, const int *iy, const int *iz, const int *nlo, const int *nhi,
                                  const unsigned char *nb, BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr A = VIEW(f1, g1), B = VIEW(f2, g2);
  orc_mask N = MVIEW(nb, nbg);
  double m = 0.0;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double d = fabs(AT(A, i, j, k) - AT(B, i, j, k));
    if (d > m) m = d;
  }
  *out = m;
}

/* :194-284 */
void oracle_lsm3dcomputestableadvectiondtlocal(double *dt, const double *vx, const double *vy, const double *vz,
                                               BOXP(vg), const double *dx, const double *dy, const double *dz,
                                               const double *cfl, const int *ix, const int *iy, const int *iz,
                                               const int *nlo, const int *nhi, const unsigned char *nb,
                                               BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr U = VIEW(vx, vg), V = VIEW(vy, vg), W = VIEW(vz, vg);
  orc_mask N = MVIEW(nb, nbg);
  double m = -1.0, ax = 1.0 / *dx, ay = 1.0 / *dy, az = 1.0 / *dz;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double c = fabs(AT(U, i, j, k)) * ax + fabs(AT(V, i, j, k)) * ay + fabs(AT(W, i, j, k)) * az;
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

/* :290-426 (skips points whose |grad phi| is not above zero_tol) */
void oracle_lsm3dcomputestablenormalveldtlocal(double *dt, const double *vn, BOXP(vg), const double *pxp,
                                               const double *pyp, const double *pzp, BOXP(gp),
                                               const double *pxm, const double *pym, const double *pzm,
                                               BOXP(gm), const double *dx, const double *dy, const double *dz,
                                               const double *cfl, const int *ix, const int *iy, const int *iz,
                                               const int *nlo, const int *nhi, const unsigned char *nb,
                                               BOXP(nbg), const unsigned char *mark_fb) {
  orc_arr VN = VIEW(vn, vg);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  orc_mask N = MVIEW(nb, nbg);
  double h = orc_max(orc_max(*dx, *dy), *dz), mdx2 = h * h;
  double m = -1.0, ax = 1.0 / *dx, ay = 1.0 / *dy, az = 1.0 / *dz;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double x = orc_max(fabs(AT(XP, i, j, k)), fabs(AT(XM, i, j, k)));
    double y = orc_max(fabs(AT(YP, i, j, k)), fabs(AT(YM, i, j, k)));
    double z = orc_max(fabs(AT(ZP, i, j, k)), fabs(AT(ZM, i, j, k)));
    double n = sqrt(x * x + y * y + z * z + mdx2);
    if (n > ORC_ZERO_TOL) {
      double c = fabs(AT(VN, i, j, k)) / n * (x * ax + y * ay + z * az);
      if (c > m) m = c;
    }
  }
  *dt = *cfl / (m + ORC_TINY);
}

/* :628-751.  The reference divides |vel_n| by a |grad phi| it never assigns
 * (abs_vel_n and norm_grad_phi are uninitialised locals there); this restates
 * the evident intent, the formula of the whole-grid routine
 * (src/toolbox/lsm_math_utils3d.f.in:519-564). */
void oracle_lsm3dcomputestableconstnormalveldtlocal(double *dt, const double *vel_n, const double *pxp,
                                                    const double *pyp, const double *pzp, BOXP(gp),
                                                    const double *pxm, const double *pym, const double *pzm,
                                                    BOXP(gm), const double *dx, const double *dy,
                                                    const double *dz, const double *cfl, const int *ix,
                                                    const int *iy, const int *iz, const int *nlo,
                                                    const int *nhi, const unsigned char *nb, BOXP(nbg),
                                                    const unsigned char *mark_fb) {
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  orc_mask N = MVIEW(nb, nbg);
  double h = orc_max(orc_max(*dx, *dy), *dz), mdx2 = h * h;
  double m = -1.0, ax = 1.0 / *dx, ay = 1.0 / *dy, az = 1.0 / *dz, av = fabs(*vel_n);
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double x = orc_max(fabs(AT(XP, i, j, k)), fabs(AT(XM, i, j, k)));
    double y = orc_max(fabs(AT(YP, i, j, k)), fabs(AT(YM, i, j, k)));
    double z = orc_max(fabs(AT(ZP, i, j, k)), fabs(AT(ZM, i, j, k)));
    double n = sqrt(x * x + y * y + z * z + mdx2);
    double c = av / n * (x * ax + y * ay + z * az);
    if (c > m) m = c;
  }
  *dt = *cfl / (m + ORC_TINY);
}

/* ---------------------------------------------------------------------------
 * Local reinitialization RHS, src/reinitialization/lsm_reinitialization3d_local.f.in:50-291
 * ------------------------------------------------------------------------- */
void oracle_lsm3dcomputereinitializationeqnrhslocal(double *rr, BOXP(rg), const double *phi, BOXP(pg),
                                                    const double *phi0, BOXP(p0g), const double *pxp,
                                                    const double *pyp, const double *pzp, BOXP(gp),
                                                    const double *pxm, const double *pym, const double *pzm,
                                                    BOXP(gm), const double *dx, const double *dy,
                                                    const double *dz, const int *use_phi0_for_sgn,
                                                    const int *ix, const int *iy, const int *iz, const int *nlo,
                                                    const int *nhi, const unsigned char *nb, BOXP(nbg),
                                                    const unsigned char *mark_fb) {
  orc_arr R = VIEW(rr, rg), F = VIEW(phi, pg), F0 = VIEW(phi0, p0g);
  orc_arr XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp);
  orc_arr XM = VIEW(pxm, gm), YM = VIEW(pym, gm), ZM = VIEW(pzm, gm);
  orc_mask N = MVIEW(nb, nbg);
  double h = orc_max(orc_max(*dx, *dy), *dz), dx_sq = h * h;
  int use0 = (*use_phi0_for_sgn == 1);
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double s = use0 ? AT(F0, i, j, k) : AT(F, i, j, k);
    double gpv[3] = {AT(XP, i, j, k), AT(YP, i, j, k), AT(ZP, i, j, k)};
    double gmv[3] = {AT(XM, i, j, k), AT(YM, i, j, k), AT(ZM, i, j, k)};
    double star[3];
    for (int d = 0; d < 3; d++) {
      double p, m;
      if (s > 0.0) { p = orc_max(-gpv[d], 0.0); m = orc_max(gmv[d], 0.0); }
      else { p = orc_max(gpv[d], 0.0); m = orc_max(-gmv[d], 0.0); }
      star[d] = orc_max(p, m);
    }
    if (fabs(s) >= ORC_ZERO_TOL) {
      double n2 = star[0] * star[0] + star[1] * star[1] + star[2] * star[2];
      double sg = use0 ? s / sqrt(s * s + dx_sq) : s / sqrt(s * s + n2 * dx_sq);
      AT(R, i, j, k) = sg * (1.0 - sqrt(n2));
    } else {
      AT(R, i, j, k) = 0.0;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Local mean curvature, src/geometry/lsm_curvature3d_local.f.in:15-144
 * ------------------------------------------------------------------------- */
void oracle_lsm3dcomputemeancurvatureorder2local(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                                 const double *px, const double *py, const double *pz,
                                                 const double *gmag, BOXP(gg), const double *dx,
                                                 const double *dy, const double *dz, const int *ix,
                                                 const int *iy, const int *iz, const int *nlo, const int *nhi,
                                                 const unsigned char *nb, BOXP(nbg),
                                                 const unsigned char *mark_fb) {
  orc_arr K = VIEW(kappa, kg), F = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg),
          G = VIEW(gmag, gg);
  orc_mask N = MVIEW(nb, nbg);
  double fxx = 1.0 / *dx / *dx, fyy = 1.0 / *dy / *dy, fzz = 1.0 / *dz / *dz;
  double fxy = 0.25 / *dx / *dy, fxz = 0.25 / *dx / *dz, fyz = 0.25 / *dy / *dz;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
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
    double g = AT(G, i, j, k), den = g * g * g;
    if (den < ORC_ZERO_TOL) AT(K, i, j, k) = 0.0;
    else AT(K, i, j, k) = curv_num(xx, xy, xz, yy, yz, zz, AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k)) / den;
  }
}

/* ---------------------------------------------------------------------------
 * Band construction, src/toolbox/lsm_localization3d.f
 * ------------------------------------------------------------------------- */
/* :48-167 -- first-claim growth of levels 1..level in parent order and
 * neighbour order +x, -x, +y, -y, +z, -z. */
void oracle_lsm3dmarknarrowbandneighbors(unsigned char *nb, BOXP(g), int *ix, int *iy, int *iz,
                                         const int *nlo, const int *nhi, int *n_lo, int *n_hi,
                                         const int *level) {
  orc_mask N = MVIEW(nb, g);
  int base = *nlo;
  (void)nhi;
  int count = n_hi[0] + 1;
  static const int dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  int lo[3] = {*ilo_g, *jlo_g, *klo_g}, hi[3] = {*ihi_g, *jhi_g, *khi_g};
  for (int L = 1; L <= *level; L++) {
    unsigned char mark = (unsigned char)(L + 1);
    for (int m = n_lo[L - 1]; m <= n_hi[L - 1]; m++) {
      int p[3] = {LIDX(ix, base, m), LIDX(iy, base, m), LIDX(iz, base, m)};
      for (int d = 0; d < 6; d++) {
        int ax = d / 2, up = (d % 2 == 0);
        if (up ? !(p[ax] < hi[ax]) : !(p[ax] > lo[ax])) continue;
        int q0 = p[0] + dirs[d][0], q1 = p[1] + dirs[d][1], q2 = p[2] + dirs[d][2];
        if (AT(N, q0, q1, q2) == 0) {
          LIDX(ix, base, count) = q0;
          LIDX(iy, base, count) = q1;
          LIDX(iz, base, count) = q2;
          AT(N, q0, q1, q2) = mark;
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

/* :174-297 -- level 0 is the (k,j,i)-ordered set |phi| < width over phi's
 * ghost box; outer-layer points (|phi| >= width_inner) are recorded in
 * index_outer, phi <= 0 from the front and phi > 0 from the back. */
void oracle_lsm3ddeterminenarrowband(const double *phi, BOXP(g), unsigned char *nb, BOXP(nbg), int *ix, int *iy,
                                     int *iz, const int *nlo, const int *nhi, int *n_lo, int *n_hi,
                                     int *index_outer, const int *nlo_outer, const int *nhi_outer,
                                     int *nlo_outer_plus, int *nhi_outer_plus, int *nlo_outer_minus,
                                     int *nhi_outer_minus, const double *width, const double *width_inner,
                                     const int *level) {
  orc_arr F = VIEW(phi, g);
  orc_mask N = MVIEW(nb, nbg);
  int base = *nlo, obase = *nlo_outer;
  int count = *nlo;
  n_lo[0] = *nlo;
  int cminus = *nlo_outer, cplus = *nhi_outer;
  *nlo_outer_minus = *nlo_outer;
  *nhi_outer_plus = *nhi_outer;
  for (int k = *klo_g; k <= *khi_g; k++)
    for (int j = *jlo_g; j <= *jhi_g; j++)
      for (int i = *ilo_g; i <= *ihi_g; i++) {
        double v = AT(F, i, j, k), a = fabs(v);
        if (a < *width) {
          LIDX(ix, base, count) = i;
          LIDX(iy, base, count) = j;
          LIDX(iz, base, count) = k;
          AT(N, i, j, k) = 1;
          if (a >= *width_inner) {
            if (v <= 0.0) { LIDX(index_outer, obase, cminus) = count; cminus++; }
            else { LIDX(index_outer, obase, cplus) = count; cplus--; }
          }
          count++;
        } else {
          AT(N, i, j, k) = 0;
        }
      }
  if (count > *nlo) {
    n_hi[0] = count - 1;
    *nhi_outer_minus = cminus - 1;
    *nlo_outer_plus = cplus + 1;
    oracle_lsm3dmarknarrowbandneighbors(nb, ilo_nbg, ihi_nbg, jlo_nbg, jhi_nbg, klo_nbg, khi_nbg, ix, iy, iz, nlo,
                                        nhi, n_lo, n_hi, level);
  } else {
    n_hi[0] = count;
    *nhi_outer_minus = cminus;
    *nlo_outer_plus = cplus;
  }
}

/* :303-377 -- the six fill-box planes, each spanning the ghost box */
void oracle_lsm3dmarknarrowbandboundarylayer(unsigned char *nb, BOXP(g), BOXP(fb), const unsigned char *mark) {
  orc_mask N = MVIEW(nb, g);
  for (int k = *klo_g; k <= *khi_g; k++)
    for (int j = *jlo_g; j <= *jhi_g; j++) {
      AT(N, *ilo_fb, j, k) = *mark;
      AT(N, *ihi_fb, j, k) = *mark;
    }
  for (int k = *klo_g; k <= *khi_g; k++)
    for (int i = *ilo_g; i <= *ihi_g; i++) {
      AT(N, i, *jlo_fb, k) = *mark;
      AT(N, i, *jhi_fb, k) = *mark;
    }
  for (int j = *jlo_g; j <= *jhi_g; j++)
    for (int i = *ilo_g; i <= *ihi_g; i++) {
      AT(N, i, j, *klo_fb) = *mark;
      AT(N, i, j, *khi_fb) = *mark;
    }
}

/* :385-475 */
void oracle_lsm3ddeterminenarrowbandfrommask(const double *mask, BOXP(mg), unsigned char *nb, BOXP(nbg), int *ix,
                                             int *iy, int *iz, const int *nlo, const int *nhi, int *n_lo,
                                             int *n_hi, const int *level, const int *use_mask_sign) {
  orc_arr Mk = VIEW(mask, mg);
  orc_mask N = MVIEW(nb, nbg);
  int base = *nlo, count = *nlo;
  n_lo[0] = *nlo;
  for (int k = *klo_mg; k <= *khi_mg; k++)
    for (int j = *jlo_mg; j <= *jhi_mg; j++)
      for (int i = *ilo_mg; i <= *ihi_mg; i++) {
        double v = AT(Mk, i, j, k);
        int in = (*use_mask_sign > 0) ? (v >= 0) : (v <= 0);
        if (in) {
          LIDX(ix, base, count) = i;
          LIDX(iy, base, count) = j;
          LIDX(iz, base, count) = k;
          AT(N, i, j, k) = 1;
          count++;
        } else {
          AT(N, i, j, k) = 0;
        }
      }
  n_hi[0] = count - 1;
  oracle_lsm3dmarknarrowbandneighbors(nb, ilo_nbg, ihi_nbg, jlo_nbg, jhi_nbg, klo_nbg, khi_nbg, ix, iy, iz, nlo,
                                      nhi, n_lo, n_hi, level);
}

/* :482-561 */
void oracle_lsm3dmultiplycutofflserhslocal(const double *phi, double *rhs, BOXP(g), const int *ix, const int *iy,
                                           const int *iz, const int *nlo, const int *nhi, const unsigned char *nb,
                                           BOXP(nbg), const unsigned char *mark_fb, const double *beta,
                                           const double *gamma) {
  orc_arr F = VIEW(phi, g), R = VIEW(rhs, g);
  orc_mask N = MVIEW(nb, nbg);
  double b = *beta, gm = *gamma;
  double c1 = gm - 3.0 * b;
  double c2 = gm - b;
  c2 = c2 * c2 * c2;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double a = fabs(AT(F, i, j, k)), cut;
    if (a <= b) cut = 1.0;
    else if (a <= gm) {
      double t = a - gm;
      cut = (t * t * (2.0 * a + c1)) / c2;
    } else cut = 0.0;
    AT(R, i, j, k) = cut * AT(R, i, j, k);
  }
}

/* :568-653 */
void oracle_lsm3dcheckouternarrowbandlayer(int *change_sign, const double *phi, BOXP(g), const int *ix,
                                           const int *iy, const int *iz, const int *nlo, const int *nhi,
                                           const int *index_outer, const int *nlo_outer, const int *nhi_outer,
                                           const int *nlo_outer_plus, const int *nhi_outer_plus,
                                           const int *nlo_outer_minus, const int *nhi_outer_minus) {
  orc_arr F = VIEW(phi, g);
  int base = *nlo, obase = *nlo_outer;
  (void)nhi; (void)nhi_outer;
  int sp = 0, sm = 0;
  *change_sign = 0;
  for (int m = *nlo_outer_plus; m <= *nhi_outer_plus; m++) {
    int l = LIDX(index_outer, obase, m);
    if (AT(F, LIDX(ix, base, l), LIDX(iy, base, l), LIDX(iz, base, l)) > 0.0) sp = 1;
    else sm = 1;
    if (sp == 1 && sm == 1) { *change_sign = 1; break; }
  }
  if (*change_sign == 0) {
    sp = 0;
    sm = 0;
    for (int m = *nlo_outer_minus; m <= *nhi_outer_minus; m++) {
      int l = LIDX(index_outer, obase, m);
      if (AT(F, LIDX(ix, base, l), LIDX(iy, base, l), LIDX(iz, base, l)) > 0.0) sp = 1;
      else sm = 1;
      if (sp == 1 && sm == 1) { *change_sign = 1; break; }
    }
  }
}

/* :660-719 */
void oracle_lsm3dimposemasklocal(double *dest, BOXP(dg), const double *src, BOXP(sg), const double *mask,
                                 BOXP(mg), const int *ix, const int *iy, const int *iz, const int *nlo,
                                 const int *nhi) {
  orc_arr D = VIEW(dest, dg), S = VIEW(src, sg), Mk = VIEW(mask, mg);
  for (int l = *nlo; l <= *nhi; l++) {
    int i = LIDX(ix, *nlo, l), j = LIDX(iy, *nlo, l), k = LIDX(iz, *nlo, l);
    AT(D, i, j, k) = orc_max(AT(Mk, i, j, k), AT(S, i, j, k));
  }
}

/* :725-774 */
void oracle_lsm3dcopydatalocal(double *dest, BOXP(dg), const double *src, BOXP(sg), const int *ix,
                               const int *iy, const int *iz, const int *nlo, const int *nhi) {
  orc_arr D = VIEW(dest, dg), S = VIEW(src, sg);
  for (int l = *nlo; l <= *nhi; l++) {
    int i = LIDX(ix, *nlo, l), j = LIDX(iy, *nlo, l), k = LIDX(iz, *nlo, l);
    AT(D, i, j, k) = AT(S, i, j, k);
  }
}
