/*
 * oracle_geometry.c -- TEST INFRASTRUCTURE ONLY (see lsm_oracle.h).
 *
 * The kernels on either side of the curvature / reinitialization path:
 * unit and signed unit normals, mean curvature from a precomputed Hessian,
 * Gaussian curvature (second and fourth order), fourth-order mean curvature,
 * the orthogonalization RHS and the field-extension RHS.  Each function cites
 * the reference routine whose operation order it keeps; arithmetic is FP64
 * without contraction, so the results compare bitwise with the reference's
 * kernels (the one pow() of the mean curvature is the C library's, as in the
 * reference's compiled Fortran).
 */
#include "lsm_oracle.h"

typedef struct {
  int ilo, ihi, jlo, jhi, klo, khi;
} gbox;

static gbox mk(const int *ilo, const int *ihi, const int *jlo, const int *jhi, const int *klo, const int *khi) {
  gbox b = {*ilo, *ihi, *jlo, *jhi, *klo, *khi};
  return b;
}
#define GBOX(n) mk(ilo_##n, ihi_##n, jlo_##n, jhi_##n, klo_##n, khi_##n)
#define FOR_BOX(b)                             \
  for (int k = (b).klo; k <= (b).khi; k++)     \
    for (int j = (b).jlo; j <= (b).jhi; j++)   \
      for (int i = (b).ilo; i <= (b).ihi; i++)

static double max3(double a, double b, double c) { return orc_max(orc_max(a, b), c); }

/* ---------------------------------------------------------------------------
 * Unit normals and mean curvature, src/geometry/lsm_geometry3d_fort.f.in
 * ------------------------------------------------------------------------- */

/* lsm3dComputeUnitNormal, :27-127.  n = grad(phi)/|grad(phi)|; (1,0,0) where
 * |grad(phi)| < zero_tol. */
void oracle_lsm3dcomputeunitnormal(double *nx, double *ny, double *nz, BOXP(ng), const double *px,
                                   const double *py, const double *pz, BOXP(gg), BOXP(fb)) {
  orc_arr NX = VIEW(nx, ng), NY = VIEW(ny, ng), NZ = VIEW(nz, ng), X = VIEW(px, gg), Y = VIEW(py, gg),
          Z = VIEW(pz, gg);
  gbox b = GBOX(fb);
  FOR_BOX(b) {
    double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
    double norm = sqrt(x * x + y * y + z * z);
    if (norm >= ORC_ZERO_TOL) {
      double inv = 1.0 / norm;
      AT(NX, i, j, k) = x * inv;
      AT(NY, i, j, k) = y * inv;
      AT(NZ, i, j, k) = z * inv;
    } else {
      AT(NX, i, j, k) = 1.0;
      AT(NY, i, j, k) = 0.0;
      AT(NZ, i, j, k) = 0.0;
    }
  }
}

/* lsm3dComputeSignedUnitNormal, :129-272.  n = sgn(phi) grad(phi)/|grad(phi)|
 * with the smoothed sign phi/sqrt(phi^2 + |grad(phi)|^2 max(dx,dy,dz)^2);
 * (0,0,0) where |phi| <= zero_tol, (1,0,0) where |grad(phi)|^2 < zero_tol. */
void oracle_lsm3dcomputesignedunitnormal(double *nx, double *ny, double *nz, BOXP(ng), const double *px,
                                         const double *py, const double *pz, BOXP(gg), const double *phi,
                                         BOXP(pg), BOXP(fb), const double *dx, const double *dy,
                                         const double *dz) {
  orc_arr NX = VIEW(nx, ng), NY = VIEW(ny, ng), NZ = VIEW(nz, ng), X = VIEW(px, gg), Y = VIEW(py, gg),
          Z = VIEW(pz, gg), P = VIEW(phi, pg);
  double h = max3(*dx, *dy, *dz), dx_sq = h * h;
  gbox b = GBOX(fb);
  FOR_BOX(b) {
    double c = AT(P, i, j, k);
    if (fabs(c) > ORC_ZERO_TOL) {
      double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
      double ngs = x * x + y * y + z * z;
      if (ngs >= ORC_ZERO_TOL) {
        double sgn = c / sqrt(c * c + ngs * dx_sq);
        double inv = 1.0 / sqrt(ngs);
        AT(NX, i, j, k) = sgn * x * inv;
        AT(NY, i, j, k) = sgn * y * inv;
        AT(NZ, i, j, k) = sgn * z * inv;
      } else {
        AT(NX, i, j, k) = 1.0;
        AT(NY, i, j, k) = 0.0;
        AT(NZ, i, j, k) = 0.0;
      }
    } else {
      AT(NX, i, j, k) = 0.0;
      AT(NY, i, j, k) = 0.0;
      AT(NZ, i, j, k) = 0.0;
    }
  }
}

/* lsm3dComputeMeanCurvature, :274-394: kappa from a precomputed gradient and
 * Hessian, numerator / |grad(phi)|^3 with |grad(phi)|^3 = pow(|grad(phi)|^2, 1.5);
 * 0 where |grad(phi)|^2 < zero_tol. */
void oracle_lsm3dcomputemeancurvature(double *kappa, BOXP(kg), const double *px, const double *py,
                                      const double *pz, BOXP(gg), const double *pxx, const double *pxy,
                                      const double *pxz, const double *pyy, const double *pyz, const double *pzz,
                                      BOXP(hg), BOXP(fb)) {
  orc_arr K = VIEW(kappa, kg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg), XX = VIEW(pxx, hg),
          XY = VIEW(pxy, hg), XZ = VIEW(pxz, hg), YY = VIEW(pyy, hg), YZ = VIEW(pyz, hg), ZZ = VIEW(pzz, hg);
  gbox b = GBOX(fb);
  FOR_BOX(b) {
    double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
    double ngs = x * x + y * y + z * z;
    if (ngs < ORC_ZERO_TOL) {
      AT(K, i, j, k) = 0.0;
    } else {
      double xx = AT(XX, i, j, k), xy = AT(XY, i, j, k), xz = AT(XZ, i, j, k), yy = AT(YY, i, j, k),
             yz = AT(YZ, i, j, k), zz = AT(ZZ, i, j, k);
      double num = xx * y * y - 2.0 * xy * x * y + yy * x * x + xx * z * z - 2.0 * xz * x * z + zz * x * x +
                   yy * z * z - 2.0 * yz * y * z + zz * y * y;
      AT(K, i, j, k) = num / pow(ngs, 1.5);
    }
  }
}

/* ---------------------------------------------------------------------------
 * Curvatures from phi, src/geometry/lsm_curvature3d.f.in
 * ------------------------------------------------------------------------- */

/* Gaussian-curvature numerator of :230-237 (second order) and :490-497
 * (fourth order, with the x-z mixed derivative taken from phi_z along x). */
static double gauss_num(double x, double y, double z, double xx, double yy, double zz, double xy, double xz,
                        double yz) {
  return x * x * (yy * zz - yz * yz) + y * y * (xx * zz - xz * xz) + z * z * (xx * yy - xy * xy) +
         2.0 * (x * y * (xz * yz - xy * zz) + y * z * (xy * xz - yz * xx) + x * z * (xy * yz - xz * yy));
}

/* lsm3dComputeGaussianCurvatureOrder2, :136-248: compact second differences
 * (the 2-point cross stencil of the mean curvature), numerator / |grad(phi)|^4,
 * 0 where |grad(phi)|^4 < zero_tol. */
void oracle_lsm3dcomputegaussiancurvatureorder2(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                                const double *px, const double *py, const double *pz,
                                                const double *gmag, BOXP(gg), BOXP(fb), const double *dx,
                                                const double *dy, const double *dz) {
  orc_arr K = VIEW(kappa, kg), F = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg),
          G = VIEW(gmag, gg);
  double fxx = 1.0 / *dx / *dx, fyy = 1.0 / *dy / *dy, fzz = 1.0 / *dz / *dz;
  double fxy = 0.25 / *dx / *dy, fxz = 0.25 / *dx / *dz, fyz = 0.25 / *dy / *dz;
  gbox b = GBOX(fb);
  FOR_BOX(b) {
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
    double den = g * g * g * g;
    if (den < ORC_ZERO_TOL)
      AT(K, i, j, k) = 0.0;
    else
      AT(K, i, j, k) = gauss_num(AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k), xx, yy, zz, xy, xz, yz) / den;
  }
}

/* Fourth-order second derivatives of :330-341 / :462-473: (-f(+2) + 16 f(+1)
 * - 30 f(0) - f(-2) + 16 f(-1)) / (12 h^2), and the mixed derivatives of
 * :342-351 / :474-483 as fourth-order first differences of the precomputed first
 * derivatives: phi_xy = d/dy phi_x, phi_yz = d/dz phi_y, phi_zx = d/dx phi_z. */
typedef struct {
  double xx, yy, zz, xy, yz, zx;
} hess4;

static hess4 hessian_order4(orc_arr F, orc_arr X, orc_arr Y, orc_arr Z, int i, int j, int k, const double f2[3],
                            const double f1[3]) {
  hess4 h;
  double c = AT(F, i, j, k);
  h.xx = (-AT(F, i + 2, j, k) + 16.0 * AT(F, i + 1, j, k) - 30.0 * c - AT(F, i - 2, j, k) +
          16.0 * AT(F, i - 1, j, k)) * f2[0];
  h.yy = (-AT(F, i, j + 2, k) + 16.0 * AT(F, i, j + 1, k) - 30.0 * c - AT(F, i, j - 2, k) +
          16.0 * AT(F, i, j - 1, k)) * f2[1];
  h.zz = (-AT(F, i, j, k + 2) + 16.0 * AT(F, i, j, k + 1) - 30.0 * c - AT(F, i, j, k - 2) +
          16.0 * AT(F, i, j, k - 1)) * f2[2];
  h.xy = (-AT(X, i, j + 2, k) + 8.0 * AT(X, i, j + 1, k) + AT(X, i, j - 2, k) - 8.0 * AT(X, i, j - 1, k)) * f1[1];
  h.yz = (-AT(Y, i, j, k + 2) + 8.0 * AT(Y, i, j, k + 1) + AT(Y, i, j, k - 2) - 8.0 * AT(Y, i, j, k - 1)) * f1[2];
  h.zx = (-AT(Z, i + 2, j, k) + 8.0 * AT(Z, i + 1, j, k) + AT(Z, i - 2, j, k) - 8.0 * AT(Z, i - 1, j, k)) * f1[0];
  return h;
}

#define ORDER4_FACTORS                                                                              \
  const double c12 = 0.0833333333333333333333;                                                     \
  double f2[3] = {c12 / *dx / *dx, c12 / *dy / *dy, c12 / *dz / *dz};                               \
  double f1[3] = {c12 / *dx, c12 / *dy, c12 / *dz}

/* lsm3dComputeMeanCurvatureOrder4, :253-376: numerator in the order of
 * :358-366, over |grad(phi)|^3. */
void oracle_lsm3dcomputemeancurvatureorder4(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                            const double *px, const double *py, const double *pz,
                                            const double *gmag, BOXP(gg), BOXP(fb), const double *dx,
                                            const double *dy, const double *dz) {
  orc_arr K = VIEW(kappa, kg), F = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg),
          G = VIEW(gmag, gg);
  ORDER4_FACTORS;
  gbox b = GBOX(fb);
  FOR_BOX(b) {
    hess4 h = hessian_order4(F, X, Y, Z, i, j, k, f2, f1);
    double g = AT(G, i, j, k);
    double den = g * g * g;
    if (den < ORC_ZERO_TOL) {
      AT(K, i, j, k) = 0.0;
    } else {
      double x = AT(X, i, j, k), y = AT(Y, i, j, k), z = AT(Z, i, j, k);
      double num = h.xx * y * y + h.yy * x * x - 2.0 * h.xy * x * y + h.xx * z * z + h.zz * x * x -
                   2.0 * h.zx * x * z + h.yy * z * z + h.zz * y * y - 2.0 * h.yz * y * z;
      AT(K, i, j, k) = num / den;
    }
  }
}

/* lsm3dComputeGaussianCurvatureOrder4, :383-507: Gaussian numerator over
 * |grad(phi)|^4 with the fourth-order Hessian above. */
void oracle_lsm3dcomputegaussiancurvatureorder4(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                                const double *px, const double *py, const double *pz,
                                                const double *gmag, BOXP(gg), BOXP(fb), const double *dx,
                                                const double *dy, const double *dz) {
  orc_arr K = VIEW(kappa, kg), F = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg),
          G = VIEW(gmag, gg);
  ORDER4_FACTORS;
  gbox b = GBOX(fb);
  FOR_BOX(b) {
    hess4 h = hessian_order4(F, X, Y, Z, i, j, k, f2, f1);
    double g = AT(G, i, j, k);
    double den = g * g * g * g;
    if (den < ORC_ZERO_TOL)
      AT(K, i, j, k) = 0.0;
    else
      AT(K, i, j, k) =
          gauss_num(AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k), h.xx, h.yy, h.zz, h.xy, h.zx, h.yz) / den;
  }
}

/* ---------------------------------------------------------------------------
 * Orthogonalization RHS, src/reinitialization/lsm_reinitialization3d.f.in:296-476
 * ------------------------------------------------------------------------- */

/* Upwinded -(sgn(psi) grad(psi)/|grad(psi)|) . grad(phi), grad(psi) the mean of
 * its plus and minus one-sided derivatives, the upwind side of phi's one-sided
 * derivatives picked by the sign of each component of sgn(psi) grad(psi);
 * 0 where |psi| <= zero_tol or |grad(psi)| <= zero_tol. */
void oracle_lsm3dcomputeorthogonalizationeqnrhs(double *rhs, BOXP(rg), const double *pxp, const double *pyp,
                                                const double *pzp, BOXP(gp), const double *pxm,
                                                const double *pym, const double *pzm, BOXP(gm),
                                                const double *psi, BOXP(sg), const double *qxp, const double *qyp,
                                                const double *qzp, BOXP(qp), const double *qxm,
                                                const double *qym, const double *qzm, BOXP(qm), BOXP(fb),
                                                const double *dx, const double *dy, const double *dz) {
  orc_arr R = VIEW(rhs, rg), XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp), XM = VIEW(pxm, gm),
          YM = VIEW(pym, gm), ZM = VIEW(pzm, gm), S = VIEW(psi, sg), QXP = VIEW(qxp, qp), QYP = VIEW(qyp, qp),
          QZP = VIEW(qzp, qp), QXM = VIEW(qxm, qm), QYM = VIEW(qym, qm), QZM = VIEW(qzm, qm);
  double h = max3(*dx, *dy, *dz), dx_sq = h * h;
  gbox b = GBOX(fb);
  FOR_BOX(b) {
    double g1 = 0.5 * (AT(QXP, i, j, k) + AT(QXM, i, j, k));
    double g2 = 0.5 * (AT(QYP, i, j, k) + AT(QYM, i, j, k));
    double g3 = 0.5 * (AT(QZP, i, j, k) + AT(QZM, i, j, k));
    double ng = sqrt(g1 * g1 + g2 * g2 + g3 * g3);
    double s = AT(S, i, j, k);
    if (fabs(s) > ORC_ZERO_TOL && ng > ORC_ZERO_TOL) {
      double sgn = s / sqrt(s * s + dx_sq * (ng * ng));
      double a1 = sgn * g1, a2 = sgn * g2, a3 = sgn * g3;
      AT(R, i, j, k) = -1.0 / ng *
                       (orc_max(a1, 0.0) * AT(XM, i, j, k) + orc_min(a1, 0.0) * AT(XP, i, j, k) +
                        orc_max(a2, 0.0) * AT(YM, i, j, k) + orc_min(a2, 0.0) * AT(YP, i, j, k) +
                        orc_max(a3, 0.0) * AT(ZM, i, j, k) + orc_min(a3, 0.0) * AT(ZP, i, j, k));
    } else {
      AT(R, i, j, k) = 0.0;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Narrow-band twins, src/geometry/lsm_curvature3d_local.f.in and
 * src/reinitialization/lsm_reinitialization3d_local.f.in: the whole-grid
 * formulas above at the band points nlo..nhi whose mark is <= mark_fb.
 * ------------------------------------------------------------------------- */
#define BAND_TAIL                                                                                   \
  const int *ix, const int *iy, const int *iz, const int *nlo, const int *nhi, const unsigned char *nb, \
      BOXP(nbg), const unsigned char *mark_fb

/* lsm3dComputeGaussianCurvatureOrder2Local, lsm_curvature3d_local.f.in:151-280 */
void oracle_lsm3dcomputegaussiancurvatureorder2local(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                                     const double *px, const double *py, const double *pz,
                                                     const double *gmag, BOXP(gg), const double *dx,
                                                     const double *dy, const double *dz, BAND_TAIL) {
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
    double g = AT(G, i, j, k), den = g * g * g * g;
    if (den < ORC_ZERO_TOL)
      AT(K, i, j, k) = 0.0;
    else
      AT(K, i, j, k) = gauss_num(AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k), xx, yy, zz, xy, xz, yz) / den;
  }
}

/* lsm3dComputeMeanCurvatureOrder4Local (:285-426) and
 * lsm3dComputeGaussianCurvatureOrder4Local (:432-573) evaluate the same
 * expression, the Gaussian numerator over |grad(phi)|^4 with the fourth-order
 * Hessian, so both symbols below run this one loop. */
static void curv4_band(double *kappa, BOXP(kg), const double *phi, BOXP(pg), const double *px, const double *py,
                       const double *pz, const double *gmag, BOXP(gg), const double *dx, const double *dy,
                       const double *dz, BAND_TAIL) {
  orc_arr K = VIEW(kappa, kg), F = VIEW(phi, pg), X = VIEW(px, gg), Y = VIEW(py, gg), Z = VIEW(pz, gg),
          G = VIEW(gmag, gg);
  orc_mask N = MVIEW(nb, nbg);
  ORDER4_FACTORS;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    hess4 h = hessian_order4(F, X, Y, Z, i, j, k, f2, f1);
    double g = AT(G, i, j, k), den = g * g * g * g;
    if (den < ORC_ZERO_TOL)
      AT(K, i, j, k) = 0.0;
    else
      AT(K, i, j, k) =
          gauss_num(AT(X, i, j, k), AT(Y, i, j, k), AT(Z, i, j, k), h.xx, h.yy, h.zz, h.xy, h.zx, h.yz) / den;
  }
}

#define CURV4_FORWARD                                                                                     \
  curv4_band(kappa, ilo_kg, ihi_kg, jlo_kg, jhi_kg, klo_kg, khi_kg, phi, ilo_pg, ihi_pg, jlo_pg, jhi_pg, klo_pg, \
             khi_pg, px, py, pz, gmag, ilo_gg, ihi_gg, jlo_gg, jhi_gg, klo_gg, khi_gg, dx, dy, dz, ix, iy, iz,   \
             nlo, nhi, nb, ilo_nbg, ihi_nbg, jlo_nbg, jhi_nbg, klo_nbg, khi_nbg, mark_fb)

void oracle_lsm3dcomputemeancurvatureorder4local(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                                 const double *px, const double *py, const double *pz,
                                                 const double *gmag, BOXP(gg), const double *dx, const double *dy,
                                                 const double *dz, BAND_TAIL) {
  CURV4_FORWARD;
}

void oracle_lsm3dcomputegaussiancurvatureorder4local(double *kappa, BOXP(kg), const double *phi, BOXP(pg),
                                                     const double *px, const double *py, const double *pz,
                                                     const double *gmag, BOXP(gg), const double *dx,
                                                     const double *dy, const double *dz, BAND_TAIL) {
  CURV4_FORWARD;
}

/* lsm3dComputeOrthogonalizationEqnRHSLocal, lsm_reinitialization3d_local.f.in:326-522 */
void oracle_lsm3dcomputeorthogonalizationeqnrhslocal(
    double *rhs, BOXP(rg), const double *pxp, const double *pyp, const double *pzp, BOXP(gp), const double *pxm,
    const double *pym, const double *pzm, BOXP(gm), const double *psi, BOXP(sg), const double *qxp,
    const double *qyp, const double *qzp, BOXP(qp), const double *qxm, const double *qym, const double *qzm,
    BOXP(qm), const double *dx, const double *dy, const double *dz, BAND_TAIL) {
  orc_arr R = VIEW(rhs, rg), XP = VIEW(pxp, gp), YP = VIEW(pyp, gp), ZP = VIEW(pzp, gp), XM = VIEW(pxm, gm),
          YM = VIEW(pym, gm), ZM = VIEW(pzm, gm), S = VIEW(psi, sg), QXP = VIEW(qxp, qp), QYP = VIEW(qyp, qp),
          QZP = VIEW(qzp, qp), QXM = VIEW(qxm, qm), QYM = VIEW(qym, qm), QZM = VIEW(qzm, qm);
  orc_mask N = MVIEW(nb, nbg);
  double h = max3(*dx, *dy, *dz), dx_sq = h * h;
  FOR_BAND(*nlo, *nhi, N, *mark_fb) {
    double g1 = 0.5 * (AT(QXP, i, j, k) + AT(QXM, i, j, k));
    double g2 = 0.5 * (AT(QYP, i, j, k) + AT(QYM, i, j, k));
    double g3 = 0.5 * (AT(QZP, i, j, k) + AT(QZM, i, j, k));
    double ng = sqrt(g1 * g1 + g2 * g2 + g3 * g3);
    double s = AT(S, i, j, k);
    if (fabs(s) > ORC_ZERO_TOL && ng > ORC_ZERO_TOL) {
      double sgn = s / sqrt(s * s + dx_sq * (ng * ng));
      double a1 = sgn * g1, a2 = sgn * g2, a3 = sgn * g3;
      AT(R, i, j, k) = -1.0 / ng *
                       (orc_max(a1, 0.0) * AT(XM, i, j, k) + orc_min(a1, 0.0) * AT(XP, i, j, k) +
                        orc_max(a2, 0.0) * AT(YM, i, j, k) + orc_min(a2, 0.0) * AT(YP, i, j, k) +
                        orc_max(a3, 0.0) * AT(ZM, i, j, k) + orc_min(a3, 0.0) * AT(ZP, i, j, k));
    } else {
      AT(R, i, j, k) = 0.0;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Field extension RHS, src/field_extension/lsm_field_extension3d.f:31-156
 * ------------------------------------------------------------------------- */

/* rhs = -(N . grad S) with N the signed unit normal and grad S upwinded along
 * it, where |phi| > 3 max(dx,dy,dz) or phi has the sign of all six face
 * neighbours; 0 in the cells the zero level set crosses. */
void oracle_lsm3dcomputefieldextensioneqnrhs(double *rhs, BOXP(rg), const double *S_, BOXP(sg), const double *phi,
                                             BOXP(pg), const double *sx, const double *sy, const double *sz,
                                             BOXP(ug), const double *nx, const double *ny, const double *nz,
                                             BOXP(ng), BOXP(fb), const double *dx, const double *dy,
                                             const double *dz) {
  orc_arr R = VIEW(rhs, rg), P = VIEW(phi, pg), SX = VIEW(sx, ug), SY = VIEW(sy, ug), SZ = VIEW(sz, ug),
          NX = VIEW(nx, ng), NY = VIEW(ny, ng), NZ = VIEW(nz, ng);
  double cutoff = 3.0 * max3(*dx, *dy, *dz);
  gbox b = GBOX(fb);
  (void)S_;
  FOR_BOX(b) {
    double c = AT(P, i, j, k);
    int away = fabs(c) > cutoff ||
               (c * AT(P, i - 1, j, k) > 0.0 && c * AT(P, i + 1, j, k) > 0.0 && c * AT(P, i, j - 1, k) > 0.0 &&
                c * AT(P, i, j + 1, k) > 0.0 && c * AT(P, i, j, k - 1) > 0.0 && c * AT(P, i, j, k + 1) > 0.0);
    AT(R, i, j, k) = away ? -(AT(NX, i, j, k) * AT(SX, i, j, k) + AT(NY, i, j, k) * AT(SY, i, j, k) +
                              AT(NZ, i, j, k) * AT(SZ, i, j, k))
                          : 0.0;
  }
}
