/*
 * lsm_spatial_derivatives3d.h -- whole-grid HJ-ENO/WENO and central
 * derivatives on the GPU.  Same macro names, symbols and argument lists as
 * the reference's src/toolbox/lsm_spatial_derivatives3d.h.
 *
 * The D1/D2/D3 scratch arguments are accepted and left untouched: each thread
 * forms the undivided differences it needs in registers with the reference's
 * floating-point operations, so results are bitwise those of the reference
 * (FP64 build) while no full-grid difference array is written.
 */
#ifndef LSMHIP_SPATIAL_DERIVATIVES_3D_H
#define LSMHIP_SPATIAL_DERIVATIVES_3D_H

#include "lsmhip_config.h"

#define LSM3D_HJ_ENO1 lsm3dhjeno1_
#define LSM3D_HJ_ENO2 lsm3dhjeno2_
#define LSM3D_HJ_ENO3 lsm3dhjeno3_
#define LSM3D_HJ_WENO5 lsm3dhjweno5_
#define LSM3D_UPWIND_HJ_ENO1 lsm3dupwindhjeno1_
#define LSM3D_UPWIND_HJ_ENO2 lsm3dupwindhjeno2_
#define LSM3D_UPWIND_HJ_ENO3 lsm3dupwindhjeno3_
#define LSM3D_UPWIND_HJ_WENO5 lsm3dupwindhjweno5_
#define LSM3D_CENTRAL_GRAD_ORDER2 lsm3dcentralgradorder2_
#define LSM3D_CENTRAL_GRAD_ORDER4 lsm3dcentralgradorder4_
#define LSM3D_LAPLACIAN_ORDER2 lsm3dlaplacianorder2_
#define LSM3D_AVERAGE_GRAD_PHI lsm3daveragegradphi_
#define LSM3D_GRADIENT_MAGNITUDE lsm3dgradientmagnitude_
#define LSM3D_PHI_UPWIND_GRAD_F lsm3dphiupwindgradf_
#define LSM3D_CENTRAL_HESSIAN lsm3dcentralhessian_

LSMHIP_BEGIN_DECLS

/* Plus/minus derivative arrays, phi, then the scheme's scratch (ignored). */
#define LSMHIP_PM_ARGS                                                                                      \
  LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), \
      LSMLIB_REAL *phi_x_minus, LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,                           \
      LSMHIP_BOX_ARGS(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb)
#define LSMHIP_UPWIND_ARGS                                                                                   \
  LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *phi, \
      LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z,     \
      LSMHIP_BOX_ARGS(vel_gb)
#define LSMHIP_DXYZ const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz

/* replaces src/toolbox/lsm_spatial_derivatives3d.h:84 */
void LSM3D_HJ_ENO1(LSMHIP_PM_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :160 */
void LSM3D_HJ_ENO2(LSMHIP_PM_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMLIB_REAL *D2, LSMHIP_BOX_ARGS(D2_gb),
                   LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :244 */
void LSM3D_HJ_ENO3(LSMHIP_PM_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMLIB_REAL *D2, LSMHIP_BOX_ARGS(D2_gb),
                   LSMLIB_REAL *D3, LSMHIP_BOX_ARGS(D3_gb), LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :333 */
void LSM3D_HJ_WENO5(LSMHIP_PM_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :396 */
void LSM3D_UPWIND_HJ_ENO1(LSMHIP_UPWIND_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMHIP_BOX_ARGS(fb),
                          LSMHIP_DXYZ);
/* replaces :460 */
void LSM3D_UPWIND_HJ_ENO2(LSMHIP_UPWIND_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMLIB_REAL *D2,
                          LSMHIP_BOX_ARGS(D2_gb), LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :532 */
void LSM3D_UPWIND_HJ_ENO3(LSMHIP_UPWIND_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMLIB_REAL *D2,
                          LSMHIP_BOX_ARGS(D2_gb), LSMLIB_REAL *D3, LSMHIP_BOX_ARGS(D3_gb), LSMHIP_BOX_ARGS(fb),
                          LSMHIP_DXYZ);
/* replaces :609 */
void LSM3D_UPWIND_HJ_WENO5(LSMHIP_UPWIND_ARGS, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMHIP_BOX_ARGS(fb),
                           LSMHIP_DXYZ);
/* replaces :674 */
void LSM3D_CENTRAL_GRAD_ORDER2(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                               LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                               LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :724 */
void LSM3D_CENTRAL_GRAD_ORDER4(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                               LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                               LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :778 */
void LSM3D_LAPLACIAN_ORDER2(LSMLIB_REAL *laplacian_phi, LSMHIP_BOX_ARGS(laplacian_phi_gb), const LSMLIB_REAL *phi,
                            LSMHIP_BOX_ARGS(phi_gb), LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);
/* replaces :894 */
void LSM3D_AVERAGE_GRAD_PHI(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb),
                            LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus,
                            LSMHIP_BOX_ARGS(grad_phi_plus_gb), LSMLIB_REAL *phi_x_minus, LSMLIB_REAL *phi_y_minus,
                            LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb), LSMHIP_BOX_ARGS(fb));
/* replaces :940 (thresholds 1e-8 on each squared component and on the result) */
void LSM3D_GRADIENT_MAGNITUDE(const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                              LSMLIB_REAL *grad_phi_mag, LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(grad_phi_fb));
/* replaces :832.  Cells where phi is exactly zero (or NaN) are left as they
 * are, as in the reference. */
void LSM3D_PHI_UPWIND_GRAD_F(LSMLIB_REAL *F_x, LSMLIB_REAL *F_y, LSMLIB_REAL *F_z, LSMHIP_BOX_ARGS(grad_F_gb),
                             LSMLIB_REAL *F_x_plus, LSMLIB_REAL *F_y_plus, LSMLIB_REAL *F_z_plus,
                             LSMHIP_BOX_ARGS(grad_F_plus_gb), LSMLIB_REAL *F_x_minus, LSMLIB_REAL *F_y_minus,
                             LSMLIB_REAL *F_z_minus, LSMHIP_BOX_ARGS(grad_F_minus_gb), const LSMLIB_REAL *phi,
                             LSMHIP_BOX_ARGS(phi_gb), LSMHIP_BOX_ARGS(fb));
/* replaces :971 */
void LSM3D_CENTRAL_HESSIAN(LSMLIB_REAL *phi_xx, LSMLIB_REAL *phi_xy, LSMLIB_REAL *phi_xz, LSMLIB_REAL *phi_yy,
                           LSMLIB_REAL *phi_yz, LSMLIB_REAL *phi_zz, LSMHIP_BOX_ARGS(hessian_gb),
                           const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), LSMHIP_DXYZ);

LSMHIP_END_DECLS

#endif
