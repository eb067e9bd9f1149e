/*
 * lsm_spatial_derivatives2d.h -- 2D HJ-ENO/WENO derivatives, central gradients, Laplacian, gradient magnitude on the GPU
 * (reference: src/toolbox/lsm_spatial_derivatives2d.h).
 */
#ifndef LSMHIP_LSM_SPATIAL_DERIVATIVES2D_H
#define LSMHIP_LSM_SPATIAL_DERIVATIVES2D_H

#include "lsmhip_config.h"

#define LSM2D_HJ_ENO1 lsm2dhjeno1_
#define LSM2D_HJ_ENO2 lsm2dhjeno2_
#define LSM2D_HJ_ENO3 lsm2dhjeno3_
#define LSM2D_HJ_WENO5 lsm2dhjweno5_
#define LSM2D_UPWIND_HJ_ENO1 lsm2dupwindhjeno1_
#define LSM2D_UPWIND_HJ_ENO2 lsm2dupwindhjeno2_
#define LSM2D_UPWIND_HJ_ENO3 lsm2dupwindhjeno3_
#define LSM2D_UPWIND_HJ_WENO5 lsm2dupwindhjweno5_
#define LSM2D_CENTRAL_GRAD_ORDER2 lsm2dcentralgradorder2_
#define LSM2D_CENTRAL_GRAD_ORDER4 lsm2dcentralgradorder4_
#define LSM2D_LAPLACIAN_ORDER2 lsm2dlaplacianorder2_
#define LSM2D_AVERAGE_GRAD_PHI lsm2daveragegradphi_
#define LSM2D_GRADIENT_MAGNITUDE lsm2dgradientmagnitude_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_spatial_derivatives2d.h:84 */
void LSM2D_HJ_ENO1(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL *phi_x_minus, LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb),
    LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:147 */
void LSM2D_HJ_ENO2(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL *phi_x_minus, LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMLIB_REAL *D2,
    LSMHIP_BOX2_ARGS(D2_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:216 */
void LSM2D_HJ_ENO3(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL *phi_x_minus, LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMLIB_REAL *D2,
    LSMHIP_BOX2_ARGS(D2_gb), LSMLIB_REAL *D3, LSMHIP_BOX2_ARGS(D3_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:287 */
void LSM2D_HJ_WENO5(LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL *phi_x_minus, LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb),
    LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:337 */
void LSM2D_UPWIND_HJ_ENO1(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    LSMHIP_BOX2_ARGS(vel_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMHIP_BOX2_ARGS(fb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:388 */
void LSM2D_UPWIND_HJ_ENO2(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    LSMHIP_BOX2_ARGS(vel_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMLIB_REAL *D2, LSMHIP_BOX2_ARGS(D2_gb),
    LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:445 */
void LSM2D_UPWIND_HJ_ENO3(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    LSMHIP_BOX2_ARGS(vel_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMLIB_REAL *D2, LSMHIP_BOX2_ARGS(D2_gb),
    LSMLIB_REAL *D3, LSMHIP_BOX2_ARGS(D3_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:505 */
void LSM2D_UPWIND_HJ_WENO5(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    LSMHIP_BOX2_ARGS(vel_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMHIP_BOX2_ARGS(fb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:557 */
void LSM2D_CENTRAL_GRAD_ORDER2(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:599 */
void LSM2D_CENTRAL_GRAD_ORDER4(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:643 */
void LSM2D_LAPLACIAN_ORDER2(LSMLIB_REAL *laplacian_phi, LSMHIP_BOX2_ARGS(laplacian_phi_gb), const LSMLIB_REAL *phi,
    LSMHIP_BOX2_ARGS(phi_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:739 */
void LSM2D_AVERAGE_GRAD_PHI(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    LSMHIP_BOX2_ARGS(fb));
/* replaces src/toolbox/lsm_spatial_derivatives2d.h:774 */
void LSM2D_GRADIENT_MAGNITUDE(const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMLIB_REAL *grad_phi_mag,
    LSMHIP_BOX2_ARGS(grad_phi_gb), LSMHIP_BOX2_ARGS(fb));

LSMHIP_END_DECLS

#endif
