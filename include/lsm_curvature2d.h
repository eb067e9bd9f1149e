/*
 * lsm_curvature2d.h -- 2D mean curvature (orders 2 and 4) on the GPU
 * (reference: src/geometry/lsm_curvature2d.h).
 */
#ifndef LSMHIP_LSM_CURVATURE2D_H
#define LSMHIP_LSM_CURVATURE2D_H

#include "lsmhip_config.h"

#define LSM2D_COMPUTE_MEAN_CURVATURE_ORDER2 lsm2dcomputemeancurvatureorder2_
#define LSM2D_COMPUTE_MEAN_CURVATURE_ORDER4 lsm2dcomputemeancurvatureorder4_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_curvature2d.h:59 */
void LSM2D_COMPUTE_MEAN_CURVATURE_ORDER2(LSMLIB_REAL *kappa, LSMHIP_BOX2_ARGS(kappa_gb), const LSMLIB_REAL *phi,
    LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *grad_phi_mag,
    LSMHIP_BOX2_ARGS(grad_phi_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/geometry/lsm_curvature2d.h:175 */
void LSM2D_COMPUTE_MEAN_CURVATURE_ORDER4(LSMLIB_REAL *kappa, LSMHIP_BOX2_ARGS(kappa_gb), const LSMLIB_REAL *phi,
    LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *grad_phi_mag,
    LSMHIP_BOX2_ARGS(grad_phi_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);

LSMHIP_END_DECLS

#endif
