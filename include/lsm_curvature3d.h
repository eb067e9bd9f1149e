/*
 * lsm_curvature3d.h -- second-order mean curvature on the GPU
 * (reference: src/geometry/lsm_curvature3d.h).
 */
#ifndef LSMHIP_CURVATURE_3D_H
#define LSMHIP_CURVATURE_3D_H

#include "lsmhip_config.h"

#define LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2 lsm3dcomputemeancurvatureorder2_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_curvature3d.h:68 */
void LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi,
                                         LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                         const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                         LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(kappa_fb),
                                         const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

LSMHIP_END_DECLS

#endif
