/*
 * lsm_geometry2d.h -- 2D unit normals, mean curvature from a Hessian, smoothed area and perimeter on the GPU
 * (reference: src/geometry/lsm_geometry2d.h).
 */
#ifndef LSMHIP_LSM_GEOMETRY2D_H
#define LSMHIP_LSM_GEOMETRY2D_H

#include "lsmhip_config.h"

#define LSM2D_COMPUTE_UNIT_NORMAL lsm2dcomputeunitnormal_
#define LSM2D_COMPUTE_SIGNED_UNIT_NORMAL lsm2dcomputesignedunitnormal_
#define LSM2D_COMPUTE_MEAN_CURVATURE lsm2dcomputemeancurvature_
#define LSM2D_AREA_REGION_PHI_LESS_THAN_ZERO lsm2darearegionphilessthanzero_
#define LSM2D_AREA_REGION_PHI_GREATER_THAN_ZERO lsm2darearegionphigreaterthanzero_
#define LSM2D_PERIMETER_ZERO_LEVEL_SET lsm2dperimeterzerolevelset_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_geometry2d.h:71 */
void LSM2D_COMPUTE_UNIT_NORMAL(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y, LSMHIP_BOX2_ARGS(normal_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), LSMHIP_BOX2_ARGS(fb));
/* replaces src/geometry/lsm_geometry2d.h:116 */
void LSM2D_COMPUTE_SIGNED_UNIT_NORMAL(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y, LSMHIP_BOX2_ARGS(normal_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *phi,
    LSMHIP_BOX2_ARGS(phi_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);
/* replaces src/geometry/lsm_geometry2d.h:166 */
void LSM2D_COMPUTE_MEAN_CURVATURE(LSMLIB_REAL *kappa, LSMHIP_BOX2_ARGS(kappa_gb), const LSMLIB_REAL *phi_x,
    const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *phi_xx, const LSMLIB_REAL *phi_xy,
    const LSMLIB_REAL *phi_yy, LSMHIP_BOX2_ARGS(hessian_phi_gb), LSMHIP_BOX2_ARGS(fb));
/* replaces src/geometry/lsm_geometry2d.h:206 */
void LSM2D_AREA_REGION_PHI_LESS_THAN_ZERO(LSMLIB_REAL *area, const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb),
    LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *epsilon);
/* replaces src/geometry/lsm_geometry2d.h:238 */
void LSM2D_AREA_REGION_PHI_GREATER_THAN_ZERO(LSMLIB_REAL *area, const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb),
    LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *epsilon);
/* replaces src/geometry/lsm_geometry2d.h:272 */
void LSM2D_PERIMETER_ZERO_LEVEL_SET(LSMLIB_REAL *perimeter, const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), LSMHIP_BOX2_ARGS(ib),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *epsilon);

LSMHIP_END_DECLS

#endif
