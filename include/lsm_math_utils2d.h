/*
 * lsm_math_utils2d.h -- 2D CFL time steps, max-norm difference, smoothed integrals on the GPU
 * (reference: src/toolbox/lsm_math_utils2d.h).
 */
#ifndef LSMHIP_LSM_MATH_UTILS2D_H
#define LSMHIP_LSM_MATH_UTILS2D_H

#include "lsmhip_config.h"

#define LSM2D_MAX_NORM_DIFF lsm2dmaxnormdiff_
#define LSM2D_COMPUTE_STABLE_ADVECTION_DT lsm2dcomputestableadvectiondt_
#define LSM2D_COMPUTE_STABLE_NORMAL_VEL_DT lsm2dcomputestablenormalveldt_
#define LSM2D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT lsm2dcomputestableconstnormalveldt_
#define LSM2D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO lsm2dvolumeintegralphilessthanzero_
#define LSM2D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO lsm2dvolumeintegralphigreaterthanzero_
#define LSM2D_SURFACE_INTEGRAL lsm2dsurfaceintegral_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_math_utils2d.h:90 */
void LSM2D_MAX_NORM_DIFF(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1, LSMHIP_BOX2_ARGS(field1_gb),
    const LSMLIB_REAL *field2, LSMHIP_BOX2_ARGS(field2_gb), LSMHIP_BOX2_ARGS(ib));
/* replaces src/toolbox/lsm_math_utils2d.h:204 */
void LSM2D_COMPUTE_STABLE_ADVECTION_DT(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    LSMHIP_BOX2_ARGS(vel_gb), LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const LSMLIB_REAL *cfl_number);
/* replaces src/toolbox/lsm_math_utils2d.h:245 */
void LSM2D_COMPUTE_STABLE_NORMAL_VEL_DT(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, LSMHIP_BOX2_ARGS(vel_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *cfl_number);
/* replaces src/toolbox/lsm_math_utils2d.h:297 */
void LSM2D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *cfl_number);
/* replaces src/toolbox/lsm_math_utils2d.h:340 */
void LSM2D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSMHIP_BOX2_ARGS(F_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *epsilon);
/* replaces src/toolbox/lsm_math_utils2d.h:379 */
void LSM2D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSMHIP_BOX2_ARGS(F_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *epsilon);
/* replaces src/toolbox/lsm_math_utils2d.h:418 */
void LSM2D_SURFACE_INTEGRAL(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSMHIP_BOX2_ARGS(F_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
    LSMHIP_BOX2_ARGS(grad_phi_gb), LSMHIP_BOX2_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const LSMLIB_REAL *epsilon);

LSMHIP_END_DECLS

#endif
