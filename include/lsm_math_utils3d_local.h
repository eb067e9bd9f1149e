/*
 * lsm_math_utils3d_local.h -- narrow-band CFL steps and max-norm difference
 * on the GPU (reference: src/toolbox/lsm_math_utils3d_local.h).
 */
#ifndef LSMHIP_MATH_UTILS_3D_LOCAL_H
#define LSMHIP_MATH_UTILS_3D_LOCAL_H

#include "lsm_spatial_derivatives3d_local.h"

#define LSM3D_MAX_NORM_DIFF_LOCAL lsm3dmaxnormdifflocal_
#define LSM3D_COMPUTE_STABLE_ADVECTION_DT_LOCAL lsm3dcomputestableadvectiondtlocal_
#define LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL lsm3dcomputestablenormalveldtlocal_
#define LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL lsm3dcomputestableconstnormalveldtlocal_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_math_utils3d_local.h:65 */
void LSM3D_MAX_NORM_DIFF_LOCAL(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1, LSMHIP_BOX_ARGS(field1_gb),
                               const LSMLIB_REAL *field2, LSMHIP_BOX_ARGS(field2_gb), LSMHIP_LOCAL_TAIL);
/* replaces :112 */
void LSM3D_COMPUTE_STABLE_ADVECTION_DT_LOCAL(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                             const LSMLIB_REAL *vel_z, LSMHIP_BOX_ARGS(vel_gb), LSMHIP_LDXYZ,
                                             const LSMLIB_REAL *cfl_number, LSMHIP_LOCAL_TAIL);
/* replaces :169 */
void LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, LSMHIP_BOX_ARGS(vel_gb),
                                              const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
                                              const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb),
                                              const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
                                              const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb),
                                              LSMHIP_LDXYZ, const LSMLIB_REAL *cfl_number, LSMHIP_LOCAL_TAIL);
/* replaces :322.  The reference divides by two locals it never assigns
 * (lsm_math_utils3d_local.f.in:628-751); this computes the whole-grid
 * routine's formula over the admissible band points. */
void LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL(
    LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
    const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus,
    const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb), LSMHIP_LDXYZ,
    const LSMLIB_REAL *cfl_number, LSMHIP_LOCAL_TAIL);

LSMHIP_END_DECLS

#endif
