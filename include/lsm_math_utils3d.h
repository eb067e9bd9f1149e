/*
 * lsm_math_utils3d.h -- CFL time steps, max-norm difference and smoothed
 * integrals on the GPU (reference: src/toolbox/lsm_math_utils3d.h).  The
 * scalar result is written to host memory when the call returns.  Maxima equal
 * the reference's bitwise; integrals are FP64 compensated sums in a fixed
 * order (the reference sums serially).
 */
#ifndef LSMHIP_MATH_UTILS_3D_H
#define LSMHIP_MATH_UTILS_3D_H

#include "lsmhip_config.h"

#define LSM3D_MAX_NORM_DIFF lsm3dmaxnormdiff_
#define LSM3D_COMPUTE_STABLE_ADVECTION_DT lsm3dcomputestableadvectiondt_
#define LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT lsm3dcomputestablenormalveldt_
#define LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT lsm3dcomputestableconstnormalveldt_
#define LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO lsm3dvolumeintegralphilessthanzero_
#define LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO lsm3dvolumeintegralphigreaterthanzero_
#define LSM3D_SURFACE_INTEGRAL lsm3dsurfaceintegral_

LSMHIP_BEGIN_DECLS

#define LSMHIP_DXYZ_EPS \
  const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon

/* replaces src/toolbox/lsm_math_utils3d.h:87 */
void LSM3D_MAX_NORM_DIFF(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1, LSMHIP_BOX_ARGS(field1_gb),
                         const LSMLIB_REAL *field2, LSMHIP_BOX_ARGS(field2_gb), LSMHIP_BOX_ARGS(ib));
/* replaces :221 */
void LSM3D_COMPUTE_STABLE_ADVECTION_DT(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                       const LSMLIB_REAL *vel_z, LSMHIP_BOX_ARGS(vel_gb), LSMHIP_BOX_ARGS(ib),
                                       const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                                       const LSMLIB_REAL *cfl_number);
/* replaces :268 */
void LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, LSMHIP_BOX_ARGS(vel_gb),
                                        const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
                                        const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb),
                                        const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
                                        const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb),
                                        LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                        const LSMLIB_REAL *dz, const LSMLIB_REAL *cfl_number);
/* replaces :331 */
void LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
                                              const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
                                              const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb),
                                              const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,
                                              const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb),
                                              LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                              const LSMLIB_REAL *dz, const LSMLIB_REAL *cfl_number);
/* replaces :383 */
void LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSMHIP_BOX_ARGS(F_gb),
                                              const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb), LSMHIP_BOX_ARGS(ib),
                                              LSMHIP_DXYZ_EPS);
/* replaces :430 */
void LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSMHIP_BOX_ARGS(F_gb),
                                                 const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                                 LSMHIP_BOX_ARGS(ib), LSMHIP_DXYZ_EPS);
/* replaces :477 */
void LSM3D_SURFACE_INTEGRAL(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSMHIP_BOX_ARGS(F_gb), const LSMLIB_REAL *phi,
                            LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                            const LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(ib),
                            LSMHIP_DXYZ_EPS);

LSMHIP_END_DECLS

#endif
