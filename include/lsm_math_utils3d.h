/*
 * lsm_math_utils3d.h -- CFL time steps, max-norm and average differences,
 * voxel counts and smoothed integrals on the GPU, over a box or over the cells
 * of a control volume (reference: src/toolbox/lsm_math_utils3d.h).  The scalar
 * result is written to host memory when the call returns.  Maxima and counts
 * equal the reference's bitwise; sums are FP64 compensated sums in a fixed
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
#define LSM3D_AVE_ABS_DIFF lsm3daveabsdiff_
#define LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO lsm3dvoxelcountgreaterthanzero_
#define LSM3D_VOXEL_COUNT_LESS_THAN_ZERO lsm3dvoxelcountlessthanzero_
#define LSM3D_MAX_NORM_DIFF_CONTROL_VOLUME lsm3dmaxnormdiffcontrolvolume_
#define LSM3D_COMPUTE_STABLE_ADVECTION_DT_CONTROL_VOLUME lsm3dcomputestableadvectiondtcontrolvolume_
#define LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_CONTROL_VOLUME lsm3dcomputestablenormalveldtcontrolvolume_
#define LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_CONTROL_VOLUME lsm3dcomputestableconstnormalveldtcontrolvolume_
#define LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO_CONTROL_VOLUME lsm3dvolumeintegralphilessthanzerocontrolvolume_
#define LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME lsm3dvolumeintegralphigreaterthanzerocontrolvolume_
#define LSM3D_SURFACE_INTEGRAL_CONTROL_VOLUME lsm3dsurfaceintegralcontrolvolume_
#define LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO_CONTROL_VOLUME lsm3dvoxelcountgreaterthanzerocontrolvolume_
#define LSM3D_VOXEL_COUNT_LESS_THAN_ZERO_CONTROL_VOLUME lsm3dvoxelcountlessthanzerocontrolvolume_

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
/* replaces :124.  The count is exact in both precisions (the reference's FP32
 * build counts in FP32, which stops at 2^24 cells). */
void LSM3D_AVE_ABS_DIFF(LSMLIB_REAL *ave_abs_diff, const LSMLIB_REAL *field1, LSMHIP_BOX_ARGS(field1_gb),
                        const LSMLIB_REAL *field2, LSMHIP_BOX_ARGS(field2_gb), LSMHIP_BOX_ARGS(ib));
/* replaces :160 */
void LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO(int *count, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb),
                                         LSMHIP_BOX_ARGS(fb));
/* replaces :190 */
void LSM3D_VOXEL_COUNT_LESS_THAN_ZERO(int *count, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb));

/* Control-volume forms: a cell takes part where control_vol > 0
 * (control_vol_sgn > 0) or control_vol < 0 (otherwise); the voxel counts, as
 * the reference's, take control_vol >= 0 and <= 0. */
#define LSMHIP_CV_ARGS \
  const LSMLIB_REAL *control_vol, LSMHIP_BOX_ARGS(control_vol_gb), const int *control_vol_sgn
/* replaces :533.  As the reference, the cell at the lower corner of ib always
 * takes part. */
void LSM3D_MAX_NORM_DIFF_CONTROL_VOLUME(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1,
                                        LSMHIP_BOX_ARGS(field1_gb), const LSMLIB_REAL *field2,
                                        LSMHIP_BOX_ARGS(field2_gb), LSMHIP_CV_ARGS, LSMHIP_BOX_ARGS(ib));
/* replaces :585 */
void LSM3D_COMPUTE_STABLE_ADVECTION_DT_CONTROL_VOLUME(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x,
                                                      const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z,
                                                      LSMHIP_BOX_ARGS(vel_gb), LSMHIP_CV_ARGS, LSMHIP_BOX_ARGS(ib),
                                                      const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                                      const LSMLIB_REAL *dz, const LSMLIB_REAL *cfl_number);
/* replaces :646 */
void LSM3D_COMPUTE_STABLE_NORMAL_VEL_DT_CONTROL_VOLUME(
    LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, LSMHIP_BOX_ARGS(vel_gb), const LSMLIB_REAL *phi_x_plus,
    const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus,
    LSMHIP_BOX_ARGS(grad_phi_minus_gb), LSMHIP_CV_ARGS, LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const LSMLIB_REAL *cfl_number);
/* replaces :723 */
void LSM3D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_CONTROL_VOLUME(
    LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
    const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus,
    const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb),
    LSMHIP_CV_ARGS, LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const LSMLIB_REAL *cfl_number);
/* replaces :789 */
void LSM3D_VOLUME_INTEGRAL_PHI_LESS_THAN_ZERO_CONTROL_VOLUME(LSMLIB_REAL *int_F, const LSMLIB_REAL *F,
                                                             LSMHIP_BOX_ARGS(F_gb), const LSMLIB_REAL *phi,
                                                             LSMHIP_BOX_ARGS(phi_gb), LSMHIP_CV_ARGS,
                                                             LSMHIP_BOX_ARGS(ib), LSMHIP_DXYZ_EPS);
/* replaces :850 */
void LSM3D_VOLUME_INTEGRAL_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME(LSMLIB_REAL *int_F, const LSMLIB_REAL *F,
                                                                LSMHIP_BOX_ARGS(F_gb), const LSMLIB_REAL *phi,
                                                                LSMHIP_BOX_ARGS(phi_gb), LSMHIP_CV_ARGS,
                                                                LSMHIP_BOX_ARGS(ib), LSMHIP_DXYZ_EPS);
/* replaces :911 */
void LSM3D_SURFACE_INTEGRAL_CONTROL_VOLUME(LSMLIB_REAL *int_F, const LSMLIB_REAL *F, LSMHIP_BOX_ARGS(F_gb),
                                           const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x,
                                           const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                           LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_CV_ARGS, LSMHIP_BOX_ARGS(ib),
                                           LSMHIP_DXYZ_EPS);
/* replaces :969 */
void LSM3D_VOXEL_COUNT_GREATER_THAN_ZERO_CONTROL_VOLUME(int *count, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb),
                                                        LSMHIP_CV_ARGS, LSMHIP_BOX_ARGS(fb));
/* replaces :1006 */
void LSM3D_VOXEL_COUNT_LESS_THAN_ZERO_CONTROL_VOLUME(int *count, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb),
                                                     LSMHIP_CV_ARGS, LSMHIP_BOX_ARGS(fb));

LSMHIP_END_DECLS

#endif
