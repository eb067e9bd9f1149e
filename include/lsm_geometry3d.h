/*
 * lsm_geometry3d.h -- unit normals, mean curvature from a given Hessian,
 * smoothed volume and surface area, over a box or a control volume, on the GPU
 * (reference: src/geometry/lsm_geometry3d.h).  Volumes and areas are FP64
 * compensated sums.
 */
#ifndef LSMHIP_GEOMETRY_3D_H
#define LSMHIP_GEOMETRY_3D_H

#include "lsmhip_config.h"

#define LSM3D_COMPUTE_UNIT_NORMAL lsm3dcomputeunitnormal_
#define LSM3D_COMPUTE_SIGNED_UNIT_NORMAL lsm3dcomputesignedunitnormal_
#define LSM3D_COMPUTE_MEAN_CURVATURE lsm3dcomputemeancurvature_
#define LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO lsm3dvolumeregionphilessthanzero_
#define LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO lsm3dvolumeregionphigreaterthanzero_
#define LSM3D_SURFACE_AREA_ZERO_LEVEL_SET lsm3dsurfaceareazerolevelset_
#define LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA lsm3dsurfaceareazerolevelsetdelta_
#define LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO_CONTROL_VOLUME lsm3dvolumeregionphilessthanzerocontrolvolume_
#define LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME lsm3dvolumeregionphigreaterthanzerocontrolvolume_
#define LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_CONTROL_VOLUME lsm3dsurfaceareazerolevelsetcontrolvolume_
#define LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA_CONTROL_VOLUME lsm3dsurfaceareazerolevelsetdeltacontrolvolume_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_geometry3d.h:76 */
void LSM3D_COMPUTE_UNIT_NORMAL(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y, LSMLIB_REAL *normal_z,
                               LSMHIP_BOX_ARGS(normal_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                               const LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(fb));
/* replaces :129 */
void LSM3D_COMPUTE_SIGNED_UNIT_NORMAL(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y, LSMLIB_REAL *normal_z,
                                      LSMHIP_BOX_ARGS(normal_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                      const LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *phi,
                                      LSMHIP_BOX_ARGS(phi_gb), LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dx,
                                      const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);
/* replaces :192.  FP64: |grad phi|^3 is the correctly rounded (|grad phi|^2)^1.5. */
void LSM3D_COMPUTE_MEAN_CURVATURE(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi_x,
                                  const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb),
                                  const LSMLIB_REAL *phi_xx, const LSMLIB_REAL *phi_xy, const LSMLIB_REAL *phi_xz,
                                  const LSMLIB_REAL *phi_yy, const LSMLIB_REAL *phi_yz, const LSMLIB_REAL *phi_zz,
                                  LSMHIP_BOX_ARGS(hessian_phi_gb), LSMHIP_BOX_ARGS(fb));
/* replaces :244 */
void LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO(LSMLIB_REAL *volume, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                            LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                            const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* replaces :281 */
void LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO(LSMLIB_REAL *volume, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                               LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
                                               const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* replaces :320 */
void LSM3D_SURFACE_AREA_ZERO_LEVEL_SET(LSMLIB_REAL *area, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                       const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                       LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx,
                                       const LSMLIB_REAL *dy, const LSMLIB_REAL *dz, const LSMLIB_REAL *epsilon);
/* replaces :365: the area from a given delta function of phi (cells where it
 * is positive) */
void LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA(LSMLIB_REAL *surface_area, const LSMLIB_REAL *delta_phi,
                                             LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x,
                                             const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                             LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx,
                                             const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

/* Control-volume forms: a cell takes part where control_vol > 0
 * (control_vol_sgn > 0) or control_vol < 0 (otherwise). */
#define LSMHIP_GEOM_CV_ARGS \
  const LSMLIB_REAL *control_vol, LSMHIP_BOX_ARGS(control_vol_gb), const int *control_vol_sgn
/* replaces :415 */
void LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO_CONTROL_VOLUME(LSMLIB_REAL *volume, const LSMLIB_REAL *phi,
                                                           LSMHIP_BOX_ARGS(phi_gb), LSMHIP_GEOM_CV_ARGS,
                                                           LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx,
                                                           const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                                                           const LSMLIB_REAL *epsilon);
/* replaces :466 */
void LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO_CONTROL_VOLUME(LSMLIB_REAL *volume, const LSMLIB_REAL *phi,
                                                              LSMHIP_BOX_ARGS(phi_gb), LSMHIP_GEOM_CV_ARGS,
                                                              LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx,
                                                              const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                                                              const LSMLIB_REAL *epsilon);
/* replaces :519 */
void LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_CONTROL_VOLUME(LSMLIB_REAL *surface_area, const LSMLIB_REAL *phi,
                                                      LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x,
                                                      const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                                      LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_GEOM_CV_ARGS,
                                                      LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx,
                                                      const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
                                                      const LSMLIB_REAL *epsilon);
/* replaces :578 */
void LSM3D_SURFACE_AREA_ZERO_LEVEL_SET_DELTA_CONTROL_VOLUME(LSMLIB_REAL *surface_area, const LSMLIB_REAL *delta_phi,
                                                            LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x,
                                                            const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                                            LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_GEOM_CV_ARGS,
                                                            LSMHIP_BOX_ARGS(ib), const LSMLIB_REAL *dx,
                                                            const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

LSMHIP_END_DECLS

#endif
