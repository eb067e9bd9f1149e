/*
 * lsm_geometry3d.h -- smoothed volume and surface area on the GPU
 * (reference: src/geometry/lsm_geometry3d.h; the other geometry routines of
 * that header are outside this path).  FP64 compensated sums.
 */
#ifndef LSMHIP_GEOMETRY_3D_H
#define LSMHIP_GEOMETRY_3D_H

#include "lsmhip_config.h"

#define LSM3D_VOLUME_REGION_PHI_LESS_THAN_ZERO lsm3dvolumeregionphilessthanzero_
#define LSM3D_VOLUME_REGION_PHI_GREATER_THAN_ZERO lsm3dvolumeregionphigreaterthanzero_
#define LSM3D_SURFACE_AREA_ZERO_LEVEL_SET lsm3dsurfaceareazerolevelset_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_geometry3d.h:244 */
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

LSMHIP_END_DECLS

#endif
