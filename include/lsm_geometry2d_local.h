/*
 * lsm_geometry2d_local.h -- 2D narrow-band unit normals on the GPU
 * (reference: src/geometry/lsm_geometry2d_local.h).
 */
#ifndef LSMHIP_LSM_GEOMETRY2D_LOCAL_H
#define LSMHIP_LSM_GEOMETRY2D_LOCAL_H

#include "lsmhip_config.h"

#define LSM2D_COMPUTE_UNIT_NORMAL_LOCAL lsm2dcomputeunitnormallocal_
#define LSM2D_COMPUTE_SIGNED_UNIT_NORMAL_LOCAL lsm2dcomputesignedunitnormallocal_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_geometry2d_local.h:69 */
void LSM2D_COMPUTE_UNIT_NORMAL_LOCAL(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y, LSMHIP_BOX2_ARGS(normal_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/geometry/lsm_geometry2d_local.h:133 */
void LSM2D_COMPUTE_SIGNED_UNIT_NORMAL_LOCAL(LSMLIB_REAL *normal_x, LSMLIB_REAL *normal_y,
    LSMHIP_BOX2_ARGS(normal_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);

LSMHIP_END_DECLS

#endif
