/*
 * lsm_curvature2d_local.h -- 2D narrow-band mean curvature (orders 2 and 4) on the GPU
 * (reference: src/geometry/lsm_curvature2d_local.h).
 */
#ifndef LSMHIP_LSM_CURVATURE2D_LOCAL_H
#define LSMHIP_LSM_CURVATURE2D_LOCAL_H

#include "lsmhip_config.h"

#define LSM2D_COMPUTE_MEAN_CURVATURE_ORDER2_LOCAL lsm2dcomputemeancurvatureorder2local_
#define LSM2D_COMPUTE_MEAN_CURVATURE_ORDER4_LOCAL lsm2dcomputemeancurvatureorder4local_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_curvature2d_local.h:64 */
void LSM2D_COMPUTE_MEAN_CURVATURE_ORDER2_LOCAL(LSMLIB_REAL *kappa, LSMHIP_BOX2_ARGS(kappa_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
    const LSMLIB_REAL *grad_phi_mag, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/geometry/lsm_curvature2d_local.h:207 */
void LSM2D_COMPUTE_MEAN_CURVATURE_ORDER4_LOCAL(LSMLIB_REAL *kappa, LSMHIP_BOX2_ARGS(kappa_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
    const LSMLIB_REAL *grad_phi_mag, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);

LSMHIP_END_DECLS

#endif
