/*
 * lsm_curvature3d_local.h -- narrow-band mean and Gaussian curvatures on the GPU
 * (reference: src/geometry/lsm_curvature3d_local.h).
 */
#ifndef LSMHIP_CURVATURE_3D_LOCAL_H
#define LSMHIP_CURVATURE_3D_LOCAL_H

#include "lsm_spatial_derivatives3d_local.h"

#define LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2_LOCAL lsm3dcomputemeancurvatureorder2local_
#define LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2_LOCAL lsm3dcomputegaussiancurvatureorder2local_
#define LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4_LOCAL lsm3dcomputemeancurvatureorder4local_
#define LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4_LOCAL lsm3dcomputegaussiancurvatureorder4local_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_curvature3d_local.h:72 */
void LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2_LOCAL(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi,
                                               LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x,
                                               const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                               const LSMLIB_REAL *grad_phi_mag, LSMHIP_BOX_ARGS(grad_phi_gb),
                                               LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);
/* replaces src/geometry/lsm_curvature3d_local.h:141 */
void LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2_LOCAL(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb),
                                                   const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                                   const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                                   const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                                   LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);
/* replaces src/geometry/lsm_curvature3d_local.h:210.  As the reference's Fortran
 * (src/geometry/lsm_curvature3d_local.f.in:285-426) it computes the Gaussian
 * curvature of the next routine, not a mean curvature. */
void LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4_LOCAL(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi,
                                               LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x,
                                               const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                               const LSMLIB_REAL *grad_phi_mag, LSMHIP_BOX_ARGS(grad_phi_gb),
                                               LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);
/* replaces src/geometry/lsm_curvature3d_local.h:279 */
void LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4_LOCAL(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb),
                                                   const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                                   const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                                   const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                                   LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);

LSMHIP_END_DECLS

#endif
