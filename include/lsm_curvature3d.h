/*
 * lsm_curvature3d.h -- mean and Gaussian curvature (second and fourth order) on the GPU
 * (reference: src/geometry/lsm_curvature3d.h).
 */
#ifndef LSMHIP_CURVATURE_3D_H
#define LSMHIP_CURVATURE_3D_H

#include "lsmhip_config.h"

#define LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2 lsm3dcomputemeancurvatureorder2_
#define LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2 lsm3dcomputegaussiancurvatureorder2_
#define LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4 lsm3dcomputemeancurvatureorder4_
#define LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4 lsm3dcomputegaussiancurvatureorder4_

LSMHIP_BEGIN_DECLS

/* replaces src/geometry/lsm_curvature3d.h:68 */
void LSM3D_COMPUTE_MEAN_CURVATURE_ORDER2(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi,
                                         LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                         const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                         LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(kappa_fb),
                                         const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

/* replaces src/geometry/lsm_curvature3d.h:128 */
void LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER2(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi,
                                         LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                         const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                         LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(kappa_fb),
                                         const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

/* replaces src/geometry/lsm_curvature3d.h:183.  phi and the gradient are read
   two cells away along each axis (the mixed derivatives are fourth-order
   differences of phi_x along y, phi_y along z and phi_z along x). */
void LSM3D_COMPUTE_MEAN_CURVATURE_ORDER4(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi,
                                         LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                         const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                         LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(kappa_fb),
                                         const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

/* replaces src/geometry/lsm_curvature3d.h:244 (stencil of the order-4 mean curvature) */
void LSM3D_COMPUTE_GAUSSIAN_CURVATURE_ORDER4(LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb), const LSMLIB_REAL *phi,
                                         LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                         const LSMLIB_REAL *phi_z, const LSMLIB_REAL *grad_phi_mag,
                                         LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(kappa_fb),
                                         const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

LSMHIP_END_DECLS

#endif
