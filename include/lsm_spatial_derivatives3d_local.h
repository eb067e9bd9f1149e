/*
 * lsm_spatial_derivatives3d_local.h -- narrow-band derivatives on the GPU
 * (reference: src/toolbox/lsm_spatial_derivatives3d_local.h).  Index arrays
 * and narrow_band are device arrays; index_x points at the list's element
 * nlo_index0 (the Fortran declaration index_x(nlo_index0:...)).  D1/D2 are
 * written over the listed points as in the reference.  Points whose stencil
 * would leave an array's ghost box are skipped and reported through
 * lsmhip_get_last_error() (LSMHIP_ERR_INDEX).
 */
#ifndef LSMHIP_SPATIAL_DERIVATIVES_3D_LOCAL_H
#define LSMHIP_SPATIAL_DERIVATIVES_3D_LOCAL_H

#include "lsmhip_config.h"

#define LSM3D_HJ_ENO1_LOCAL lsm3dhjeno1local_
#define LSM3D_HJ_ENO2_LOCAL lsm3dhjeno2local_
#define LSM3D_HJ_WENO5_LOCAL lsm3dhjweno5local_
#define LSM3D_CENTRAL_GRAD_ORDER2_LOCAL lsm3dcentralgradorder2local_
#define LSM3D_CENTRAL_GRAD_ORDER4_LOCAL lsm3dcentralgradorder4local_
#define LSM3D_LAPLACIAN_ORDER2_LOCAL lsm3dlaplacianorder2local_
#define LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL lsm3dcomputeavegradphilocal_
#define LSM3D_GRADIENT_MAGNITUDE_LOCAL lsm3dgradientmagnitudelocal_

LSMHIP_BEGIN_DECLS

#define LSMHIP_LOCAL_PM                                                                                       \
  LSMLIB_REAL *phi_x_plus, LSMLIB_REAL *phi_y_plus, LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), \
      LSMLIB_REAL *phi_x_minus, LSMLIB_REAL *phi_y_minus, LSMLIB_REAL *phi_z_minus,                           \
      LSMHIP_BOX_ARGS(grad_phi_minus_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb)
#define LSMHIP_LOCAL_TAIL                                                                                    \
  const int *index_x, const int *index_y, const int *index_z, const int *nlo_index, const int *nhi_index,   \
      const unsigned char *narrow_band, LSMHIP_BOX_ARGS(nb_gb), const unsigned char *mark_fb
#define LSMHIP_LDXYZ const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz

/* replaces src/toolbox/lsm_spatial_derivatives3d_local.h:78 */
void LSM3D_HJ_ENO1_LOCAL(LSMHIP_LOCAL_PM, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMHIP_LDXYZ, const int *index_x,
                         const int *index_y, const int *index_z, const int *nlo_index0, const int *nhi_index0,
                         const int *nlo_index1, const int *nhi_index1, const unsigned char *narrow_band,
                         LSMHIP_BOX_ARGS(nb_gb), const unsigned char *mark_fb, const unsigned char *mark_D1);
/* replaces :159 */
void LSM3D_HJ_ENO2_LOCAL(LSMHIP_LOCAL_PM, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMLIB_REAL *D2,
                         LSMHIP_BOX_ARGS(D2_gb), LSMHIP_LDXYZ, const int *index_x, const int *index_y,
                         const int *index_z, const int *nlo_index0, const int *nhi_index0, const int *nlo_index1,
                         const int *nhi_index1, const int *nlo_index2, const int *nhi_index2,
                         const unsigned char *narrow_band, LSMHIP_BOX_ARGS(nb_gb), const unsigned char *mark_fb,
                         const unsigned char *mark_D1, const unsigned char *mark_D2);
/* No counterpart in the reference, which has no 3D WENO5 narrow-band routine:
 * the 3D form of lsm2dHJWENO5LOCAL (src/toolbox/lsm_spatial_derivatives2d_local.h:301),
 * D1 over index levels 0..3 filtered by mark_D1, WENO5 on level 0 filtered by mark_fb. */
void LSM3D_HJ_WENO5_LOCAL(LSMHIP_LOCAL_PM, LSMLIB_REAL *D1, LSMHIP_BOX_ARGS(D1_gb), LSMHIP_LDXYZ, const int *index_x,
                          const int *index_y, const int *index_z, const int *nlo_index0, const int *nhi_index0,
                          const int *nlo_index1, const int *nhi_index1, const int *nlo_index2, const int *nhi_index2,
                          const int *nlo_index3, const int *nhi_index3, const unsigned char *narrow_band,
                          LSMHIP_BOX_ARGS(nb_gb), const unsigned char *mark_fb, const unsigned char *mark_D1);
/* replaces :240 */
void LSM3D_CENTRAL_GRAD_ORDER2_LOCAL(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                                     LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                     LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);
/* replaces :293 */
void LSM3D_CENTRAL_GRAD_ORDER4_LOCAL(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMLIB_REAL *phi_z,
                                     LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                     LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);
/* replaces :345 */
void LSM3D_LAPLACIAN_ORDER2_LOCAL(LSMLIB_REAL *laplacian_phi, LSMHIP_BOX_ARGS(laplacian_phi_gb),
                                  const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb), LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);
/* replaces :397 (FP64 compensated sum; the reference sums serially) */
void LSM3D_COMPUTE_AVE_GRAD_PHI_LOCAL(LSMLIB_REAL *grad_phi_ave, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
                                      LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);
/* replaces :438 (thresholds zero_tol, unlike the whole-grid 1e-8) */
void LSM3D_GRADIENT_MAGNITUDE_LOCAL(const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                    LSMLIB_REAL *grad_phi_mag, LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_LOCAL_TAIL);

LSMHIP_END_DECLS

#endif
