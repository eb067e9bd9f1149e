/*
 * lsm_spatial_derivatives2d_local.h -- 2D narrow-band HJ-ENO/WENO derivatives, central gradients, Laplacian, gradient magnitude on the GPU
 * (reference: src/toolbox/lsm_spatial_derivatives2d_local.h).
 */
#ifndef LSMHIP_LSM_SPATIAL_DERIVATIVES2D_LOCAL_H
#define LSMHIP_LSM_SPATIAL_DERIVATIVES2D_LOCAL_H

#include "lsmhip_config.h"

#define LSM2D_HJ_ENO1_LOCAL lsm2dhjeno1local_
#define LSM2D_HJ_ENO2_LOCAL lsm2dhjeno2local_
#define LSM2D_HJ_ENO3_LOCAL lsm2dhjeno3local_
#define LSM2D_HJ_WENO5_LOCAL lsm2dhjweno5local_
#define LSM2D_UPWIND_HJ_ENO2_LOCAL lsm2dupwindhjeno2local_
#define LSM2D_CENTRAL_GRAD_ORDER2_LOCAL lsm2dcentralgradorder2local_
#define LSM2D_CENTRAL_GRAD_ORDER4_LOCAL lsm2dcentralgradorder4local_
#define LSM2D_LAPLACIAN_ORDER2_LOCAL lsm2dlaplacianorder2local_
#define LSM2D_COMPUTE_AVE_GRAD_PHI_LOCAL lsm2dcomputeavegradphilocal_
#define LSM2D_GRADIENT_MAGNITUDE_LOCAL lsm2dgradientmagnitudelocal_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:77 */
void LSM2D_HJ_ENO1_LOCAL(LSMLIB_REAL  *phi_x_plus, LSMLIB_REAL  *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL  *phi_x_minus, LSMLIB_REAL  *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL  *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int *index_x, const int *index_y, const int *nlo_index0,
    const int *nhi_index0, const int *nlo_index1, const int *nhi_index1, const unsigned char *narrow_band,
    LSMHIP_BOX2_ARGS(nb_gb), const unsigned char *mark_fb, const unsigned char *mark_D1);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:145 */
void LSM2D_HJ_ENO2_LOCAL(LSMLIB_REAL  *phi_x_plus, LSMLIB_REAL  *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL  *phi_x_minus, LSMLIB_REAL  *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL  *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMLIB_REAL *D2,
    LSMHIP_BOX2_ARGS(D2_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int *index_x, const int *index_y,
    const int *nlo_index0, const int *nhi_index0, const int *nlo_index1, const int *nhi_index1,
    const int *nlo_index2, const int *nhi_index2, const unsigned char *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char *mark_fb, const unsigned char *mark_D1, const unsigned char *mark_D2);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:220 */
void LSM2D_HJ_ENO3_LOCAL(LSMLIB_REAL  *phi_x_plus, LSMLIB_REAL  *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL  *phi_x_minus, LSMLIB_REAL  *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL  *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMLIB_REAL *D2,
    LSMHIP_BOX2_ARGS(D2_gb), LSMLIB_REAL *D3, LSMHIP_BOX2_ARGS(D3_gb), const LSMLIB_REAL *dx,
    const LSMLIB_REAL *dy, const int *index_x, const int *index_y, const int *nlo_index0, const int *nhi_index0,
    const int *nlo_index1, const int *nhi_index1, const int *nlo_index2, const int *nhi_index2,
    const int *nlo_index3, const int *nhi_index3, const unsigned char *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char *mark_fb, const unsigned char *mark_D1, const unsigned char *mark_D2,
    const unsigned char *mark_D3);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:301 */
void LSM2D_HJ_WENO5_LOCAL(LSMLIB_REAL  *phi_x_plus, LSMLIB_REAL  *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    LSMLIB_REAL  *phi_x_minus, LSMLIB_REAL  *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL  *phi, LSMHIP_BOX2_ARGS(phi_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int *index_x, const int *index_y, const int *nlo_index0,
    const int *nhi_index0, const int *nlo_index1, const int *nhi_index1, const int *nlo_index2,
    const int *nhi_index2, const int *nlo_index3, const int *nhi_index3, const unsigned char *narrow_band,
    LSMHIP_BOX2_ARGS(nb_gb), const unsigned char *mark_fb, const unsigned char *mark_D1);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:361 */
void LSM2D_UPWIND_HJ_ENO2_LOCAL(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    LSMHIP_BOX2_ARGS(vel_gb), LSMLIB_REAL *D1, LSMHIP_BOX2_ARGS(D1_gb), LSMLIB_REAL *D2, LSMHIP_BOX2_ARGS(D2_gb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int *index_x, const int *index_y, const int *nlo_index0,
    const int *nhi_index0, const int *nlo_index1, const int *nhi_index1, const int *nlo_index2,
    const int *nhi_index2, const unsigned char *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char *mark_fb,
    const unsigned char *mark_D1, const unsigned char *mark_D2);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:426 */
void LSM2D_CENTRAL_GRAD_ORDER2_LOCAL(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:469 */
void LSM2D_CENTRAL_GRAD_ORDER4_LOCAL(LSMLIB_REAL *phi_x, LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:512 */
void LSM2D_LAPLACIAN_ORDER2_LOCAL(LSMLIB_REAL *laplacian_phi, LSMHIP_BOX2_ARGS(laplacian_phi_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:555 */
void LSM2D_COMPUTE_AVE_GRAD_PHI_LOCAL(LSMLIB_REAL *grad_phi_ave, const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int  *index_x, const int  *index_y, const int  *nlo_index,
    const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_spatial_derivatives2d_local.h:591 */
void LSM2D_GRADIENT_MAGNITUDE_LOCAL(const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMLIB_REAL *grad_phi_mag,
    LSMHIP_BOX2_ARGS(grad_phi_gb), const int  *index_x, const int  *index_y, const int  *nlo_index,
    const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);

LSMHIP_END_DECLS

#endif
