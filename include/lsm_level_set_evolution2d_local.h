/*
 * lsm_level_set_evolution2d_local.h -- 2D narrow-band level-set equation RHS terms on the GPU
 * (reference: src/toolbox/lsm_level_set_evolution2d_local.h).
 */
#ifndef LSMHIP_LSM_LEVEL_SET_EVOLUTION2D_LOCAL_H
#define LSMHIP_LSM_LEVEL_SET_EVOLUTION2D_LOCAL_H

#include "lsmhip_config.h"

#define LSM2D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL lsm2dzerooutlevelseteqnrhslocal_
#define LSM2D_ADD_ADVECTION_TERM_TO_LSE_RHS_LOCAL lsm2daddadvectiontermtolserhslocal_
#define LSM2D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL lsm2daddnormalveltermtolserhslocal_
#define LSM2D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL lsm2daddconstnormalveltermtolserhslocal_
#define LSM2D_ADD_CONST_CURV_TERM_TO_LSE_RHS_LOCAL lsm2daddconstcurvtermtolserhslocal_
#define LSM2D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS_LOCAL lsm2daddconstprecomputedcurvtermtolserhslocal_
#define LSM2D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL lsm2daddexternalandnormalveltermtolserhslocal_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_level_set_evolution2d_local.h:70 */
void LSM2D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb), const int *index_x,
    const int *index_y, const int *nlo_index, const int *nhi_index);
/* replaces src/toolbox/lsm_level_set_evolution2d_local.h:106 */
void LSM2D_ADD_ADVECTION_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *vel_x,
    const LSMLIB_REAL *vel_y, LSMHIP_BOX2_ARGS(vel_gb), const int  *index_x, const int  *index_y,
    const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_level_set_evolution2d_local.h:163 */
void LSM2D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *vel_n, LSMHIP_BOX2_ARGS(vel_gb), const int  *index_x, const int  *index_y,
    const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_level_set_evolution2d_local.h:226 */
void LSM2D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *vel_n, const int  *index_x, const int  *index_y, const int  *nlo_index,
    const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_level_set_evolution2d_local.h:289 */
void LSM2D_ADD_CONST_CURV_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *phi_xx,
    const LSMLIB_REAL *phi_xy, const LSMLIB_REAL *phi_yy, LSMHIP_BOX2_ARGS(grad2_phi_gb), const LSMLIB_REAL *b,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_level_set_evolution2d_local.h:349 (declared there as LSM2D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSERHS_LOCAL) */
void LSM2D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *kappa, LSMHIP_BOX2_ARGS(kappa_gb), const LSMLIB_REAL *grad_mag_phi,
    LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *b, const int  *index_x, const int  *index_y,
    const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_level_set_evolution2d_local.h:414 */
void LSM2D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *vel_n, const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y, LSMHIP_BOX2_ARGS(vel_gb),
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);

LSMHIP_END_DECLS

#endif
