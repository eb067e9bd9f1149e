/*
 * lsm_level_set_evolution2d.h -- 2D level-set equation RHS terms on the GPU
 * (reference: src/toolbox/lsm_level_set_evolution2d.h).
 */
#ifndef LSMHIP_LSM_LEVEL_SET_EVOLUTION2D_H
#define LSMHIP_LSM_LEVEL_SET_EVOLUTION2D_H

#include "lsmhip_config.h"

#define LSM2D_ZERO_OUT_LEVEL_SET_EQN_RHS lsm2dzerooutlevelseteqnrhs_
#define LSM2D_ADD_ADVECTION_TERM_TO_LSE_RHS lsm2daddadvectiontermtolserhs_
#define LSM2D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS lsm2daddnormalveltermtolserhs_
#define LSM2D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS lsm2daddconstnormalveltermtolserhs_
#define LSM2D_ADD_CONST_CURV_TERM_TO_LSE_RHS lsm2daddconstcurvtermtolserhs_
#define LSM2D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS lsm2daddconstprecomputedcurvtermtolserhs_
#define LSM2D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS lsm2daddexternalandnormalveltermtolserhs_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_level_set_evolution2d.h:64 */
void LSM2D_ZERO_OUT_LEVEL_SET_EQN_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb));
/* replaces src/toolbox/lsm_level_set_evolution2d.h:92 */
void LSM2D_ADD_ADVECTION_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *vel_x,
    const LSMLIB_REAL *vel_y, LSMHIP_BOX2_ARGS(vel_gb), LSMHIP_BOX2_ARGS(fb));
/* replaces src/toolbox/lsm_level_set_evolution2d.h:139 */
void LSM2D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *vel_n, LSMHIP_BOX2_ARGS(vel_gb), LSMHIP_BOX2_ARGS(fb));
/* replaces src/toolbox/lsm_level_set_evolution2d.h:193 */
void LSM2D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *vel_n, LSMHIP_BOX2_ARGS(fb));
/* replaces src/toolbox/lsm_level_set_evolution2d.h:243 */
void LSM2D_ADD_CONST_CURV_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(lse_rhs_gb),
    const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y, LSMHIP_BOX2_ARGS(grad_phi_gb), const LSMLIB_REAL *phi_xx,
    const LSMLIB_REAL *phi_xy, const LSMLIB_REAL *phi_yy, LSMHIP_BOX2_ARGS(grad2_phi_gb), const LSMLIB_REAL *b,
    LSMHIP_BOX2_ARGS(fb));
/* replaces src/toolbox/lsm_level_set_evolution2d.h:288 */
void LSM2D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *kappa, LSMHIP_BOX2_ARGS(kappa_gb), const LSMLIB_REAL *grad_mag_phi,
    LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *b, LSMHIP_BOX2_ARGS(rhs_fb));
/* replaces src/toolbox/lsm_level_set_evolution2d.h:341 */
void LSM2D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *vel_n, const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y, LSMHIP_BOX2_ARGS(vel_gb),
    LSMHIP_BOX2_ARGS(rhs_fb));

LSMHIP_END_DECLS

#endif
