/*
 * lsm_level_set_evolution3d_local.h -- narrow-band RHS terms on the GPU
 * (reference: src/toolbox/lsm_level_set_evolution3d_local.h).
 */
#ifndef LSMHIP_LEVEL_SET_EVOLUTION_3D_LOCAL_H
#define LSMHIP_LEVEL_SET_EVOLUTION_3D_LOCAL_H

#include "lsm_spatial_derivatives3d_local.h"

#define LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL lsm3dzerooutlevelseteqnrhslocal_
#define LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS_LOCAL lsm3daddadvectiontermtolserhslocal_
#define LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL lsm3daddnormalveltermtolserhslocal_
#define LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL lsm3daddconstnormalveltermtolserhslocal_
#define LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS_LOCAL lsm3daddconstcurvtermtolserhslocal_
#define LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS_LOCAL lsm3daddconstprecomputedcurvtermtolserhslocal_
#define LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL lsm3daddexternalandnormalveltermtolserhslocal_

LSMHIP_BEGIN_DECLS

#define LSMHIP_LGRAD_PM                                                                                      \
  const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus,               \
      LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,     \
      const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb)

/* replaces src/toolbox/lsm_level_set_evolution3d_local.h:67 (no mark filter) */
void LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb), const int *index_x,
                                            const int *index_y, const int *index_z, const int *nlo_index,
                                            const int *nhi_index);
/* replaces :102 */
void LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb),
                                               const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                               const LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb),
                                               const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                               const LSMLIB_REAL *vel_z, LSMHIP_BOX_ARGS(vel_gb), LSMHIP_LOCAL_TAIL);
/* replaces :166 */
void LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb), LSMHIP_LGRAD_PM,
                                                const LSMLIB_REAL *vel_n, LSMHIP_BOX_ARGS(vel_gb), LSMHIP_LOCAL_TAIL);
/* replaces :237 */
void LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb),
                                                      LSMHIP_LGRAD_PM, const LSMLIB_REAL *vel_n, LSMHIP_LOCAL_TAIL);
/* replaces :300 */
void LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb),
                                                const LSMLIB_REAL *phi_x, const LSMLIB_REAL *phi_y,
                                                const LSMLIB_REAL *phi_z, LSMHIP_BOX_ARGS(grad_phi_gb),
                                                const LSMLIB_REAL *phi_xx, const LSMLIB_REAL *phi_xy,
                                                const LSMLIB_REAL *phi_xz, const LSMLIB_REAL *phi_yy,
                                                const LSMLIB_REAL *phi_yz, const LSMLIB_REAL *phi_zz,
                                                LSMHIP_BOX_ARGS(grad2_phi_gb), const LSMLIB_REAL *b,
                                                LSMHIP_LOCAL_TAIL);
/* replaces :374 */
void LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(rhs_gb),
                                                            const LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb),
                                                            const LSMLIB_REAL *grad_mag_phi,
                                                            LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *b,
                                                            LSMHIP_LOCAL_TAIL);
/* replaces :449 */
void LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS_LOCAL(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(rhs_gb),
                                                             LSMHIP_LGRAD_PM, const LSMLIB_REAL *vel_n,
                                                             const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
                                                             const LSMLIB_REAL *vel_z, LSMHIP_BOX_ARGS(vel_gb),
                                                             LSMHIP_LOCAL_TAIL);

LSMHIP_END_DECLS

#endif
