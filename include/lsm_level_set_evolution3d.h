/*
 * lsm_level_set_evolution3d.h -- whole-grid RHS terms of the level-set
 * equation on the GPU (reference: src/toolbox/lsm_level_set_evolution3d.h).
 */
#ifndef LSMHIP_LEVEL_SET_EVOLUTION_3D_H
#define LSMHIP_LEVEL_SET_EVOLUTION_3D_H

#include "lsmhip_config.h"

#define LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS lsm3dzerooutlevelseteqnrhs_
#define LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS lsm3daddadvectiontermtolserhs_
#define LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS lsm3daddnormalveltermtolserhs_
#define LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS lsm3daddconstnormalveltermtolserhs_
#define LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS lsm3daddconstcurvtermtolserhs_
#define LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS lsm3daddconstprecomputedcurvtermtolserhs_
#define LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS lsm3daddexternalandnormalveltermtolserhs_

LSMHIP_BEGIN_DECLS

#define LSMHIP_GRAD_PM                                                                                       \
  const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, const LSMLIB_REAL *phi_z_plus,               \
      LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus,     \
      const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb)

/* replaces src/toolbox/lsm_level_set_evolution3d.h:64 (zeroes the whole ghost box) */
void LSM3D_ZERO_OUT_LEVEL_SET_EQN_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb));
/* replaces :94 */
void LSM3D_ADD_ADVECTION_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb), const LSMLIB_REAL *phi_x,
                                         const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                         LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *vel_x,
                                         const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z, LSMHIP_BOX_ARGS(vel_gb),
                                         LSMHIP_BOX_ARGS(fb));
/* replaces :151 */
void LSM3D_ADD_NORMAL_VEL_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb), LSMHIP_GRAD_PM,
                                          const LSMLIB_REAL *vel_n, LSMHIP_BOX_ARGS(vel_gb), LSMHIP_BOX_ARGS(fb));
/* replaces :217 */
void LSM3D_ADD_CONST_NORMAL_VEL_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb), LSMHIP_GRAD_PM,
                                                const LSMLIB_REAL *vel_n, LSMHIP_BOX_ARGS(fb));
/* replaces :277 */
void LSM3D_ADD_CONST_CURV_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(lse_rhs_gb), const LSMLIB_REAL *phi_x,
                                          const LSMLIB_REAL *phi_y, const LSMLIB_REAL *phi_z,
                                          LSMHIP_BOX_ARGS(grad_phi_gb), const LSMLIB_REAL *phi_xx,
                                          const LSMLIB_REAL *phi_xy, const LSMLIB_REAL *phi_xz,
                                          const LSMLIB_REAL *phi_yy, const LSMLIB_REAL *phi_yz,
                                          const LSMLIB_REAL *phi_zz, LSMHIP_BOX_ARGS(grad2_phi_gb),
                                          const LSMLIB_REAL *b, LSMHIP_BOX_ARGS(fb));
/* replaces :334; updates k = klo_fb..khi_rhs_gb, j = jlo_fb..jhi_rhs_gb,
 * i = ilo_rhs_gb..ihi_rhs_gb, the loop bounds of lsm_level_set_evolution3d.f.in:514-516 */
void LSM3D_ADD_CONST_PRECOMPUTED_CURV_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(rhs_gb),
                                                      const LSMLIB_REAL *kappa, LSMHIP_BOX_ARGS(kappa_gb),
                                                      const LSMLIB_REAL *grad_mag_phi, LSMHIP_BOX_ARGS(grad_phi_gb),
                                                      const LSMLIB_REAL *b, LSMHIP_BOX_ARGS(rhs_fb));
/* replaces :398; same loop bounds as above (lsm_level_set_evolution3d.f.in:620-622) */
void LSM3D_ADD_EXTERNAL_AND_NORMAL_VEL_TERM_TO_LSE_RHS(LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(rhs_gb), LSMHIP_GRAD_PM,
                                                       const LSMLIB_REAL *vel_n, const LSMLIB_REAL *vel_x,
                                                       const LSMLIB_REAL *vel_y, const LSMLIB_REAL *vel_z,
                                                       LSMHIP_BOX_ARGS(vel_gb), LSMHIP_BOX_ARGS(rhs_fb));

LSMHIP_END_DECLS

#endif
