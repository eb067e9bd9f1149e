/*
 * lsm_reinitialization2d_local.h -- 2D narrow-band reinitialization and orthogonalization RHS on the GPU
 * (reference: src/reinitialization/lsm_reinitialization2d_local.h).
 */
#ifndef LSMHIP_LSM_REINITIALIZATION2D_LOCAL_H
#define LSMHIP_LSM_REINITIALIZATION2D_LOCAL_H

#include "lsmhip_config.h"

#define LSM2D_COMPUTE_REINITIALIZATION_EQN_RHS_LOCAL lsm2dcomputereinitializationeqnrhslocal_
#define LSM2D_COMPUTE_ORTHOGONALIZATION_EQN_RHS_LOCAL lsm2dcomputeorthogonalizationeqnrhslocal_

LSMHIP_BEGIN_DECLS

/* replaces src/reinitialization/lsm_reinitialization2d_local.h:41 */
void LSM2D_COMPUTE_REINITIALIZATION_EQN_RHS_LOCAL(LSMLIB_REAL *reinit_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *phi0, LSMHIP_BOX2_ARGS(phi0_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int *use_phi0_for_sgn, const int  *index_x,
    const int  *index_y, const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band,
    LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/reinitialization/lsm_reinitialization2d_local.h:82 */
void LSM2D_COMPUTE_ORTHOGONALIZATION_EQN_RHS_LOCAL(LSMLIB_REAL *ortho_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *psi, LSMHIP_BOX2_ARGS(psi_gb), const LSMLIB_REAL *psi_x_plus, const LSMLIB_REAL *psi_y_plus,
    LSMHIP_BOX2_ARGS(grad_psi_plus_gb), const LSMLIB_REAL *psi_x_minus, const LSMLIB_REAL *psi_y_minus,
    LSMHIP_BOX2_ARGS(grad_psi_minus_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int  *index_x,
    const int  *index_y, const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band,
    LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);

LSMHIP_END_DECLS

#endif
