/*
 * lsm_reinitialization2d.h -- 2D reinitialization and orthogonalization RHS on the GPU
 * (reference: src/reinitialization/lsm_reinitialization2d.h).
 */
#ifndef LSMHIP_LSM_REINITIALIZATION2D_H
#define LSMHIP_LSM_REINITIALIZATION2D_H

#include "lsmhip_config.h"

#define LSM2D_COMPUTE_REINITIALIZATION_EQN_RHS lsm2dcomputereinitializationeqnrhs_
#define LSM2D_COMPUTE_ORTHOGONALIZATION_EQN_RHS lsm2dcomputeorthogonalizationeqnrhs_

LSMHIP_BEGIN_DECLS

/* replaces src/reinitialization/lsm_reinitialization2d.h:78 */
void LSM2D_COMPUTE_REINITIALIZATION_EQN_RHS(LSMLIB_REAL *reinit_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *phi0, LSMHIP_BOX2_ARGS(phi0_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const int *use_phi0_for_sgn);
/* replaces src/reinitialization/lsm_reinitialization2d.h:149 */
void LSM2D_COMPUTE_ORTHOGONALIZATION_EQN_RHS(LSMLIB_REAL *ortho_rhs, LSMHIP_BOX2_ARGS(rhs_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *psi, LSMHIP_BOX2_ARGS(psi_gb), const LSMLIB_REAL *psi_x_plus, const LSMLIB_REAL *psi_y_plus,
    LSMHIP_BOX2_ARGS(grad_psi_plus_gb), const LSMLIB_REAL *psi_x_minus, const LSMLIB_REAL *psi_y_minus,
    LSMHIP_BOX2_ARGS(grad_psi_minus_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);

LSMHIP_END_DECLS

#endif
