/*
 * lsm_reinitialization3d.h -- reinitialization and orthogonalization RHS on the GPU
 * (reference: src/reinitialization/lsm_reinitialization3d.h).
 */
#ifndef LSMHIP_REINITIALIZATION_3D_H
#define LSMHIP_REINITIALIZATION_3D_H

#include "lsmhip_config.h"

#define LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS lsm3dcomputereinitializationeqnrhs_
#define LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS lsm3dcomputeorthogonalizationeqnrhs_

LSMHIP_BEGIN_DECLS

/* replaces src/reinitialization/lsm_reinitialization3d.h:77 */
void LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS(
    LSMLIB_REAL *reinit_rhs, LSMHIP_BOX_ARGS(rhs_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
    const LSMLIB_REAL *phi0, LSMHIP_BOX_ARGS(phi0_gb), const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
    const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus,
    const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb),
    LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz,
    const int *use_phi0_for_sgn);

/* replaces src/reinitialization/lsm_reinitialization3d.h:163 */
void LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS(
    LSMLIB_REAL *ortho_rhs, LSMHIP_BOX_ARGS(rhs_gb), const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
    const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus,
    const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *psi, LSMHIP_BOX_ARGS(psi_gb), const LSMLIB_REAL *psi_x_plus, const LSMLIB_REAL *psi_y_plus,
    const LSMLIB_REAL *psi_z_plus, LSMHIP_BOX_ARGS(grad_psi_plus_gb), const LSMLIB_REAL *psi_x_minus,
    const LSMLIB_REAL *psi_y_minus, const LSMLIB_REAL *psi_z_minus, LSMHIP_BOX_ARGS(grad_psi_minus_gb),
    LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

LSMHIP_END_DECLS

#endif
