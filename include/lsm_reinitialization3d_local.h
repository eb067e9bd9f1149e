/*
 * lsm_reinitialization3d_local.h -- narrow-band reinitialization and
 * orthogonalization RHS on the GPU
 * (reference: src/reinitialization/lsm_reinitialization3d_local.h).
 */
#ifndef LSMHIP_REINITIALIZATION_3D_LOCAL_H
#define LSMHIP_REINITIALIZATION_3D_LOCAL_H

#include "lsm_spatial_derivatives3d_local.h"

#define LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_LOCAL lsm3dcomputereinitializationeqnrhslocal_
#define LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS_LOCAL lsm3dcomputeorthogonalizationeqnrhslocal_

LSMHIP_BEGIN_DECLS

/* replaces src/reinitialization/lsm_reinitialization3d_local.h:41 */
void LSM3D_COMPUTE_REINITIALIZATION_EQN_RHS_LOCAL(
    LSMLIB_REAL *reinit_rhs, LSMHIP_BOX_ARGS(rhs_gb), const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(phi_gb),
    const LSMLIB_REAL *phi0, LSMHIP_BOX_ARGS(phi0_gb), const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
    const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus,
    const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb), LSMHIP_LDXYZ,
    const int *use_phi0_for_sgn, LSMHIP_LOCAL_TAIL);
/* replaces src/reinitialization/lsm_reinitialization3d_local.h:99 */
void LSM3D_COMPUTE_ORTHOGONALIZATION_EQN_RHS_LOCAL(
    LSMLIB_REAL *ortho_rhs, LSMHIP_BOX_ARGS(rhs_gb), const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus,
    const LSMLIB_REAL *phi_z_plus, LSMHIP_BOX_ARGS(grad_phi_plus_gb), const LSMLIB_REAL *phi_x_minus,
    const LSMLIB_REAL *phi_y_minus, const LSMLIB_REAL *phi_z_minus, LSMHIP_BOX_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *psi, LSMHIP_BOX_ARGS(psi_gb), const LSMLIB_REAL *psi_x_plus, const LSMLIB_REAL *psi_y_plus,
    const LSMLIB_REAL *psi_z_plus, LSMHIP_BOX_ARGS(grad_psi_plus_gb), const LSMLIB_REAL *psi_x_minus,
    const LSMLIB_REAL *psi_y_minus, const LSMLIB_REAL *psi_z_minus, LSMHIP_BOX_ARGS(grad_psi_minus_gb),
    LSMHIP_LDXYZ, LSMHIP_LOCAL_TAIL);

LSMHIP_END_DECLS

#endif
