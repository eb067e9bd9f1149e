/*
 * lsm_field_extension3d.h -- field-extension RHS on the GPU
 * (reference: src/field_extension/lsm_field_extension3d.h).
 */
#ifndef LSMHIP_FIELD_EXTENSION_3D_H
#define LSMHIP_FIELD_EXTENSION_3D_H

#include "lsmhip_config.h"

#define LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS lsm3dcomputefieldextensioneqnrhs_

LSMHIP_BEGIN_DECLS

/* replaces src/field_extension/lsm_field_extension3d.h:65.  S is not read
   (only its upwind gradient); phi is read one cell away along each axis. */
void LSM3D_COMPUTE_FIELD_EXTENSION_EQN_RHS(
    LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb), const LSMLIB_REAL *S, LSMHIP_BOX_ARGS(S_gb), const LSMLIB_REAL *phi,
    LSMHIP_BOX_ARGS(phi_gb), const LSMLIB_REAL *S_x_upwind, const LSMLIB_REAL *S_y_upwind,
    const LSMLIB_REAL *S_z_upwind, LSMHIP_BOX_ARGS(grad_S_upwind_gb), const LSMLIB_REAL *signed_normal_x,
    const LSMLIB_REAL *signed_normal_y, const LSMLIB_REAL *signed_normal_z, LSMHIP_BOX_ARGS(signed_normal_gb),
    LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

LSMHIP_END_DECLS

#endif
