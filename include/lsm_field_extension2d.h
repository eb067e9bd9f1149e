/*
 * lsm_field_extension2d.h -- 2D field-extension RHS on the GPU
 * (reference: src/field_extension/lsm_field_extension2d.h).
 */
#ifndef LSMHIP_LSM_FIELD_EXTENSION2D_H
#define LSMHIP_LSM_FIELD_EXTENSION2D_H

#include "lsmhip_config.h"

#define LSM2D_COMPUTE_FIELD_EXTENSION_EQN_RHS lsm2dcomputefieldextensioneqnrhs_

LSMHIP_BEGIN_DECLS

/* replaces src/field_extension/lsm_field_extension2d.h:65 */
void LSM2D_COMPUTE_FIELD_EXTENSION_EQN_RHS(LSMLIB_REAL *rhs, LSMHIP_BOX2_ARGS(rhs_gb), const LSMLIB_REAL *S,
    LSMHIP_BOX2_ARGS(S_gb), const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(phi_gb), const LSMLIB_REAL *S_x_upwind,
    const LSMLIB_REAL *S_y_upwind, LSMHIP_BOX2_ARGS(grad_S_upwind_gb), const LSMLIB_REAL *signed_normal_x,
    const LSMLIB_REAL *signed_normal_y, LSMHIP_BOX2_ARGS(signed_normal_gb), LSMHIP_BOX2_ARGS(fb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy);

LSMHIP_END_DECLS

#endif
