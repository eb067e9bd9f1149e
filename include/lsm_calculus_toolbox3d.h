/*
 * lsm_calculus_toolbox3d.h -- first-order delta function of the zero level set
 * on the GPU (reference: src/toolbox/lsm_calculus_toolbox3d.h).
 */
#ifndef LSMHIP_CALCULUS_TOOLBOX_3D_H
#define LSMHIP_CALCULUS_TOOLBOX_3D_H

#include "lsmhip_config.h"

#define LSM3D_DELTA_FUNCTION_ORDER1 lsm3ddeltafunctionorder1_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_calculus_toolbox3d.h:47.  delta (written over fb)
 * shares phi's ghost box.  As the reference's Fortran
 * (src/toolbox/lsm_calculus_toolbox3d.f.in:15-141), the z terms use norm_phi_x
 * and 1/dx. */
void LSM3D_DELTA_FUNCTION_ORDER1(const LSMLIB_REAL *phi, LSMLIB_REAL *delta, LSMHIP_BOX_ARGS(gb),
                                 const LSMLIB_REAL *norm_phi_x, const LSMLIB_REAL *norm_phi_y,
                                 const LSMLIB_REAL *norm_phi_z, LSMHIP_BOX_ARGS(grad_phi_gb), LSMHIP_BOX_ARGS(fb),
                                 const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *dz);

LSMHIP_END_DECLS

#endif
