/*
 * lsm_calculus_toolbox.h -- smoothed Heaviside and delta functions (host
 * scalars), the Fortran routines of src/toolbox/lsm_calculus_toolbox.f:25 and
 * :68 with their truncated 1/pi = 0.31830988618379.
 */
#ifndef LSMHIP_CALCULUS_TOOLBOX_H
#define LSMHIP_CALCULUS_TOOLBOX_H

#include "lsmhip_config.h"

LSMHIP_BEGIN_DECLS

/* replaces lsmHeaviside, src/toolbox/lsm_calculus_toolbox.f:25 */
void lsmheaviside_(LSMLIB_REAL *H, const LSMLIB_REAL *x, const LSMLIB_REAL *epsilon);
/* replaces lsmDeltaFunction, src/toolbox/lsm_calculus_toolbox.f:68 */
void lsmdeltafunction_(LSMLIB_REAL *delta, const LSMLIB_REAL *x, const LSMLIB_REAL *epsilon);

LSMHIP_END_DECLS

#endif
