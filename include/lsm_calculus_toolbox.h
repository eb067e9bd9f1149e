/*
 * lsm_calculus_toolbox.h -- smoothed Heaviside and delta functions (host
 * scalars): the Fortran routines of src/toolbox/lsm_calculus_toolbox.f:25 and
 * :68 with their truncated 1/pi = 0.31830988618379, and the C macros of
 * src/toolbox/lsm_calculus_toolbox.h:41-95 (full-precision pi from math.h).
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

#include <math.h>

/* Sine-smoothed Heaviside of x with smoothing half-width eps
 * (src/toolbox/lsm_calculus_toolbox.h:41). */
#define LSM_HEAVISIDE(x, eps) \
  (((x) < -(eps)) ? 0 : (((x) > (eps)) ? 1 : 0.5 * (1 + (x) / (eps) + 1 / M_PI * sin(M_PI * (x) / (eps)))))
/* Cosine-smoothed delta function (:58). */
#define LSM_DELTA_FUNCTION(x, eps) \
  (((x) < -(eps)) ? 0 : (((x) > (eps)) ? 0 : 0.5 / (eps) * (1 + cos(M_PI * (x) / (eps)))))
/* Piecewise-quadratic ("hat") Heaviside (:74).  The reference compares x with a
 * variable named zero that the caller must provide; here it is the constant 0. */
#define LSM_HEAVISIDE_HAT(x, eps)                                             \
  (((x) < -(eps)) ? 0                                                         \
                  : (((x) > (eps)) ? 1                                        \
                                   : (((x) > 0) ? 0.5 * ((x) + (eps)) * ((x) + (eps)) / ((eps) * (eps)) \
                                                : 1.0 - 0.5 * ((eps) - (x)) * ((eps) - (x)) / ((eps) * (eps)))))
/* Piecewise-linear ("hat") delta function (:91). */
#define LSM_DELTA_FUNCTION_HAT(x, eps) \
  (((x) < -(eps)) ? 0 : (((x) > (eps)) ? 0 : ((eps) - fabs(x)) / ((eps) * (eps))))

#endif
