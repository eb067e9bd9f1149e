/*
 * lsm_math_utils2d_local.h -- 2D narrow-band CFL time steps and max-norm difference on the GPU
 * (reference: src/toolbox/lsm_math_utils2d_local.h).
 */
#ifndef LSMHIP_LSM_MATH_UTILS2D_LOCAL_H
#define LSMHIP_LSM_MATH_UTILS2D_LOCAL_H

#include "lsmhip_config.h"

#define LSM2D_MAX_NORM_DIFF_LOCAL lsm2dmaxnormdifflocal_
#define LSM2D_COMPUTE_STABLE_ADVECTION_DT_LOCAL lsm2dcomputestableadvectiondtlocal_
#define LSM2D_COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL lsm2dcomputestablenormalveldtlocal_
#define LSM2D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL lsm2dcomputestableconstnormalveldtlocal_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_math_utils2d_local.h:64 */
void LSM2D_MAX_NORM_DIFF_LOCAL(LSMLIB_REAL *max_norm_diff, const LSMLIB_REAL *field1, LSMHIP_BOX2_ARGS(field1_gb),
    const LSMLIB_REAL *field2, LSMHIP_BOX2_ARGS(field2_gb), const int  *index_x, const int  *index_y,
    const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_math_utils2d_local.h:146 */
void LSM2D_COMPUTE_STABLE_ADVECTION_DT_LOCAL(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_x, const LSMLIB_REAL *vel_y,
    LSMHIP_BOX2_ARGS(vel_gb), const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *cfl_number,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_math_utils2d_local.h:196 */
void LSM2D_COMPUTE_STABLE_NORMAL_VEL_DT_LOCAL(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n, LSMHIP_BOX2_ARGS(vel_gb),
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *cfl_number, const int  *index_x,
    const int  *index_y, const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band,
    LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_math_utils2d_local.h:257 */
void LSM2D_COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL(LSMLIB_REAL *dt, const LSMLIB_REAL *vel_n,
    const LSMLIB_REAL *phi_x_plus, const LSMLIB_REAL *phi_y_plus, LSMHIP_BOX2_ARGS(grad_phi_plus_gb),
    const LSMLIB_REAL *phi_x_minus, const LSMLIB_REAL *phi_y_minus, LSMHIP_BOX2_ARGS(grad_phi_minus_gb),
    const LSMLIB_REAL *dx, const LSMLIB_REAL *dy, const LSMLIB_REAL *cfl_number, const int  *index_x,
    const int  *index_y, const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band,
    LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);

LSMHIP_END_DECLS

#endif
