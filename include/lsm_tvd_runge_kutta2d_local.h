/*
 * lsm_tvd_runge_kutta2d_local.h -- 2D narrow-band TVD Runge-Kutta stages on the GPU
 * (reference: src/toolbox/lsm_tvd_runge_kutta2d_local.h).
 */
#ifndef LSMHIP_LSM_TVD_RUNGE_KUTTA2D_LOCAL_H
#define LSMHIP_LSM_TVD_RUNGE_KUTTA2D_LOCAL_H

#include "lsmhip_config.h"

#define LSM2D_RK1_STEP_LOCAL lsm2drk1steplocal_
#define LSM2D_TVD_RK2_STAGE1_LOCAL lsm2dtvdrk2stage1local_
#define LSM2D_TVD_RK2_STAGE2_LOCAL lsm2dtvdrk2stage2local_
#define LSM2D_TVD_RK3_STAGE1_LOCAL lsm2dtvdrk3stage1local_
#define LSM2D_TVD_RK3_STAGE2_LOCAL lsm2dtvdrk3stage2local_
#define LSM2D_TVD_RK3_STAGE3_LOCAL lsm2dtvdrk3stage3local_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_tvd_runge_kutta2d_local.h:35 */
void LSM2D_RK1_STEP_LOCAL(LSMLIB_REAL *u_next, LSMHIP_BOX2_ARGS(u_next_gb), const LSMLIB_REAL *u_cur,
    LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX2_ARGS(rhs_gb), const LSMLIB_REAL *dt,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d_local.h:63 */
void LSM2D_TVD_RK2_STAGE1_LOCAL(LSMLIB_REAL *u_stage1, LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur,
    LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX2_ARGS(rhs_gb), const LSMLIB_REAL *dt,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d_local.h:92 */
void LSM2D_TVD_RK2_STAGE2_LOCAL(LSMLIB_REAL *u_next, LSMHIP_BOX2_ARGS(u_next_gb), const LSMLIB_REAL *u_stage1,
    LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs,
    LSMHIP_BOX2_ARGS(rhs_gb), const LSMLIB_REAL *dt, const int  *index_x, const int  *index_y,
    const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d_local.h:125 */
void LSM2D_TVD_RK3_STAGE1_LOCAL(LSMLIB_REAL *u_stage1, LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur,
    LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX2_ARGS(rhs_gb), const LSMLIB_REAL *dt,
    const int  *index_x, const int  *index_y, const int  *nlo_index, const int  *nhi_index,
    const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d_local.h:154 */
void LSM2D_TVD_RK3_STAGE2_LOCAL(LSMLIB_REAL *u_stage2, LSMHIP_BOX2_ARGS(u_stage2_gb), const LSMLIB_REAL *u_stage1,
    LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs,
    LSMHIP_BOX2_ARGS(rhs_gb), const LSMLIB_REAL *dt, const int  *index_x, const int  *index_y,
    const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d_local.h:188 */
void LSM2D_TVD_RK3_STAGE3_LOCAL(LSMLIB_REAL *u_next, LSMHIP_BOX2_ARGS(u_next_gb), const LSMLIB_REAL *u_stage2,
    LSMHIP_BOX2_ARGS(u_stage2_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs,
    LSMHIP_BOX2_ARGS(rhs_gb), const LSMLIB_REAL *dt, const int  *index_x, const int  *index_y,
    const int  *nlo_index, const int  *nhi_index, const unsigned char  *narrow_band, LSMHIP_BOX2_ARGS(nb_gb),
    const unsigned char  *mark_fb);

LSMHIP_END_DECLS

#endif
