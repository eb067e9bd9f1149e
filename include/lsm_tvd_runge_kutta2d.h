/*
 * lsm_tvd_runge_kutta2d.h -- 2D TVD Runge-Kutta stages on the GPU
 * (reference: src/toolbox/lsm_tvd_runge_kutta2d.h).
 */
#ifndef LSMHIP_LSM_TVD_RUNGE_KUTTA2D_H
#define LSMHIP_LSM_TVD_RUNGE_KUTTA2D_H

#include "lsmhip_config.h"

#define LSM2D_RK1_STEP lsm2drk1step_
#define LSM2D_TVD_RK2_STAGE1 lsm2dtvdrk2stage1_
#define LSM2D_TVD_RK2_STAGE2 lsm2dtvdrk2stage2_
#define LSM2D_TVD_RK3_STAGE1 lsm2dtvdrk3stage1_
#define LSM2D_TVD_RK3_STAGE2 lsm2dtvdrk3stage2_
#define LSM2D_TVD_RK3_STAGE3 lsm2dtvdrk3stage3_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_tvd_runge_kutta2d.h:59 */
void LSM2D_RK1_STEP(LSMLIB_REAL *u_next, LSMHIP_BOX2_ARGS(u_next_gb), const LSMLIB_REAL *u_cur,
    LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX2_ARGS(rhs_gb), LSMHIP_BOX2_ARGS(fb),
    const LSMLIB_REAL *dt);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d.h:100 */
void LSM2D_TVD_RK2_STAGE1(LSMLIB_REAL *u_stage1, LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur,
    LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX2_ARGS(rhs_gb), LSMHIP_BOX2_ARGS(fb),
    const LSMLIB_REAL *dt);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d.h:138 */
void LSM2D_TVD_RK2_STAGE2(LSMLIB_REAL *u_next, LSMHIP_BOX2_ARGS(u_next_gb), const LSMLIB_REAL *u_stage1,
    LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs,
    LSMHIP_BOX2_ARGS(rhs_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dt);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d.h:184 */
void LSM2D_TVD_RK3_STAGE1(LSMLIB_REAL *u_stage1, LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur,
    LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX2_ARGS(rhs_gb), LSMHIP_BOX2_ARGS(fb),
    const LSMLIB_REAL *dt);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d.h:222 */
void LSM2D_TVD_RK3_STAGE2(LSMLIB_REAL *u_stage2, LSMHIP_BOX2_ARGS(u_stage2_gb), const LSMLIB_REAL *u_stage1,
    LSMHIP_BOX2_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs,
    LSMHIP_BOX2_ARGS(rhs_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dt);
/* replaces src/toolbox/lsm_tvd_runge_kutta2d.h:265 */
void LSM2D_TVD_RK3_STAGE3(LSMLIB_REAL *u_next, LSMHIP_BOX2_ARGS(u_next_gb), const LSMLIB_REAL *u_stage2,
    LSMHIP_BOX2_ARGS(u_stage2_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX2_ARGS(u_cur_gb), const LSMLIB_REAL *rhs,
    LSMHIP_BOX2_ARGS(rhs_gb), LSMHIP_BOX2_ARGS(fb), const LSMLIB_REAL *dt);

LSMHIP_END_DECLS

#endif
