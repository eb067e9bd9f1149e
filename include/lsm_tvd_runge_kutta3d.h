/*
 * lsm_tvd_runge_kutta3d.h -- TVD Runge-Kutta stages on the GPU
 * (reference: src/toolbox/lsm_tvd_runge_kutta3d.h).
 */
#ifndef LSMHIP_TVD_RUNGE_KUTTA_3D_H
#define LSMHIP_TVD_RUNGE_KUTTA_3D_H

#include "lsmhip_config.h"

#define LSM3D_RK1_STEP lsm3drk1step_
#define LSM3D_TVD_RK2_STAGE1 lsm3dtvdrk2stage1_
#define LSM3D_TVD_RK2_STAGE2 lsm3dtvdrk2stage2_
#define LSM3D_TVD_RK3_STAGE1 lsm3dtvdrk3stage1_
#define LSM3D_TVD_RK3_STAGE2 lsm3dtvdrk3stage2_
#define LSM3D_TVD_RK3_STAGE3 lsm3dtvdrk3stage3_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_tvd_runge_kutta3d.h:59 */
void LSM3D_RK1_STEP(LSMLIB_REAL *u_next, LSMHIP_BOX_ARGS(u_next_gb), const LSMLIB_REAL *u_cur,
                    LSMHIP_BOX_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb), LSMHIP_BOX_ARGS(fb),
                    const LSMLIB_REAL *dt);
/* replaces :108 */
void LSM3D_TVD_RK2_STAGE1(LSMLIB_REAL *u_stage1, LSMHIP_BOX_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur,
                          LSMHIP_BOX_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb),
                          LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dt);
/* replaces :154 */
void LSM3D_TVD_RK2_STAGE2(LSMLIB_REAL *u_next, LSMHIP_BOX_ARGS(u_next_gb), const LSMLIB_REAL *u_stage1,
                          LSMHIP_BOX_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX_ARGS(u_cur_gb),
                          const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb), LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dt);
/* replaces :210 */
void LSM3D_TVD_RK3_STAGE1(LSMLIB_REAL *u_stage1, LSMHIP_BOX_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur,
                          LSMHIP_BOX_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb),
                          LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dt);
/* replaces :256 */
void LSM3D_TVD_RK3_STAGE2(LSMLIB_REAL *u_stage2, LSMHIP_BOX_ARGS(u_stage2_gb), const LSMLIB_REAL *u_stage1,
                          LSMHIP_BOX_ARGS(u_stage1_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX_ARGS(u_cur_gb),
                          const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb), LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dt);
/* replaces :309 */
void LSM3D_TVD_RK3_STAGE3(LSMLIB_REAL *u_next, LSMHIP_BOX_ARGS(u_next_gb), const LSMLIB_REAL *u_stage2,
                          LSMHIP_BOX_ARGS(u_stage2_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX_ARGS(u_cur_gb),
                          const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb), LSMHIP_BOX_ARGS(fb), const LSMLIB_REAL *dt);

LSMHIP_END_DECLS

#endif
