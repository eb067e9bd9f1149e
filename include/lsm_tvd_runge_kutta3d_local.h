/*
 * lsm_tvd_runge_kutta3d_local.h -- narrow-band TVD Runge-Kutta stages on the
 * GPU (reference: src/toolbox/lsm_tvd_runge_kutta3d_local.h).
 */
#ifndef LSMHIP_TVD_RUNGE_KUTTA_3D_LOCAL_H
#define LSMHIP_TVD_RUNGE_KUTTA_3D_LOCAL_H

#include "lsm_spatial_derivatives3d_local.h"

#define LSM3D_RK1_STEP_LOCAL lsm3drk1steplocal_
#define LSM3D_TVD_RK2_STAGE1_LOCAL lsm3dtvdrk2stage1local_
#define LSM3D_TVD_RK2_STAGE2_LOCAL lsm3dtvdrk2stage2local_
#define LSM3D_TVD_RK3_STAGE1_LOCAL lsm3dtvdrk3stage1local_
#define LSM3D_TVD_RK3_STAGE2_LOCAL lsm3dtvdrk3stage2local_
#define LSM3D_TVD_RK3_STAGE3_LOCAL lsm3dtvdrk3stage3local_

LSMHIP_BEGIN_DECLS

#define LSMHIP_RK_LOCAL_1(out)                                                                            \
  LSMLIB_REAL *out, LSMHIP_BOX_ARGS(out##_gb), const LSMLIB_REAL *u_cur, LSMHIP_BOX_ARGS(u_cur_gb),       \
      const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb), const LSMLIB_REAL *dt, LSMHIP_LOCAL_TAIL
#define LSMHIP_RK_LOCAL_2(out, prev)                                                                      \
  LSMLIB_REAL *out, LSMHIP_BOX_ARGS(out##_gb), const LSMLIB_REAL *prev, LSMHIP_BOX_ARGS(prev##_gb),       \
      const LSMLIB_REAL *u_cur, LSMHIP_BOX_ARGS(u_cur_gb), const LSMLIB_REAL *rhs, LSMHIP_BOX_ARGS(rhs_gb), \
      const LSMLIB_REAL *dt, LSMHIP_LOCAL_TAIL

/* replaces src/toolbox/lsm_tvd_runge_kutta3d_local.h:35 */
void LSM3D_RK1_STEP_LOCAL(LSMHIP_RK_LOCAL_1(u_next));
/* replaces :73 */
void LSM3D_TVD_RK2_STAGE1_LOCAL(LSMHIP_RK_LOCAL_1(u_stage1));
/* replaces :111 */
void LSM3D_TVD_RK2_STAGE2_LOCAL(LSMHIP_RK_LOCAL_2(u_next, u_stage1));
/* replaces :155 */
void LSM3D_TVD_RK3_STAGE1_LOCAL(LSMHIP_RK_LOCAL_1(u_stage1));
/* replaces :192 */
void LSM3D_TVD_RK3_STAGE2_LOCAL(LSMHIP_RK_LOCAL_2(u_stage2, u_stage1));
/* replaces :237 */
void LSM3D_TVD_RK3_STAGE3_LOCAL(LSMHIP_RK_LOCAL_2(u_next, u_stage2));

LSMHIP_END_DECLS

#endif
