/*
 * lsm_boundary_conditions2d.h -- 2D ghost-cell fills on the GPU
 * (reference: src/boundary_conditions/lsm_boundary_conditions2d.h).
 */
#ifndef LSMHIP_LSM_BOUNDARY_CONDITIONS2D_H
#define LSMHIP_LSM_BOUNDARY_CONDITIONS2D_H

#include "lsmhip_config.h"

#define LSM2D_LINEAR_EXTRAPOLATION lsm2dlinearextrapolation_
#define LSM2D_SIGNED_LINEAR_EXTRAPOLATION lsm2dsignedlinearextrapolation_
#define LSM2D_COPY_EXTRAPOLATION lsm2dcopyextrapolation_
#define LSM2D_HOMOGENEOUS_NEUMANN_ENO1 lsm2dhomogeneousneumanneno1_
#define LSM2D_HOMOGENEOUS_NEUMANN_ENO2 lsm2dhomogeneousneumanneno2_
#define LSM2D_HOMOGENEOUS_NEUMANN_ENO3 lsm2dhomogeneousneumanneno3_
#define LSM2D_HOMOGENEOUS_NEUMANN_WENO5 lsm2dhomogeneousneumannweno5_

LSMHIP_BEGIN_DECLS

/* replaces src/boundary_conditions/lsm_boundary_conditions2d.h:69 */
void LSM2D_LINEAR_EXTRAPOLATION(LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const int *bdry_location_idx);
/* replaces src/boundary_conditions/lsm_boundary_conditions2d.h:107 */
void LSM2D_SIGNED_LINEAR_EXTRAPOLATION(LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const int *bdry_location_idx);
/* replaces src/boundary_conditions/lsm_boundary_conditions2d.h:129 */
void LSM2D_COPY_EXTRAPOLATION(LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const int *bdry_location_idx);
/* replaces src/boundary_conditions/lsm_boundary_conditions2d.h:153 */
void LSM2D_HOMOGENEOUS_NEUMANN_ENO1(LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const int *bdry_location_idx);
/* replaces src/boundary_conditions/lsm_boundary_conditions2d.h:176 */
void LSM2D_HOMOGENEOUS_NEUMANN_ENO2(LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const int *bdry_location_idx);
/* replaces src/boundary_conditions/lsm_boundary_conditions2d.h:199 */
void LSM2D_HOMOGENEOUS_NEUMANN_ENO3(LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const int *bdry_location_idx);
/* replaces src/boundary_conditions/lsm_boundary_conditions2d.h:222 */
void LSM2D_HOMOGENEOUS_NEUMANN_WENO5(LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const int *bdry_location_idx);

LSMHIP_END_DECLS

#endif
