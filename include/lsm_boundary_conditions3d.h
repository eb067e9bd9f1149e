/*
 * lsm_boundary_conditions3d.h -- ghost-cell fills of one face on the GPU
 * (reference: src/boundary_conditions/lsm_boundary_conditions3d.h).
 * bdry_location_idx 0..5 = x_lo, x_hi, y_lo, y_hi, z_lo, z_hi; other values
 * do nothing, as in the reference.  lsmhip_fill_ghosts (lsmhip.h) applies a
 * fill to the face groups 0..9 of the reference's C helpers.
 */
#ifndef LSMHIP_BOUNDARY_CONDITIONS_3D_H
#define LSMHIP_BOUNDARY_CONDITIONS_3D_H

#include "lsmhip_config.h"

#define LSM3D_LINEAR_EXTRAPOLATION lsm3dlinearextrapolation_
#define LSM3D_SIGNED_LINEAR_EXTRAPOLATION lsm3dsignedlinearextrapolation_
#define LSM3D_COPY_EXTRAPOLATION lsm3dcopyextrapolation_
#define LSM3D_HOMOGENEOUS_NEUMANN_ENO1 lsm3dhomogeneousneumanneno1_
#define LSM3D_HOMOGENEOUS_NEUMANN_ENO2 lsm3dhomogeneousneumanneno2_
#define LSM3D_HOMOGENEOUS_NEUMANN_ENO3 lsm3dhomogeneousneumanneno3_
#define LSM3D_HOMOGENEOUS_NEUMANN_WENO5 lsm3dhomogeneousneumannweno5_

LSMHIP_BEGIN_DECLS

#define LSMHIP_BC_ARGS LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb), const int *bdry_location_idx

/* replaces src/boundary_conditions/lsm_boundary_conditions3d.h:71 */
void LSM3D_LINEAR_EXTRAPOLATION(LSMHIP_BC_ARGS);
/* replaces :111 */
void LSM3D_SIGNED_LINEAR_EXTRAPOLATION(LSMHIP_BC_ARGS);
/* replaces :135 */
void LSM3D_COPY_EXTRAPOLATION(LSMHIP_BC_ARGS);
/* replaces :161, :186, :211, :236 (all four are copy extrapolation) */
void LSM3D_HOMOGENEOUS_NEUMANN_ENO1(LSMHIP_BC_ARGS);
void LSM3D_HOMOGENEOUS_NEUMANN_ENO2(LSMHIP_BC_ARGS);
void LSM3D_HOMOGENEOUS_NEUMANN_ENO3(LSMHIP_BC_ARGS);
void LSM3D_HOMOGENEOUS_NEUMANN_WENO5(LSMHIP_BC_ARGS);

LSMHIP_END_DECLS

#endif
