/*
 * lsm_boundary_conditions.h -- the reference's ghost-fill dispatch over a Grid
 * under its own names (src/boundary_conditions/lsm_boundary_conditions.h:70-176).
 * phi is a device array on grid's ghost box; the fills act on the boundaries
 * of grid's fill box.
 *
 * Location index (the reference's BOUNDARY_LOCATION_IDX): 0..5 one face
 * (x_lo, x_hi, y_lo, y_hi, z_lo, z_hi), 6/7/8 both faces of x/y/z, 9 every
 * face in the order x_lo, x_hi, y_lo, y_hi, z_lo, z_hi (edge and corner
 * ghosts depend on that order).  A 2D grid has the faces 0..3 and the groups
 * 6, 7 and 9; like the reference, any other index (and 4, 5, 8 on a 2D grid)
 * does nothing.  In 3D, a group is filled in one launch whose values are
 * bitwise those of its faces filled one after the other.  A NULL phi or grid
 * sets LSMHIP_ERR_ARG and does nothing.
 */
#ifndef LSMHIP_LSM_BOUNDARY_CONDITIONS_H
#define LSMHIP_LSM_BOUNDARY_CONDITIONS_H

#include "lsm_grid.h"

typedef enum {
  X_LO = 0,
  X_HI = 1,
  Y_LO = 2,
  Y_HI = 3,
  Z_LO = 4,
  Z_HI = 5,
  X_LO_AND_X_HI = 6,
  Y_LO_AND_Y_HI = 7,
  Z_LO_AND_Z_HI = 8,
  ALL_BOUNDARIES = 9
} BOUNDARY_LOCATION_IDX;

LSMHIP_BEGIN_DECLS

/* replaces linearExtrapolationBC (src/boundary_conditions/lsm_boundary_conditions.h:95-98) */
void linearExtrapolationBC(LSMLIB_REAL *phi, Grid *grid, int bdry_location_idx);
/* replaces signedLinearExtrapolationBC (src/boundary_conditions/lsm_boundary_conditions.h:130-133) */
void signedLinearExtrapolationBC(LSMLIB_REAL *phi, Grid *grid, int bdry_location_idx);
/* replaces copyExtrapolationBC (src/boundary_conditions/lsm_boundary_conditions.h:150-153) */
void copyExtrapolationBC(LSMLIB_REAL *phi, Grid *grid, int bdry_location_idx);
/* replaces homogeneousNeumannBC (src/boundary_conditions/lsm_boundary_conditions.h:173-176): copy extrapolation */
void homogeneousNeumannBC(LSMLIB_REAL *phi, Grid *grid, int bdry_location_idx);

LSMHIP_END_DECLS

#endif
