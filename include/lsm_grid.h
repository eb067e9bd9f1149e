/*
 * lsm_grid.h -- the reference's Grid and its construction and file functions
 * under their own names (src/utils/lsm_grid.h:43-107, 145-336).  Grid is the
 * reference's struct, field for field (LsmhipGrid of lsm_data_io.h), so a Grid
 * built by either library can be passed to the other.  Host functions; grids
 * are allocated with calloc, as the reference's are, and may be released with
 * free() or destroyGrid().
 *
 * Behaviour kept from the reference: num_dims other than 2 or 3 is replaced
 * by 2 with a message; a grid file that cannot be opened gives a zeroed Grid
 * and the message "Could not open file <name>".  Added: an accuracy outside
 * LOW..VERY_HIGH (which the reference uses as an array index unchecked)
 * returns NULL with LSMHIP_ERR_ARG, and a short or unwritable file sets
 * LSMHIP_ERR_ARG (lsmhip_get_last_error).
 */
#ifndef LSMHIP_LSM_GRID_H
#define LSMHIP_LSM_GRID_H

#include "lsm_data_io.h"
#include "lsm_file.h"

typedef LsmhipGrid Grid;

/* src/utils/lsm_grid.h:106-107: HJ ENO1, ENO2, ENO3, WENO5 */
typedef enum { LOW = 0, MEDIUM = 1, HIGH = 2, VERY_HIGH = 3 } LSMLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE;

LSMHIP_BEGIN_DECLS

/* replaces createGridSetDx (src/utils/lsm_grid.h:145-147) */
Grid *createGridSetDx(int num_dims, LSMLIB_REAL dx, LSMLIB_REAL *x_lo, LSMLIB_REAL *x_hi,
                      LSMLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy);
/* replaces createGridSetDxDyDz (src/utils/lsm_grid.h:177-180) */
Grid *createGridSetDxDyDz(int num_dims, LSMLIB_REAL dx, LSMLIB_REAL dy, LSMLIB_REAL dz, LSMLIB_REAL *x_lo,
                          LSMLIB_REAL *x_hi, LSMLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy);
/* replaces createGridSetGridDims (src/utils/lsm_grid.h:207-209) */
Grid *createGridSetGridDims(int num_dims, int *grid_dims, LSMLIB_REAL *x_lo, LSMLIB_REAL *x_hi,
                            LSMLIB_SPATIAL_DERIVATIVE_ACCURACY_TYPE accuracy);
/* replaces destroyGrid (src/utils/lsm_grid.h:231) */
void destroyGrid(Grid *grid);
/* replaces writeGridToBinaryFile (src/utils/lsm_grid.h:313) */
void writeGridToBinaryFile(Grid *grid, char *file_name, int zip_status);
/* replaces readGridFromBinaryFile (src/utils/lsm_grid.h:336) */
Grid *readGridFromBinaryFile(char *file_name);

LSMHIP_END_DECLS

#endif
