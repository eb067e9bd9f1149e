/*
 * lsm_data_arrays.h -- the reference's data-array files under their own names
 * (src/utils/lsm_data_arrays.h:190-244): three ints (the ghost-box
 * dimensions), then the reals of the array, column-major.
 *
 * data may be a device array or a host array (the writers look the pointer up
 * with hipPointerGetAttributes and copy device data through the host).
 * readDataArray returns host memory from malloc, as the reference's does
 * (release it with free()); copy it to the device before passing it to the
 * kernels, or read a file straight into a device array with
 * lsmhip_read_data_array (lsm_data_io.h).  A file that cannot be opened gives
 * NULL and the reference's message "Could not open file <name>"; a short file
 * or a header with a negative dimension gives NULL with LSMHIP_ERR_ARG.
 */
#ifndef LSMHIP_LSM_DATA_ARRAYS_H
#define LSMHIP_LSM_DATA_ARRAYS_H

#include "lsm_grid.h"

LSMHIP_BEGIN_DECLS

/* replaces writeDataArray (src/utils/lsm_data_arrays.h:190-191): grid->num_gridpts reals */
void writeDataArray(LSMLIB_REAL *data, Grid *grid, char *file_name, int zip_status);
/* replaces writeDataArrayNoGrid (src/utils/lsm_data_arrays.h:218-219): n[0] n[1] n[2] reals */
void writeDataArrayNoGrid(LSMLIB_REAL *data, int *n, char *file_name, int zip_status);
/* replaces readDataArray (src/utils/lsm_data_arrays.h:244) */
LSMLIB_REAL *readDataArray(int *grid_dims, char *file_name);

LSMHIP_END_DECLS

#endif
