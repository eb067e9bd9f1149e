/*
 * lsm_data_io.h -- the reference's binary files for level-set data arrays and
 * grids (src/utils/lsm_data_arrays.h:166-244, src/utils/lsm_grid.h:43-95,
 * 296-336), written and read by the HIP library so that fields computed on
 * the GPU interoperate with the reference's tools and MATLAB readers
 * (spikes/matlab/examples/matlab/visualization/readDataArray.m,
 * readGridFromBinaryFile.m).  Host functions; data arrays may be device
 * arrays (on_device != 0) or host arrays.  No compression: the reference's
 * zip_status shells out to gzip, which the caller can do the same way.
 */
#ifndef LSMHIP_DATA_IO_H
#define LSMHIP_DATA_IO_H

#include "lsmhip_config.h"

/* The reference's Grid (src/utils/lsm_grid.h:43-95): the same fields in the
 * same order, so a Grid of the reference can be passed where an LsmhipGrid is
 * expected. */
typedef struct LsmhipGrid {
  int num_dims;
  LSMLIB_REAL x_lo[3];
  LSMLIB_REAL x_hi[3];
  LSMLIB_REAL x_lo_ghostbox[3];
  LSMLIB_REAL x_hi_ghostbox[3];
  int grid_dims[3];
  int grid_dims_ghostbox[3];
  LSMLIB_REAL dx[3];
  int num_gridpts;
  int ilo_gb, ihi_gb, jlo_gb, jhi_gb, klo_gb, khi_gb;
  int ilo_fb, ihi_fb, jlo_fb, jhi_fb, klo_fb, khi_fb;
  int ilo_D1_fb, ihi_D1_fb, jlo_D1_fb, jhi_D1_fb, klo_D1_fb, khi_D1_fb;
  int ilo_D2_fb, ihi_D2_fb, jlo_D2_fb, jhi_D2_fb, klo_D2_fb, khi_D2_fb;
  int ilo_D3_fb, ihi_D3_fb, jlo_D3_fb, jhi_D3_fb, klo_D3_fb, khi_D3_fb;
  int num_nb_levels;
  unsigned char mark_gb, mark_D1, mark_D2, mark_D3, mark_fb;
  LSMLIB_REAL beta, gamma;
} LsmhipGrid;

LSMHIP_BEGIN_DECLS

/* replaces writeDataArray / writeDataArrayNoGrid (src/utils/lsm_data_arrays.h:190, 218) */
int lsmhip_write_data_array(const LSMLIB_REAL *data, const int *grid_dims_ghostbox, const char *file_name,
                            int on_device);
/* the dimensions readDataArray (src/utils/lsm_data_arrays.h:244) returns */
int lsmhip_read_data_array_dims(int *grid_dims_ghostbox, const char *file_name);
/* replaces readDataArray (src/utils/lsm_data_arrays.h:244), into a caller array of capacity reals */
int lsmhip_read_data_array(LSMLIB_REAL *data, long long capacity, int *grid_dims_ghostbox, const char *file_name,
                           int on_device);
/* replaces createGridSetGridDims (src/utils/lsm_grid.h:207); accuracy 0 ENO1 .. 3 WENO5 */
int lsmhip_grid_set_grid_dims(LsmhipGrid *grid, int num_dims, const int *grid_dims, const LSMLIB_REAL *x_lo,
                              const LSMLIB_REAL *x_hi, int accuracy);
/* replaces createGridSetDx (src/utils/lsm_grid.h:145) */
int lsmhip_grid_set_dx(LsmhipGrid *grid, int num_dims, LSMLIB_REAL dx, const LSMLIB_REAL *x_lo,
                       const LSMLIB_REAL *x_hi, int accuracy);
/* replaces writeGridToBinaryFile (src/utils/lsm_grid.h:313) */
int lsmhip_write_grid(const LsmhipGrid *grid, const char *file_name);
/* replaces readGridFromBinaryFile (src/utils/lsm_grid.h:336) */
int lsmhip_read_grid(LsmhipGrid *grid, const char *file_name);

LSMHIP_END_DECLS

#endif
