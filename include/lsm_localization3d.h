/*
 * lsm_localization3d.h -- narrow-band construction and band helpers on the
 * GPU (reference: src/toolbox/lsm_localization3d.h).  phi, narrow_band, the
 * index lists and index_outer are device arrays; n_lo/n_hi (level tables of
 * length level+1) and the outer-layer counters are host memory.  Index lists
 * and marks are bit-exact to the reference's sequential construction.  The
 * builder refuses (LSMHIP_ERR_ARG) when the lists would overflow
 * [nlo_index, nhi_index] or [nlo_index_outer, nhi_index_outer], which the
 * reference does not check.
 */
#ifndef LSMHIP_LOCALIZATION_3D_H
#define LSMHIP_LOCALIZATION_3D_H

#include "lsmhip_config.h"

#define LSM3D_DETERMINE_NARROW_BAND lsm3ddeterminenarrowband_
#define LSM3D_MARK_NARROW_BAND_BOUNDARY_LAYER lsm3dmarknarrowbandboundarylayer_
#define LSM3D_DETERMINE_NARROW_BAND_FROM_MASK lsm3ddeterminenarrowbandfrommask_
#define LSM3D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL lsm3dmultiplycutofflserhslocal_
#define LSM3D_CHECK_OUTER_NARROW_BAND_LAYER lsm3dcheckouternarrowbandlayer_
#define LSM3D_IMPOSE_MASK_LOCAL lsm3dimposemasklocal_
#define LSM3D_COPY_DATA_LOCAL lsm3dcopydatalocal_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_localization3d.h:89 */
void LSM3D_DETERMINE_NARROW_BAND(const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb), unsigned char *narrow_band,
                                 LSMHIP_BOX_ARGS(nb_gb), int *index_x, int *index_y, int *index_z,
                                 const int *nlo_index, const int *nhi_index, int *n_lo, int *n_hi, int *index_outer,
                                 const int *nlo_index_outer, const int *nhi_index_outer, int *nlo_index_outer_plus,
                                 int *nhi_index_outer_plus, int *nlo_index_outer_minus, int *nhi_index_outer_minus,
                                 const LSMLIB_REAL *width, const LSMLIB_REAL *width_inner, const int *level);
/* replaces :138 */
void LSM3D_MARK_NARROW_BAND_BOUNDARY_LAYER(unsigned char *narrow_band, LSMHIP_BOX_ARGS(gb), LSMHIP_BOX_ARGS(fb),
                                           const unsigned char *mark_boundary_layer);
/* replaces :187 */
void LSM3D_DETERMINE_NARROW_BAND_FROM_MASK(const LSMLIB_REAL *mask, LSMHIP_BOX_ARGS(mask_gb),
                                           unsigned char *narrow_band, LSMHIP_BOX_ARGS(nb_gb), int *index_x,
                                           int *index_y, int *index_z, const int *nlo_index, const int *nhi_index,
                                           int *n_lo, int *n_hi, const int *level, const int *use_mask_sign);
/* replaces :230 */
void LSM3D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL(const LSMLIB_REAL *phi, LSMLIB_REAL *lse_rhs, LSMHIP_BOX_ARGS(gb),
                                          const int *index_x, const int *index_y, const int *index_z,
                                          const int *nlo_index, const int *nhi_index,
                                          const unsigned char *narrow_band, LSMHIP_BOX_ARGS(nb_gb),
                                          const unsigned char *mark_fb, const LSMLIB_REAL *beta,
                                          const LSMLIB_REAL *gamma);
/* replaces :270 */
void LSM3D_CHECK_OUTER_NARROW_BAND_LAYER(int *change_sign, const LSMLIB_REAL *phi, LSMHIP_BOX_ARGS(gb),
                                         const int *index_x, const int *index_y, const int *index_z,
                                         const int *nlo_index, const int *nhi_index, const int *index_outer,
                                         const int *nlo_index_outer, const int *nhi_index_outer,
                                         const int *nlo_index_outer_plus, const int *nhi_index_outer_plus,
                                         const int *nlo_index_outer_minus, const int *nhi_index_outer_minus);
/* replaces :307 */
void LSM3D_IMPOSE_MASK_LOCAL(LSMLIB_REAL *dest, LSMHIP_BOX_ARGS(dest_gb), const LSMLIB_REAL *src,
                             LSMHIP_BOX_ARGS(src_gb), const LSMLIB_REAL *mask, LSMHIP_BOX_ARGS(mask_gb),
                             const int *index_x, const int *index_y, const int *index_z, const int *nlo_index,
                             const int *nhi_index);
/* replaces :347 */
void LSM3D_COPY_DATA_LOCAL(LSMLIB_REAL *dest, LSMHIP_BOX_ARGS(dest_gb), const LSMLIB_REAL *src,
                           LSMHIP_BOX_ARGS(src_gb), const int *index_x, const int *index_y, const int *index_z,
                           const int *nlo_index, const int *nhi_index);

LSMHIP_END_DECLS

#endif
