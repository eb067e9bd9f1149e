/*
 * lsm_localization2d.h -- 2D narrow-band construction, checks, cut-off, mask and copy on the GPU
 * (reference: src/toolbox/lsm_localization2d.h).
 */
#ifndef LSMHIP_LSM_LOCALIZATION2D_H
#define LSMHIP_LSM_LOCALIZATION2D_H

#include "lsmhip_config.h"

#define LSM2D_DETERMINE_NARROW_BAND lsm2ddeterminenarrowband_
#define LSM2D_MARK_NARROW_BAND_BOUNDARY_LAYER lsm2dmarknarrowbandboundarylayer_
#define LSM2D_DETERMINE_NARROW_BAND_FROM_MASK lsm2ddeterminenarrowbandfrommask_
#define LSM2D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL lsm2dmultiplycutofflserhslocal_
#define LSM2D_CHECK_OUTER_NARROW_BAND_LAYER lsm2dcheckouternarrowbandlayer_
#define LSM2D_IMPOSE_MASK_LOCAL lsm2dimposemasklocal_
#define LSM2D_COPY_DATA_LOCAL lsm2dcopydatalocal_

LSMHIP_BEGIN_DECLS

/* replaces src/toolbox/lsm_localization2d.h:94 */
void LSM2D_DETERMINE_NARROW_BAND(const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb), unsigned char *narrow_band,
    LSMHIP_BOX2_ARGS(nb_gb), int *index_x, int *index_y, const int *nlo_index, const int *nhi_index, int *n_lo,
    int *n_hi, int *index_outer, const int *nlo_index_outer, const int *nhi_index_outer, int *nlo_index_outer_plus,
    int *nhi_index_outer_plus, int *nlo_index_outer_minus, int *nhi_index_outer_minus, const LSMLIB_REAL *width,
    const LSMLIB_REAL *width_inner, const int *level);
/* replaces src/toolbox/lsm_localization2d.h:283 */
void LSM2D_MARK_NARROW_BAND_BOUNDARY_LAYER(unsigned char *narrow_band, LSMHIP_BOX2_ARGS(gb), LSMHIP_BOX2_ARGS(fb),
    const unsigned char *mark_boundary_layer);
/* replaces src/toolbox/lsm_localization2d.h:328 */
void LSM2D_DETERMINE_NARROW_BAND_FROM_MASK(const LSMLIB_REAL *mask, LSMHIP_BOX2_ARGS(mask_gb),
    unsigned char *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), int *index_x, int *index_y, const int *nlo_index,
    const int *nhi_index, int *n_lo, int *n_hi, const int *level, const int *use_mask_sign);
/* replaces src/toolbox/lsm_localization2d.h:367 */
void LSM2D_MULTIPLY_CUT_OFF_LSE_RHS_LOCAL(const LSMLIB_REAL *phi, LSMLIB_REAL *lse_rhs, LSMHIP_BOX2_ARGS(gb),
    const int *index_x, const int *index_y, const int *nlo_index, const int *nhi_index,
    const unsigned char *narrow_band, LSMHIP_BOX2_ARGS(nb_gb), const unsigned char *mark_fb,
    const LSMLIB_REAL *beta, const LSMLIB_REAL *gamma);
/* replaces src/toolbox/lsm_localization2d.h:403 */
void LSM2D_CHECK_OUTER_NARROW_BAND_LAYER(int *change_sign, const LSMLIB_REAL *phi, LSMHIP_BOX2_ARGS(gb),
    const int *index_x, const int *index_y, const int *nlo_index, const int *nhi_index, const int *index_outer,
    const int *nlo_index_outer, const int *nhi_index_outer, const int *nlo_index_outer_plus,
    const int *nhi_index_outer_plus, const int *nlo_index_outer_minus, const int *nhi_index_outer_minus);
/* replaces src/toolbox/lsm_localization2d.h:421 */
void LSM2D_IMPOSE_MASK_LOCAL(LSMLIB_REAL *dest, LSMHIP_BOX2_ARGS(dest_gb), const LSMLIB_REAL *src,
    LSMHIP_BOX2_ARGS(src_gb), const LSMLIB_REAL *mask, LSMHIP_BOX2_ARGS(mask_gb), const int *index_x,
    const int *index_y, const int *nlo_index, const int *nhi_index);
/* replaces src/toolbox/lsm_localization2d.h:491 */
void LSM2D_COPY_DATA_LOCAL(LSMLIB_REAL *dest, LSMHIP_BOX2_ARGS(dest_gb), const LSMLIB_REAL *src,
    LSMHIP_BOX2_ARGS(src_gb), const int *index_x, const int *index_y, const int *nlo_index, const int *nhi_index);

LSMHIP_END_DECLS

#endif
