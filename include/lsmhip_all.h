/* lsmhip_all.h -- every header of the HIP level-set library. */
#ifndef LSMHIP_ALL_H
#define LSMHIP_ALL_H
#include "lsm_boundary_conditions3d.h"
#include "lsm_calculus_toolbox.h"
#include "lsm_curvature3d.h"
#include "lsm_curvature3d_local.h"
#include "lsm_data_io.h"
#include "lsm_field_extension3d.h"
#include "lsm_geometry3d.h"
#include "lsm_level_set_evolution3d.h"
#include "lsm_level_set_evolution3d_local.h"
#include "lsm_localization3d.h"
#include "lsm_math_utils3d.h"
#include "lsm_math_utils3d_local.h"
#include "lsm_reinitialization3d.h"
#include "lsm_reinitialization3d_local.h"
#include "lsm_spatial_derivatives3d.h"
#include "lsm_spatial_derivatives3d_local.h"
#include "lsm_tvd_runge_kutta3d.h"
#include "lsm_tvd_runge_kutta3d_local.h"
#include "lsmhip.h"
#endif
