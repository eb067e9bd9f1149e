"""C-ABI table of the level-set time-stepping path and a ctypes binder for it.

Every entry point keeps the Fortran calling convention of the reference
(``src/*/lsm_*3d*.h``): all arguments are pointers, arrays are column-major and
addressed from the first element of their own ghost box, boxes are six ints
(ilo, ihi, jlo, jhi, klo, khi).  ``SPECS`` lists the arguments of each symbol
in order with a kind tag:

  R   real array                  (device pointer in liblsmhip)
  B   box: six ints               (host)
  Q   2D box: four ints (ilo, ihi, jlo, jhi)  (host)
  S   real scalar, input          (host)
  SO  real scalar, output         (host)
  I   int scalar, input           (host)
  IO  int scalar, output          (host)
  IH  int array on the host       (n_lo / n_hi level tables)
  IA  int array                   (device pointer in liblsmhip)
  M   uint8 array                 (device pointer in liblsmhip)
  U8  uint8 scalar, input         (host)

``bind(lib, real)`` returns an object whose attributes are callables taking the
arguments in spec order: arrays as raw addresses, numpy arrays or torch
tensors, boxes as 6-tuples, scalars as Python numbers; scalar outputs (SO, IO)
are passed as ``None`` and returned (a single value, or a tuple in order).
"""
from __future__ import annotations

import ctypes

_RAW = {
    # --- spatial derivatives, src/toolbox/lsm_spatial_derivatives3d.h ---------
    "lsm3dhjeno1_": "R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:phi B:phi R:D1 B:D1 B:fb S:dx S:dy S:dz",
    "lsm3dhjeno2_": "R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:phi B:phi R:D1 B:D1 R:D2 B:D2 B:fb S:dx S:dy S:dz",
    "lsm3dhjeno3_": "R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:phi B:phi R:D1 B:D1 R:D2 B:D2 R:D3 B:D3 B:fb "
                    "S:dx S:dy S:dz",
    "lsm3dhjweno5_": "R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:phi B:phi R:D1 B:D1 B:fb S:dx S:dy S:dz",
    "lsm3dupwindhjeno1_": "R:px R:py R:pz B:g R:phi B:phi R:vx R:vy R:vz B:vel R:D1 B:D1 B:fb S:dx S:dy S:dz",
    "lsm3dupwindhjeno2_": "R:px R:py R:pz B:g R:phi B:phi R:vx R:vy R:vz B:vel R:D1 B:D1 R:D2 B:D2 B:fb S:dx S:dy S:dz",
    "lsm3dupwindhjeno3_": "R:px R:py R:pz B:g R:phi B:phi R:vx R:vy R:vz B:vel R:D1 B:D1 R:D2 B:D2 R:D3 B:D3 B:fb "
                          "S:dx S:dy S:dz",
    "lsm3dupwindhjweno5_": "R:px R:py R:pz B:g R:phi B:phi R:vx R:vy R:vz B:vel R:D1 B:D1 B:fb S:dx S:dy S:dz",
    "lsm3dcentralgradorder2_": "R:px R:py R:pz B:g R:phi B:phi B:fb S:dx S:dy S:dz",
    "lsm3dcentralgradorder4_": "R:px R:py R:pz B:g R:phi B:phi B:fb S:dx S:dy S:dz",
    "lsm3dlaplacianorder2_": "R:lap B:lap R:phi B:phi B:fb S:dx S:dy S:dz",
    "lsm3daveragegradphi_": "R:px R:py R:pz B:g R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm B:fb",
    "lsm3dgradientmagnitude_": "R:px R:py R:pz R:mag B:g B:fb",
    # --- RHS terms, src/toolbox/lsm_level_set_evolution3d.h --------------------
    "lsm3dzerooutlevelseteqnrhs_": "R:rhs B:rhs",
    "lsm3daddadvectiontermtolserhs_": "R:rhs B:rhs R:px R:py R:pz B:g R:vx R:vy R:vz B:vel B:fb",
    "lsm3daddnormalveltermtolserhs_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:vn B:vel B:fb",
    "lsm3daddconstnormalveltermtolserhs_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm S:vn B:fb",
    "lsm3daddconstcurvtermtolserhs_": "R:rhs B:rhs R:px R:py R:pz B:g R:pxx R:pxy R:pxz R:pyy R:pyz R:pzz B:g2 S:b B:fb",
    "lsm3daddconstprecomputedcurvtermtolserhs_": "R:rhs B:rhs R:kappa B:kappa R:gmag B:gmag S:b B:fb",
    "lsm3daddexternalandnormalveltermtolserhs_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:vn "
                                                 "R:vx R:vy R:vz B:vel B:fb",
    # --- reinitialization / curvature ------------------------------------------
    "lsm3dcomputereinitializationeqnrhs_": "R:rhs B:rhs R:phi B:phi R:phi0 B:phi0 R:pxp R:pyp R:pzp B:gp R:pxm R:pym "
                                           "R:pzm B:gm B:fb S:dx S:dy S:dz I:use_phi0",
    "lsm3dcomputemeancurvatureorder2_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g B:fb S:dx S:dy S:dz",
    # --- normals and curvatures, src/geometry/lsm_geometry3d.h, lsm_curvature3d.h
    "lsm3dcomputeunitnormal_": "R:nx R:ny R:nz B:nrm R:px R:py R:pz B:g B:fb",
    "lsm3dcomputesignedunitnormal_": "R:nx R:ny R:nz B:nrm R:px R:py R:pz B:g R:phi B:phi B:fb S:dx S:dy S:dz",
    "lsm3dcomputemeancurvature_": "R:kappa B:kappa R:px R:py R:pz B:g R:pxx R:pxy R:pxz R:pyy R:pyz R:pzz B:g2 "
                                  "B:fb",
    "lsm3dcomputegaussiancurvatureorder2_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g B:fb "
                                            "S:dx S:dy S:dz",
    "lsm3dcomputemeancurvatureorder4_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g B:fb S:dx S:dy S:dz",
    "lsm3dcomputegaussiancurvatureorder4_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g B:fb "
                                            "S:dx S:dy S:dz",
    # --- orthogonalization, src/reinitialization/lsm_reinitialization3d.h ------
    "lsm3dcomputeorthogonalizationeqnrhs_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:psi "
                                            "B:psi R:qxp R:qyp R:qzp B:qp R:qxm R:qym R:qzm B:qm B:fb "
                                            "S:dx S:dy S:dz",
    # --- field extension, src/field_extension/lsm_field_extension3d.h ----------
    "lsm3dcomputefieldextensioneqnrhs_": "R:rhs B:rhs R:S B:S R:phi B:phi R:sx R:sy R:sz B:sg R:nx R:ny R:nz "
                                         "B:nrm B:fb S:dx S:dy S:dz",
    # --- TVD Runge-Kutta, src/toolbox/lsm_tvd_runge_kutta3d.h -----------------
    "lsm3drk1step_": "R:un B:un R:uc B:uc R:rhs B:rhs B:fb S:dt",
    "lsm3dtvdrk2stage1_": "R:un B:un R:uc B:uc R:rhs B:rhs B:fb S:dt",
    "lsm3dtvdrk2stage2_": "R:un B:un R:u1 B:u1 R:uc B:uc R:rhs B:rhs B:fb S:dt",
    "lsm3dtvdrk3stage1_": "R:un B:un R:uc B:uc R:rhs B:rhs B:fb S:dt",
    "lsm3dtvdrk3stage2_": "R:u2 B:u2 R:u1 B:u1 R:uc B:uc R:rhs B:rhs B:fb S:dt",
    "lsm3dtvdrk3stage3_": "R:un B:un R:u2 B:u2 R:uc B:uc R:rhs B:rhs B:fb S:dt",
    # --- reductions ---------------------------------------------------------------
    "lsm3dmaxnormdiff_": "SO:out R:f1 B:f1 R:f2 B:f2 B:ib",
    "lsm3daveabsdiff_": "SO:out R:f1 B:f1 R:f2 B:f2 B:ib",
    "lsm3dvoxelcountgreaterthanzero_": "IO:count R:phi B:phi B:fb",
    "lsm3dvoxelcountlessthanzero_": "IO:count R:phi B:phi B:fb",
    "lsm3dmaxnormdiffcontrolvolume_": "SO:out R:f1 B:f1 R:f2 B:f2 R:cv B:cv I:cvs B:ib",
    "lsm3dcomputestableadvectiondtcontrolvolume_": "SO:dt R:vx R:vy R:vz B:vel R:cv B:cv I:cvs B:ib S:dx S:dy "
                                                   "S:dz S:cfl",
    "lsm3dcomputestablenormalveldtcontrolvolume_": "SO:dt R:vn B:vel R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm "
                                                   "B:gm R:cv B:cv I:cvs B:ib S:dx S:dy S:dz S:cfl",
    "lsm3dcomputestableconstnormalveldtcontrolvolume_": "SO:dt S:vn R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm "
                                                        "B:gm R:cv B:cv I:cvs B:ib S:dx S:dy S:dz S:cfl",
    "lsm3dvolumeintegralphilessthanzerocontrolvolume_": "SO:int R:F B:F R:phi B:phi R:cv B:cv I:cvs B:ib S:dx "
                                                        "S:dy S:dz S:eps",
    "lsm3dvolumeintegralphigreaterthanzerocontrolvolume_": "SO:int R:F B:F R:phi B:phi R:cv B:cv I:cvs B:ib S:dx "
                                                           "S:dy S:dz S:eps",
    "lsm3dsurfaceintegralcontrolvolume_": "SO:int R:F B:F R:phi B:phi R:px R:py R:pz B:g R:cv B:cv I:cvs B:ib "
                                          "S:dx S:dy S:dz S:eps",
    "lsm3dvoxelcountgreaterthanzerocontrolvolume_": "IO:count R:phi B:phi R:cv B:cv I:cvs B:fb",
    "lsm3dvoxelcountlessthanzerocontrolvolume_": "IO:count R:phi B:phi R:cv B:cv I:cvs B:fb",
    "lsm3dsurfaceareazerolevelsetdelta_": "SO:area R:dphi B:phi R:px R:py R:pz B:g B:ib S:dx S:dy S:dz",
    "lsm3dvolumeregionphilessthanzerocontrolvolume_": "SO:vol R:phi B:phi R:cv B:cv I:cvs B:ib S:dx S:dy S:dz "
                                                      "S:eps",
    "lsm3dvolumeregionphigreaterthanzerocontrolvolume_": "SO:vol R:phi B:phi R:cv B:cv I:cvs B:ib S:dx S:dy "
                                                         "S:dz S:eps",
    "lsm3dsurfaceareazerolevelsetcontrolvolume_": "SO:area R:phi B:phi R:px R:py R:pz B:g R:cv B:cv I:cvs B:ib "
                                                  "S:dx S:dy S:dz S:eps",
    "lsm3dsurfaceareazerolevelsetdeltacontrolvolume_": "SO:area R:dphi B:phi R:px R:py R:pz B:g R:cv B:cv I:cvs "
                                                       "B:ib S:dx S:dy S:dz",
    "lsm3dphiupwindgradf_": "R:Fx R:Fy R:Fz B:g R:Fxp R:Fyp R:Fzp B:gp R:Fxm R:Fym R:Fzm B:gm R:phi B:phi B:fb",
    "lsm3dcentralhessian_": "R:pxx R:pxy R:pxz R:pyy R:pyz R:pzz B:g2 R:phi B:phi B:fb S:dx S:dy S:dz",
    "lsm3ddeltafunctionorder1_": "R:phi R:delta B:phi R:nx R:ny R:nz B:g B:fb S:dx S:dy S:dz",
    "lsm3dcomputestableadvectiondt_": "SO:dt R:vx R:vy R:vz B:vel B:ib S:dx S:dy S:dz S:cfl",
    "lsm3dcomputestablenormalveldt_": "SO:dt R:vn B:vel R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm B:ib "
                                      "S:dx S:dy S:dz S:cfl",
    "lsm3dcomputestableconstnormalveldt_": "SO:dt S:vn R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm B:ib "
                                           "S:dx S:dy S:dz S:cfl",
    "lsm3dvolumeintegralphilessthanzero_": "SO:int R:F B:F R:phi B:phi B:ib S:dx S:dy S:dz S:eps",
    "lsm3dvolumeintegralphigreaterthanzero_": "SO:int R:F B:F R:phi B:phi B:ib S:dx S:dy S:dz S:eps",
    "lsm3dsurfaceintegral_": "SO:int R:F B:F R:phi B:phi R:px R:py R:pz B:g B:ib S:dx S:dy S:dz S:eps",
    "lsm3dvolumeregionphilessthanzero_": "SO:vol R:phi B:phi B:ib S:dx S:dy S:dz S:eps",
    "lsm3dvolumeregionphigreaterthanzero_": "SO:vol R:phi B:phi B:ib S:dx S:dy S:dz S:eps",
    "lsm3dsurfaceareazerolevelset_": "SO:area R:phi B:phi R:px R:py R:pz B:g B:ib S:dx S:dy S:dz S:eps",
    # --- ghost fills, src/boundary_conditions/lsm_boundary_conditions3d.h -------
    "lsm3dlinearextrapolation_": "R:phi B:gb B:fb I:idx",
    "lsm3dsignedlinearextrapolation_": "R:phi B:gb B:fb I:idx",
    "lsm3dcopyextrapolation_": "R:phi B:gb B:fb I:idx",
    "lsm3dhomogeneousneumanneno1_": "R:phi B:gb B:fb I:idx",
    "lsm3dhomogeneousneumanneno2_": "R:phi B:gb B:fb I:idx",
    "lsm3dhomogeneousneumanneno3_": "R:phi B:gb B:fb I:idx",
    "lsm3dhomogeneousneumannweno5_": "R:phi B:gb B:fb I:idx",
    # --- smoothed Heaviside / delta, src/toolbox/lsm_calculus_toolbox.f --------
    "lsmheaviside_": "SO:H S:x S:eps",
    "lsmdeltafunction_": "SO:delta S:x S:eps",
    # --- narrow-band derivatives, src/toolbox/lsm_spatial_derivatives3d_local.h -
    "lsm3dhjeno1local_": "R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:phi B:phi R:D1 B:D1 S:dx S:dy S:dz "
                         "IA:ix IA:iy IA:iz I:nlo0 I:nhi0 I:nlo1 I:nhi1 M:nb B:nb U8:mark_fb U8:mark_D1",
    "lsm3dhjeno2local_": "R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:phi B:phi R:D1 B:D1 R:D2 B:D2 "
                         "S:dx S:dy S:dz IA:ix IA:iy IA:iz I:nlo0 I:nhi0 I:nlo1 I:nhi1 I:nlo2 I:nhi2 M:nb B:nb "
                         "U8:mark_fb U8:mark_D1 U8:mark_D2",
    "lsm3dhjweno5local_": "R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:phi B:phi R:D1 B:D1 S:dx S:dy S:dz "
                          "IA:ix IA:iy IA:iz I:nlo0 I:nhi0 I:nlo1 I:nhi1 I:nlo2 I:nhi2 I:nlo3 I:nhi3 M:nb B:nb "
                          "U8:mark_fb U8:mark_D1",
    "lsm3dcomputegaussiancurvatureorder2local_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g S:dx S:dy "
                                                 "S:dz IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputemeancurvatureorder4local_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g S:dx S:dy "
                                             "S:dz IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputegaussiancurvatureorder4local_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g S:dx S:dy "
                                                 "S:dz IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputeorthogonalizationeqnrhslocal_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:psi "
                                                 "B:psi R:qxp R:qyp R:qzp B:qp R:qxm R:qym R:qzm B:qm S:dx S:dy S:dz "
                                                 "IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcentralgradorder2local_": "R:px R:py R:pz B:g R:phi B:phi S:dx S:dy S:dz IA:ix IA:iy IA:iz I:nlo I:nhi "
                                    "M:nb B:nb U8:mark_fb",
    "lsm3dcentralgradorder4local_": "R:px R:py R:pz B:g R:phi B:phi S:dx S:dy S:dz IA:ix IA:iy IA:iz I:nlo I:nhi "
                                    "M:nb B:nb U8:mark_fb",
    "lsm3dlaplacianorder2local_": "R:lap B:lap R:phi B:phi S:dx S:dy S:dz IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb "
                                  "U8:mark_fb",
    "lsm3dcomputeavegradphilocal_": "SO:ave R:phi B:phi S:dx S:dy S:dz IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb "
                                    "U8:mark_fb",
    "lsm3dgradientmagnitudelocal_": "R:px R:py R:pz R:mag B:g IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    # --- narrow-band RHS, src/toolbox/lsm_level_set_evolution3d_local.h --------
    "lsm3dzerooutlevelseteqnrhslocal_": "R:rhs B:rhs IA:ix IA:iy IA:iz I:nlo I:nhi",
    "lsm3daddadvectiontermtolserhslocal_": "R:rhs B:rhs R:px R:py R:pz B:g R:vx R:vy R:vz B:vel IA:ix IA:iy IA:iz "
                                           "I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3daddnormalveltermtolserhslocal_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm R:vn B:vel "
                                           "IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3daddconstnormalveltermtolserhslocal_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm S:vn "
                                                "IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3daddconstcurvtermtolserhslocal_": "R:rhs B:rhs R:px R:py R:pz B:g R:pxx R:pxy R:pxz R:pyy R:pyz R:pzz B:g2 "
                                           "S:b IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3daddconstprecomputedcurvtermtolserhslocal_": "R:rhs B:rhs R:kappa B:kappa R:gmag B:gmag S:b IA:ix IA:iy "
                                                      "IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3daddexternalandnormalveltermtolserhslocal_": "R:rhs B:rhs R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm "
                                                      "R:vn R:vx R:vy R:vz B:vel IA:ix IA:iy IA:iz I:nlo I:nhi M:nb "
                                                      "B:nb U8:mark_fb",
    # --- narrow-band Runge-Kutta, src/toolbox/lsm_tvd_runge_kutta3d_local.h ----
    "lsm3drk1steplocal_": "R:un B:un R:uc B:uc R:rhs B:rhs S:dt IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dtvdrk2stage1local_": "R:un B:un R:uc B:uc R:rhs B:rhs S:dt IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb "
                               "U8:mark_fb",
    "lsm3dtvdrk2stage2local_": "R:un B:un R:u1 B:u1 R:uc B:uc R:rhs B:rhs S:dt IA:ix IA:iy IA:iz I:nlo I:nhi M:nb "
                               "B:nb U8:mark_fb",
    "lsm3dtvdrk3stage1local_": "R:un B:un R:uc B:uc R:rhs B:rhs S:dt IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb "
                               "U8:mark_fb",
    "lsm3dtvdrk3stage2local_": "R:u2 B:u2 R:u1 B:u1 R:uc B:uc R:rhs B:rhs S:dt IA:ix IA:iy IA:iz I:nlo I:nhi M:nb "
                               "B:nb U8:mark_fb",
    "lsm3dtvdrk3stage3local_": "R:un B:un R:u2 B:u2 R:uc B:uc R:rhs B:rhs S:dt IA:ix IA:iy IA:iz I:nlo I:nhi M:nb "
                               "B:nb U8:mark_fb",
    # --- narrow-band reductions, src/toolbox/lsm_math_utils3d_local.h ----------
    "lsm3dmaxnormdifflocal_": "SO:out R:f1 B:f1 R:f2 B:f2 IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputestableadvectiondtlocal_": "SO:dt R:vx R:vy R:vz B:vel S:dx S:dy S:dz S:cfl IA:ix IA:iy IA:iz I:nlo "
                                           "I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputestablenormalveldtlocal_": "SO:dt R:vn B:vel R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm S:dx "
                                           "S:dy S:dz S:cfl IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputestableconstnormalveldtlocal_": "SO:dt S:vn R:pxp R:pyp R:pzp B:gp R:pxm R:pym R:pzm B:gm S:dx "
                                                "S:dy S:dz S:cfl IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputereinitializationeqnrhslocal_": "R:rhs B:rhs R:phi B:phi R:phi0 B:phi0 R:pxp R:pyp R:pzp B:gp "
                                                "R:pxm R:pym R:pzm B:gm S:dx S:dy S:dz I:use_phi0 IA:ix IA:iy "
                                                "IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    "lsm3dcomputemeancurvatureorder2local_": "R:kappa B:kappa R:phi B:phi R:px R:py R:pz R:gmag B:g S:dx S:dy S:dz "
                                             "IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb",
    # --- band construction, src/toolbox/lsm_localization3d.h --------------------
    "lsm3ddeterminenarrowband_": "R:phi B:phi M:nb B:nb IA:ix IA:iy IA:iz I:nlo I:nhi IH:n_lo IH:n_hi IA:outer "
                                 "I:nlo_outer I:nhi_outer IO:nlo_outer_plus IO:nhi_outer_plus IO:nlo_outer_minus "
                                 "IO:nhi_outer_minus S:width S:width_inner I:level",
    "lsm3dmarknarrowbandboundarylayer_": "M:nb B:gb B:fb U8:mark",
    "lsm3ddeterminenarrowbandfrommask_": "R:mask B:mask M:nb B:nb IA:ix IA:iy IA:iz I:nlo I:nhi IH:n_lo IH:n_hi "
                                         "I:level I:use_mask_sign",
    "lsm3dmultiplycutofflserhslocal_": "R:phi R:rhs B:gb IA:ix IA:iy IA:iz I:nlo I:nhi M:nb B:nb U8:mark_fb S:beta "
                                       "S:gamma",
    "lsm3dcheckouternarrowbandlayer_": "IO:change_sign R:phi B:gb IA:ix IA:iy IA:iz I:nlo I:nhi IA:outer "
                                       "I:nlo_outer I:nhi_outer I:nlo_outer_plus I:nhi_outer_plus "
                                       "I:nlo_outer_minus I:nhi_outer_minus",
    "lsm3dimposemasklocal_": "R:dest B:dest R:src B:src R:mask B:mask IA:ix IA:iy IA:iz I:nlo I:nhi",
    "lsm3dcopydatalocal_": "R:dest B:dest R:src B:src IA:ix IA:iy IA:iz I:nlo I:nhi",
    # --- 2D twins (src/*/lsm_*2d.h); arrays over (i, j), boxes of four ints ----
    "lsm2dhjeno1_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 Q:fb S:dx S:dy",
    "lsm2dhjeno2_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 R:D2 Q:D2 Q:fb S:dx S:dy",
    "lsm2dhjeno3_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 R:D2 Q:D2 R:D3 Q:D3 Q:fb S:dx S:dy",
    "lsm2dhjweno5_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 Q:fb S:dx S:dy",
    "lsm2dupwindhjeno1_": "R:px R:py Q:g R:phi Q:phi R:vx R:vy Q:vel R:D1 Q:D1 Q:fb S:dx S:dy",
    "lsm2dupwindhjeno2_": "R:px R:py Q:g R:phi Q:phi R:vx R:vy Q:vel R:D1 Q:D1 R:D2 Q:D2 Q:fb S:dx S:dy",
    "lsm2dupwindhjeno3_": "R:px R:py Q:g R:phi Q:phi R:vx R:vy Q:vel R:D1 Q:D1 R:D2 Q:D2 R:D3 Q:D3 Q:fb S:dx S:dy",
    "lsm2dupwindhjweno5_": "R:px R:py Q:g R:phi Q:phi R:vx R:vy Q:vel R:D1 Q:D1 Q:fb S:dx S:dy",
    "lsm2dcentralgradorder2_": "R:px R:py Q:g R:phi Q:phi Q:fb S:dx S:dy",
    "lsm2dcentralgradorder4_": "R:px R:py Q:g R:phi Q:phi Q:fb S:dx S:dy",
    "lsm2dlaplacianorder2_": "R:lap Q:lap R:phi Q:phi Q:fb S:dx S:dy",
    "lsm2daveragegradphi_": "R:px R:py Q:g R:pxp R:pyp Q:gp R:pxm R:pym Q:gm Q:fb",
    "lsm2dgradientmagnitude_": "R:px R:py R:mag Q:g Q:fb",
    "lsm2dzerooutlevelseteqnrhs_": "R:rhs Q:rhs",
    "lsm2daddadvectiontermtolserhs_": "R:rhs Q:rhs R:px R:py Q:g R:vx R:vy Q:vel Q:fb",
    "lsm2daddnormalveltermtolserhs_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:vn Q:vel Q:fb",
    "lsm2daddconstnormalveltermtolserhs_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm S:vn Q:fb",
    "lsm2daddconstcurvtermtolserhs_": "R:rhs Q:rhs R:px R:py Q:g R:pxx R:pxy R:pyy Q:g2 S:b Q:fb",
    "lsm2daddconstprecomputedcurvtermtolserhs_": "R:rhs Q:rhs R:kappa Q:kappa R:gmag Q:gmag S:b Q:fb",
    "lsm2daddexternalandnormalveltermtolserhs_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:vn R:vx R:vy "
                                                 "Q:vel Q:fb",
    "lsm2drk1step_": "R:un Q:un R:uc Q:uc R:rhs Q:rhs Q:fb S:dt",
    "lsm2dtvdrk2stage1_": "R:un Q:un R:uc Q:uc R:rhs Q:rhs Q:fb S:dt",
    "lsm2dtvdrk2stage2_": "R:un Q:un R:u1 Q:u1 R:uc Q:uc R:rhs Q:rhs Q:fb S:dt",
    "lsm2dtvdrk3stage1_": "R:un Q:un R:uc Q:uc R:rhs Q:rhs Q:fb S:dt",
    "lsm2dtvdrk3stage2_": "R:u2 Q:u2 R:u1 Q:u1 R:uc Q:uc R:rhs Q:rhs Q:fb S:dt",
    "lsm2dtvdrk3stage3_": "R:un Q:un R:u2 Q:u2 R:uc Q:uc R:rhs Q:rhs Q:fb S:dt",
    "lsm2dcomputereinitializationeqnrhs_": "R:rhs Q:rhs R:phi Q:phi R:phi0 Q:phi0 R:pxp R:pyp Q:gp R:pxm R:pym Q:gm "
                                           "Q:fb S:dx S:dy I:use_phi0",
    "lsm2dcomputeorthogonalizationeqnrhs_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:psi Q:psi R:qxp R:qyp "
                                            "Q:qp R:qxm R:qym Q:qm Q:fb S:dx S:dy",
    "lsm2dcomputemeancurvatureorder2_": "R:kappa Q:kappa R:phi Q:phi R:px R:py R:gmag Q:g Q:fb S:dx S:dy",
    "lsm2dcomputemeancurvatureorder4_": "R:kappa Q:kappa R:phi Q:phi R:px R:py R:gmag Q:g Q:fb S:dx S:dy",
    "lsm2dcomputeunitnormal_": "R:nx R:ny Q:nrm R:px R:py Q:g Q:fb",
    "lsm2dcomputesignedunitnormal_": "R:nx R:ny Q:nrm R:px R:py Q:g R:phi Q:phi Q:fb S:dx S:dy",
    "lsm2dcomputemeancurvature_": "R:kappa Q:kappa R:px R:py Q:g R:pxx R:pxy R:pyy Q:g2 Q:fb",
    "lsm2dcomputefieldextensioneqnrhs_": "R:rhs Q:rhs R:S Q:S R:phi Q:phi R:sx R:sy Q:sg R:nx R:ny Q:nrm Q:fb "
                                         "S:dx S:dy",
    "lsm2dmaxnormdiff_": "SO:out R:f1 Q:f1 R:f2 Q:f2 Q:ib",
    "lsm2dcomputestableadvectiondt_": "SO:dt R:vx R:vy Q:vel Q:ib S:dx S:dy S:cfl",
    "lsm2dcomputestablenormalveldt_": "SO:dt R:vn Q:vel R:pxp R:pyp Q:gp R:pxm R:pym Q:gm Q:ib S:dx S:dy S:cfl",
    "lsm2dcomputestableconstnormalveldt_": "SO:dt S:vn R:pxp R:pyp Q:gp R:pxm R:pym Q:gm Q:ib S:dx S:dy S:cfl",
    "lsm2dvolumeintegralphilessthanzero_": "SO:int R:F Q:F R:phi Q:phi Q:ib S:dx S:dy S:eps",
    "lsm2dvolumeintegralphigreaterthanzero_": "SO:int R:F Q:F R:phi Q:phi Q:ib S:dx S:dy S:eps",
    "lsm2dsurfaceintegral_": "SO:int R:F Q:F R:phi Q:phi R:px R:py Q:g Q:ib S:dx S:dy S:eps",
    "lsm2darearegionphilessthanzero_": "SO:area R:phi Q:phi Q:ib S:dx S:dy S:eps",
    "lsm2darearegionphigreaterthanzero_": "SO:area R:phi Q:phi Q:ib S:dx S:dy S:eps",
    "lsm2dperimeterzerolevelset_": "SO:perim R:phi Q:phi R:px R:py Q:g Q:ib S:dx S:dy S:eps",
    "lsm2dlinearextrapolation_": "R:phi Q:gb Q:fb I:idx",
    "lsm2dsignedlinearextrapolation_": "R:phi Q:gb Q:fb I:idx",
    "lsm2dcopyextrapolation_": "R:phi Q:gb Q:fb I:idx",
    "lsm2dhomogeneousneumanneno1_": "R:phi Q:gb Q:fb I:idx",
    "lsm2dhomogeneousneumanneno2_": "R:phi Q:gb Q:fb I:idx",
    "lsm2dhomogeneousneumanneno3_": "R:phi Q:gb Q:fb I:idx",
    "lsm2dhomogeneousneumannweno5_": "R:phi Q:gb Q:fb I:idx",
    # --- 2D narrow-band twins (src/*/lsm_*2d_local.h, lsm_localization2d.h) -----
    "lsm2dhjeno1local_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 S:dx S:dy IA:ix IA:iy "
                         "I:nlo0 I:nhi0 I:nlo1 I:nhi1 M:nb Q:nb U8:mark_fb U8:mark_D1",
    "lsm2dhjeno2local_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 R:D2 Q:D2 S:dx S:dy IA:ix IA:iy "
                         "I:nlo0 I:nhi0 I:nlo1 I:nhi1 I:nlo2 I:nhi2 M:nb Q:nb U8:mark_fb U8:mark_D1 U8:mark_D2",
    "lsm2dhjeno3local_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 R:D2 Q:D2 R:D3 Q:D3 S:dx S:dy "
                         "IA:ix IA:iy I:nlo0 I:nhi0 I:nlo1 I:nhi1 I:nlo2 I:nhi2 I:nlo3 I:nhi3 M:nb Q:nb "
                         "U8:mark_fb U8:mark_D1 U8:mark_D2 U8:mark_D3",
    "lsm2dhjweno5local_": "R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:phi Q:phi R:D1 Q:D1 S:dx S:dy IA:ix IA:iy "
                          "I:nlo0 I:nhi0 I:nlo1 I:nhi1 I:nlo2 I:nhi2 I:nlo3 I:nhi3 M:nb Q:nb U8:mark_fb U8:mark_D1",
    "lsm2dupwindhjeno2local_": "R:px R:py Q:g R:phi Q:phi R:vx R:vy Q:vel R:D1 Q:D1 R:D2 Q:D2 S:dx S:dy IA:ix IA:iy "
                               "I:nlo0 I:nhi0 I:nlo1 I:nhi1 I:nlo2 I:nhi2 M:nb Q:nb U8:mark_fb U8:mark_D1 "
                               "U8:mark_D2",
    "lsm2dcentralgradorder2local_": "R:px R:py Q:g R:phi Q:phi S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcentralgradorder4local_": "R:px R:py Q:g R:phi Q:phi S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dlaplacianorder2local_": "R:lap Q:lap R:phi Q:phi S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputeavegradphilocal_": "SO:ave R:phi Q:phi S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dgradientmagnitudelocal_": "R:px R:py R:mag Q:g IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dzerooutlevelseteqnrhslocal_": "R:rhs Q:rhs IA:ix IA:iy I:nlo I:nhi",
    "lsm2daddadvectiontermtolserhslocal_": "R:rhs Q:rhs R:px R:py Q:g R:vx R:vy Q:vel IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2daddnormalveltermtolserhslocal_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:vn Q:vel IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2daddconstnormalveltermtolserhslocal_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm S:vn IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2daddconstcurvtermtolserhslocal_": "R:rhs Q:rhs R:px R:py Q:g R:pxx R:pxy R:pyy Q:g2 S:b IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2daddconstprecomputedcurvtermtolserhslocal_": "R:rhs Q:rhs R:kappa Q:kappa R:gmag Q:gmag S:b IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2daddexternalandnormalveltermtolserhslocal_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:vn R:vx R:vy "
                                                      "Q:vel IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2drk1steplocal_": "R:un Q:un R:uc Q:uc R:rhs Q:rhs S:dt IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dtvdrk2stage1local_": "R:un Q:un R:uc Q:uc R:rhs Q:rhs S:dt IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dtvdrk2stage2local_": "R:un Q:un R:u1 Q:u1 R:uc Q:uc R:rhs Q:rhs S:dt IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dtvdrk3stage1local_": "R:un Q:un R:uc Q:uc R:rhs Q:rhs S:dt IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dtvdrk3stage2local_": "R:u2 Q:u2 R:u1 Q:u1 R:uc Q:uc R:rhs Q:rhs S:dt IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dtvdrk3stage3local_": "R:un Q:un R:u2 Q:u2 R:uc Q:uc R:rhs Q:rhs S:dt IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dmaxnormdifflocal_": "SO:out R:f1 Q:f1 R:f2 Q:f2 IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputestableadvectiondtlocal_": "SO:dt R:vx R:vy Q:vel S:dx S:dy S:cfl IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputestablenormalveldtlocal_": "SO:dt R:vn Q:vel R:pxp R:pyp Q:gp R:pxm R:pym Q:gm S:dx S:dy S:cfl "
                                           "IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputestableconstnormalveldtlocal_": "SO:dt S:vn R:pxp R:pyp Q:gp R:pxm R:pym Q:gm S:dx S:dy S:cfl "
                                                "IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputereinitializationeqnrhslocal_": "R:rhs Q:rhs R:phi Q:phi R:phi0 Q:phi0 R:pxp R:pyp Q:gp R:pxm R:pym "
                                                "Q:gm S:dx S:dy I:use_phi0 IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputeorthogonalizationeqnrhslocal_": "R:rhs Q:rhs R:pxp R:pyp Q:gp R:pxm R:pym Q:gm R:psi Q:psi R:qxp "
                                                 "R:qyp Q:qp R:qxm R:qym Q:qm S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputemeancurvatureorder2local_": "R:kappa Q:kappa R:phi Q:phi R:px R:py R:gmag Q:g S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputemeancurvatureorder4local_": "R:kappa Q:kappa R:phi Q:phi R:px R:py R:gmag Q:g S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputeunitnormallocal_": "R:nx R:ny Q:nrm R:px R:py Q:g S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2dcomputesignedunitnormallocal_": "R:nx R:ny Q:nrm R:px R:py Q:g R:phi Q:phi S:dx S:dy IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb",
    "lsm2ddeterminenarrowband_": "R:phi Q:phi M:nb Q:nb IA:ix IA:iy I:nlo I:nhi IH:n_lo IH:n_hi IA:outer "
                                 "I:nlo_outer I:nhi_outer IO:nlo_outer_plus IO:nhi_outer_plus IO:nlo_outer_minus "
                                 "IO:nhi_outer_minus S:width S:width_inner I:level",
    "lsm2dmarknarrowbandboundarylayer_": "M:nb Q:gb Q:fb U8:mark",
    "lsm2ddeterminenarrowbandfrommask_": "R:mask Q:mask M:nb Q:nb IA:ix IA:iy I:nlo I:nhi IH:n_lo IH:n_hi "
                                         "I:level I:use_mask_sign",
    "lsm2dmultiplycutofflserhslocal_": "R:phi R:rhs Q:gb IA:ix IA:iy I:nlo I:nhi M:nb Q:nb U8:mark_fb S:beta S:gamma",
    "lsm2dcheckouternarrowbandlayer_": "IO:change_sign R:phi Q:gb IA:ix IA:iy I:nlo I:nhi IA:outer "
                                       "I:nlo_outer I:nhi_outer I:nlo_outer_plus I:nhi_outer_plus "
                                       "I:nlo_outer_minus I:nhi_outer_minus",
    "lsm2dimposemasklocal_": "R:dest Q:dest R:src Q:src R:mask Q:mask IA:ix IA:iy I:nlo I:nhi",
    "lsm2dcopydatalocal_": "R:dest Q:dest R:src Q:src IA:ix IA:iy I:nlo I:nhi",
}

# Entry points with no reference counterpart (periodic fill is absent from the
# serial reference; the 3D WENO5 local routine is authored from the 2D one).
AUTHORED = {"lsmhip_periodic_face_", "lsm3dhjweno5local_"}
_RAW["lsmhip_periodic_face_"] = "R:phi B:gb B:fb I:idx"


def _parse(spec: str):
    out = []
    for tok in spec.split():
        kind, name = tok.split(":")
        out.append((kind, name))
    return out


SPECS = {name: _parse(s) for name, s in _RAW.items()}


def _addr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):  # torch tensor
        return x.data_ptr()
    if hasattr(x, "ctypes"):  # numpy array
        return x.ctypes.data
    raise TypeError("cannot take the address of %r" % type(x))


class _Fn:
    def __init__(self, cfn, spec, real):
        self._f = cfn
        self._spec = spec
        self._real = real

    def __call__(self, *args):
        if len(args) != len(self._spec):
            raise TypeError("expected %d arguments, got %d" % (len(self._spec), len(args)))
        keep, cargs, outs = [], [], []
        for (kind, name), a in zip(self._spec, args):
            if kind in ("B", "Q"):
                if len(a) != (6 if kind == "B" else 4):
                    raise ValueError("box %s needs %d ints" % (name, 6 if kind == "B" else 4))
                for v in a:
                    c = ctypes.c_int(int(v))
                    keep.append(c)
                    cargs.append(ctypes.byref(c))
            elif kind == "S":
                c = self._real(float(a))
                keep.append(c)
                cargs.append(ctypes.byref(c))
            elif kind == "SO":
                c = self._real(0.0)
                keep.append(c)
                outs.append(c)
                cargs.append(ctypes.byref(c))
            elif kind == "I":
                c = ctypes.c_int(int(a))
                keep.append(c)
                cargs.append(ctypes.byref(c))
            elif kind == "IO":
                c = ctypes.c_int(0 if a is None else int(a))
                keep.append(c)
                outs.append(c)
                cargs.append(ctypes.byref(c))
            elif kind == "U8":
                c = ctypes.c_ubyte(int(a))
                keep.append(c)
                cargs.append(ctypes.byref(c))
            else:  # R, IA, IH, M: addresses
                ptr = _addr(a)
                if not ptr:  # every array of the table is required; a NULL one would be dereferenced
                    raise ValueError("array argument %s is NULL" % name)
                cargs.append(ctypes.c_void_p(ptr))
        self._f(*cargs)
        vals = tuple(o.value for o in outs)
        if not vals:
            return None
        return vals[0] if len(vals) == 1 else vals


class Bound:
    """Callables for every SPECS symbol a library exports."""

    def __init__(self, lib, real=ctypes.c_double, prefix="", strip_underscore=False, names=None):
        self.lib = lib
        self.missing = []
        for sym, spec in SPECS.items():
            if names is not None and sym not in names:
                continue
            cname = prefix + (sym[:-1] if strip_underscore else sym)
            try:
                cfn = getattr(lib, cname)
            except AttributeError:
                self.missing.append(sym)
                continue
            cfn.argtypes = None
            cfn.restype = None
            setattr(self, sym.rstrip("_"), _Fn(cfn, spec, real))


def arg_count(sym):
    return sum({"B": 6, "Q": 4}.get(k, 1) for k, _ in SPECS[sym])
