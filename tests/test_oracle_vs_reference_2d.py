"""Pin the 2D restatement (oracle/oracle_2d.c) bitwise against the reference's
own 2D Fortran kernels, compiled by oracle/Makefile into oracle/_ref/.

Runs on the CPU.  Skipped when the reference build is absent.
"""
import numpy as np
import pytest

import harness as H
from lsmlib_amd import capi

pytestmark = pytest.mark.skipif(H.reference() is None, reason="oracle/_ref/liblsm_ref.so not built")

WHOLE_GRID_2D = [
    "lsm2dhjeno1_", "lsm2dhjeno2_", "lsm2dhjeno3_", "lsm2dhjweno5_",
    "lsm2dupwindhjeno1_", "lsm2dupwindhjeno2_", "lsm2dupwindhjeno3_", "lsm2dupwindhjweno5_",
    "lsm2dcentralgradorder2_", "lsm2dcentralgradorder4_", "lsm2dlaplacianorder2_", "lsm2daveragegradphi_",
    "lsm2dgradientmagnitude_", "lsm2dzerooutlevelseteqnrhs_", "lsm2daddadvectiontermtolserhs_",
    "lsm2daddnormalveltermtolserhs_", "lsm2daddconstnormalveltermtolserhs_", "lsm2daddconstcurvtermtolserhs_",
    "lsm2daddexternalandnormalveltermtolserhs_",
    "lsm2drk1step_", "lsm2dtvdrk2stage1_", "lsm2dtvdrk2stage2_", "lsm2dtvdrk3stage1_", "lsm2dtvdrk3stage2_",
    "lsm2dtvdrk3stage3_",
    "lsm2dcomputereinitializationeqnrhs_", "lsm2dcomputeorthogonalizationeqnrhs_",
    "lsm2dcomputemeancurvatureorder2_", "lsm2dcomputemeancurvatureorder4_", "lsm2dcomputeunitnormal_",
    "lsm2dcomputesignedunitnormal_", "lsm2dcomputemeancurvature_", "lsm2dcomputefieldextensioneqnrhs_",
    "lsm2dmaxnormdiff_", "lsm2dcomputestableadvectiondt_", "lsm2dcomputestablenormalveldt_",
    "lsm2dcomputestableconstnormalveldt_", "lsm2dvolumeintegralphilessthanzero_",
    "lsm2dvolumeintegralphigreaterthanzero_", "lsm2dsurfaceintegral_", "lsm2darearegionphilessthanzero_",
    "lsm2darearegionphigreaterthanzero_", "lsm2dperimeterzerolevelset_",
]
QUIRKY_BOUNDS_2D = ["lsm2daddconstprecomputedcurvtermtolserhs_"]
BCS_2D = ["lsm2dlinearextrapolation_", "lsm2dsignedlinearextrapolation_", "lsm2dcopyextrapolation_",
          "lsm2dhomogeneousneumanneno1_", "lsm2dhomogeneousneumanneno2_", "lsm2dhomogeneousneumanneno3_",
          "lsm2dhomogeneousneumannweno5_"]
ORDER4_2D = {"lsm2dcomputemeancurvatureorder4_"}


def prepare2d(case, sym):
    """Inputs that take every branch: gradient components exactly zero in some
    cells and of size 1e-6 or 3e-5 in others, phi exactly +-1e-11 (the zero
    tolerance) in some, gradient magnitudes exactly zero, psi
    with zeros and tiny values, a negative constant normal speed for the CFL
    bound, and a gradient box covering the fill box +-2 in j for the order-4
    curvature (phi_x is read two rows away)."""
    if sym in ORDER4_2D:
        case.boxes["g"] = case.gb
    boxes = H.box_of_array(sym)
    names = [n for k, n in capi.SPECS[sym] if k == "R"]
    for group in (("px", "py"), ("qxp", "qyp"), ("qxm", "qym"), ("pxp", "pyp"), ("pxm", "pym")):
        if group[0] in names and sym not in ("lsm2daveragegradphi_",):
            for n in group:
                f = case.array(n, boxes[n]).reshape(-1)
                f[::13] = 0.0
                f[5::29] *= 1e-6
                f[9::31] *= 3e-5  # squared between the 1e-11 and 1e-8 cut-offs
    for n in ("phi", "phi0"):
        if n in names:  # exactly +-zero_tol: the strict / non-strict comparisons differ here
            f = case.array(n, boxes[n]).reshape(-1)
            f[4::37] = 1e-11
            f[8::41] = -1e-11
    if "psi" in names:
        f = case.array("psi", boxes["psi"]).reshape(-1)
        f[3::11] = 0.0
        f[7::23] *= 1e-12
    if "gmag" in names:
        case.array("gmag", boxes["gmag"]).reshape(-1)[::17] = 0.0
    if sym.startswith("lsm2dcomputestableconstnormalveldt"):
        case.scalars["vn"] = -0.75


def test_reference_exports_the_2d_table():
    ref = H.reference()
    missing = [s for s in capi.SPECS if s.startswith("lsm2d") and s in ref.missing]
    assert missing == []
    assert [s for s in capi.SPECS if s.startswith("lsm2d") and s in H.oracle().missing] == []


@pytest.mark.parametrize("sym", WHOLE_GRID_2D)
@pytest.mark.parametrize("seed", [1, 2])
def test_whole_grid_2d_bitwise(sym, seed):
    c = H.grid_case2d(seed)
    prepare2d(c, sym)
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", ["lsm2daddconstnormalveltermtolserhs_", "lsm2dcomputereinitializationeqnrhs_"])
@pytest.mark.parametrize("variant", [0, 1])
def test_sign_variants_2d(sym, variant):
    c = H.grid_case2d(7)
    c.scalars["vn"] = -0.6 if variant else 1e-13
    c.scalars["use_phi0"] = variant
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", QUIRKY_BOUNDS_2D)
def test_quirky_loop_bounds_2d(sym):
    c = H.grid_case2d(3, mixed_boxes=False)
    c.boxes["rhs"] = (-1, 21, -2, 17)
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", BCS_2D)
@pytest.mark.parametrize("idx", range(4))
def test_ghost_fill_faces_2d(sym, idx):
    c = H.grid_case2d(11, mixed_boxes=False)
    c.boxes["gb"] = c.gb
    c.scalars["idx"] = idx
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)
