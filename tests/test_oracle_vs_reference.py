"""Pin the oracle (plain-C restatement) bitwise against the reference's own
FP64 Fortran kernels, compiled by oracle/Makefile into oracle/_ref/.

Runs on the CPU.  Skipped when the reference build is absent.
"""
import numpy as np
import pytest

import harness as H
from lsmlib_amd import capi

pytestmark = pytest.mark.skipif(H.reference() is None, reason="oracle/_ref/liblsm_ref.so not built")

WHOLE_GRID = [
    "lsm3dhjeno1_", "lsm3dhjeno2_", "lsm3dhjeno3_", "lsm3dhjweno5_",
    "lsm3dupwindhjeno1_", "lsm3dupwindhjeno2_", "lsm3dupwindhjeno3_", "lsm3dupwindhjweno5_",
    "lsm3dcentralgradorder2_", "lsm3dcentralgradorder4_", "lsm3dlaplacianorder2_", "lsm3daveragegradphi_",
    "lsm3dgradientmagnitude_", "lsm3dzerooutlevelseteqnrhs_", "lsm3daddadvectiontermtolserhs_",
    "lsm3daddnormalveltermtolserhs_", "lsm3daddconstnormalveltermtolserhs_", "lsm3daddconstcurvtermtolserhs_",
    "lsm3dcomputereinitializationeqnrhs_", "lsm3dcomputemeancurvatureorder2_",
    "lsm3drk1step_", "lsm3dtvdrk2stage1_", "lsm3dtvdrk2stage2_", "lsm3dtvdrk3stage1_", "lsm3dtvdrk3stage2_",
    "lsm3dtvdrk3stage3_", "lsm3dmaxnormdiff_", "lsm3dcomputestableadvectiondt_",
    "lsm3dcomputestablenormalveldt_", "lsm3dcomputestableconstnormalveldt_",
    "lsm3dvolumeintegralphilessthanzero_", "lsm3dvolumeintegralphigreaterthanzero_", "lsm3dsurfaceintegral_",
    "lsm3dvolumeregionphilessthanzero_", "lsm3dvolumeregionphigreaterthanzero_",
    "lsm3dsurfaceareazerolevelset_",
]
# Averages, counts, the delta-based area, the Hessian, the phi-upwind gradient
# and the first-order delta function, and the control-volume forms.
UTILS = [
    "lsm3daveabsdiff_", "lsm3dvoxelcountgreaterthanzero_", "lsm3dvoxelcountlessthanzero_",
    "lsm3dsurfaceareazerolevelsetdelta_", "lsm3dcentralhessian_", "lsm3dphiupwindgradf_",
    "lsm3ddeltafunctionorder1_",
]
CONTROL_VOLUME = [
    "lsm3dmaxnormdiffcontrolvolume_", "lsm3dcomputestableadvectiondtcontrolvolume_",
    "lsm3dcomputestablenormalveldtcontrolvolume_", "lsm3dcomputestableconstnormalveldtcontrolvolume_",
    "lsm3dvolumeintegralphilessthanzerocontrolvolume_", "lsm3dvolumeintegralphigreaterthanzerocontrolvolume_",
    "lsm3dsurfaceintegralcontrolvolume_", "lsm3dvoxelcountgreaterthanzerocontrolvolume_",
    "lsm3dvoxelcountlessthanzerocontrolvolume_", "lsm3dvolumeregionphilessthanzerocontrolvolume_",
    "lsm3dvolumeregionphigreaterthanzerocontrolvolume_", "lsm3dsurfaceareazerolevelsetcontrolvolume_",
    "lsm3dsurfaceareazerolevelsetdeltacontrolvolume_",
]
# Kernels either side of the curvature / reinitialization path: normals,
# curvatures, orthogonalization and field extension.
NEIGHBOURS = [
    "lsm3dcomputeunitnormal_", "lsm3dcomputesignedunitnormal_", "lsm3dcomputemeancurvature_",
    "lsm3dcomputegaussiancurvatureorder2_", "lsm3dcomputemeancurvatureorder4_",
    "lsm3dcomputegaussiancurvatureorder4_", "lsm3dcomputeorthogonalizationeqnrhs_",
    "lsm3dcomputefieldextensioneqnrhs_",
]
ORDER4 = {"lsm3dcomputemeancurvatureorder4_", "lsm3dcomputegaussiancurvatureorder4_"}
QUIRKY_BOUNDS = ["lsm3daddconstprecomputedcurvtermtolserhs_", "lsm3daddexternalandnormalveltermtolserhs_"]
BCS = ["lsm3dlinearextrapolation_", "lsm3dsignedlinearextrapolation_", "lsm3dcopyextrapolation_",
       "lsm3dhomogeneousneumanneno1_", "lsm3dhomogeneousneumanneno2_", "lsm3dhomogeneousneumanneno3_",
       "lsm3dhomogeneousneumannweno5_"]


def test_oracle_exports_everything():
    assert H.oracle().missing == []
    ref_missing = set(H.reference().missing)
    assert ref_missing == capi.AUTHORED, ref_missing


def _prepare(case, sym):
    names = [n for k, n in capi.SPECS[sym] if k in ("R", "M")]
    boxes = H.box_of_array(sym)
    # phi exactly +-zero_tol, where strict and non-strict comparisons differ
    for n in ("phi", "phi0"):
        if n in names:
            f = case.array(n, boxes[n]).reshape(-1)
            f[4::37] = 1e-11
            f[8::41] = -1e-11
    # gradient components whose squares fall between the 1e-11 and 1e-8 cut-offs
    if sym == "lsm3dgradientmagnitude_":
        for n in ("px", "py", "pz"):
            case.array(n, boxes[n]).reshape(-1)[9::31] *= 5e-5
    # exact zeros in a gradient magnitude exercise the curvature's zero branch
    if "gmag" in names:
        g = case.array("gmag", H.box_of_array(sym)["gmag"])
        g.reshape(-1)[::17] = 0.0
    if sym.startswith("lsm3dcomputestableconstnormalveldt"):
        case.scalars["vn"] = -0.75
    # control volumes with exact zeros, where the counts' >= / <= and the
    # others' > / < differ
    if "cv" in names:
        case.array("cv", boxes["cv"]).reshape(-1)[::7] = 0.0
    # delta functions with exact zeros and negatives (only positive values count)
    if "dphi" in names:
        case.array("dphi", boxes["dphi"]).reshape(-1)[::5] = 0.0

def prepare_neighbours(case, sym):
    """Inputs for NEIGHBOURS that take every branch: gradients exactly zero in
    some cells and of size 1e-6 in others (|grad|^2 below the 1e-11 zero
    tolerance while |grad| is above it), psi with zeros and tiny values, and
    a gradient box two cells wider than the fill box for the fourth-order
    curvatures, whose mixed derivatives read the gradient two cells away."""
    if sym in ORDER4:
        case.boxes["g"] = case.gb
    boxes = H.box_of_array(sym)
    names = [n for k, n in capi.SPECS[sym] if k == "R"]
    for group in (("px", "py", "pz"), ("qxp", "qyp", "qzp"), ("qxm", "qym", "qzm")):
        if group[0] in names:
            for n in group:
                f = case.array(n, boxes[n]).reshape(-1)
                f[::13] = 0.0
                f[5::29] *= 1e-6
    if "psi" in names:
        f = case.array("psi", boxes["psi"]).reshape(-1)
        f[3::11] = 0.0
        f[7::23] *= 1e-12
    _prepare(case, sym)


@pytest.mark.parametrize("sym", NEIGHBOURS)
@pytest.mark.parametrize("seed", [1, 2])
def test_neighbour_kernels_bitwise(sym, seed):
    c = H.grid_case(seed)
    prepare_neighbours(c, sym)
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", WHOLE_GRID)
@pytest.mark.parametrize("seed", [1, 2])
def test_whole_grid_bitwise(sym, seed):
    c = H.grid_case(seed)
    _prepare(c, sym)
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", UTILS)
@pytest.mark.parametrize("seed", [1, 2])
def test_utils_bitwise(sym, seed):
    c = H.grid_case(seed)
    _prepare(c, sym)
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", CONTROL_VOLUME)
@pytest.mark.parametrize("sgn", [1, -1, 0])
def test_control_volume_bitwise(sym, sgn):
    c = H.grid_case(5)
    c.scalars["cvs"] = sgn
    _prepare(c, sym)
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


def test_control_volume_with_no_cell():
    """No cell in the control volume: the CFL steps divide by the initial
    maximum -1 (a negative step), the max-norm difference keeps the first
    cell's value, the sums and counts are zero."""
    for sym in CONTROL_VOLUME:
        c = H.grid_case(6)
        _prepare(c, sym)
        c.arrays["cv"] = np.full(H.shape_of(c.box("cv")), -1.0)
        c.scalars["cvs"] = 1
        a = H.run(H.reference(), sym, c)
        H.assert_same(a, H.run(H.oracle(), sym, c), sym)
        if "dt" in sym:
            assert a[1] < 0, sym


@pytest.mark.parametrize("sym", ["lsm3daddconstnormalveltermtolserhs_", "lsm3dcomputereinitializationeqnrhs_"])
@pytest.mark.parametrize("variant", [0, 1])
def test_sign_variants(sym, variant):
    c = H.grid_case(7)
    c.scalars["vn"] = -0.6 if variant else 1e-13
    c.scalars["use_phi0"] = variant
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", QUIRKY_BOUNDS)
def test_quirky_loop_bounds(sym):
    c = H.grid_case(3, mixed_boxes=False)
    c.boxes["rhs"] = (-1, 19, -2, 14, -3, 10)
    c.boxes["fb"] = (0, 17, 0, 13, 0, 9)
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", BCS)
@pytest.mark.parametrize("idx", range(6))
def test_ghost_fill_faces(sym, idx):
    c = H.grid_case(11, mixed_boxes=False)
    c.boxes["gb"] = c.gb
    c.scalars["idx"] = idx
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("x", [-0.5, -0.3, -0.29, -0.1, 0.0, 1e-9, 0.05, 0.2999, 0.3, 0.31])
def test_heaviside_delta(x):
    c = H.Case((0, 0, 0, 0, 0, 0), (0, 0, 0, 0, 0, 0))
    c.scalars.update(x=x, eps=0.3)
    for sym in ("lsmheaviside_", "lsmdeltafunction_"):
        H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)
