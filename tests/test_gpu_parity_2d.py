"""The 2D twins on the device through the C-ABI: FP64 bitwise against the
oracle (integrals to 1e-13 absolute, the mean curvature from a Hessian to 1 ulp), FP32
within tolerance on smooth inputs, box validation, and a circle at 4096^2
against closed forms."""
import numpy as np
import pytest

import harness as H
from lsmlib_amd import capi
from test_gpu_reference_cases import _cmp32
from test_oracle_vs_reference_2d import BCS_2D, ORDER4_2D, QUIRKY_BOUNDS_2D, WHOLE_GRID_2D, prepare2d

pytestmark = pytest.mark.gpu

SUMS_2D = {"lsm2dvolumeintegralphilessthanzero_", "lsm2dvolumeintegralphigreaterthanzero_", "lsm2dsurfaceintegral_",
           "lsm2darearegionphilessthanzero_", "lsm2darearegionphigreaterthanzero_", "lsm2dperimeterzerolevelset_"}
POW_2D = "lsm2dcomputemeancurvature_"


@pytest.fixture(scope="module")
def hip():
    import lsmlib_amd
    lib = lsmlib_amd.load("f64")
    lib.set_stream(0, True)
    return lib


@pytest.fixture(scope="module")
def hip32():
    import lsmlib_amd
    lib = lsmlib_amd.load("f32")
    lib.set_stream(0, True)
    return lib


def _ulps(a, b):
    ia = a.view(np.int64).astype(np.float64)
    ib = b.view(np.int64).astype(np.float64)
    ia = np.where(ia < 0, -(ia + 2.0 ** 63), ia)
    ib = np.where(ib < 0, -(ib + 2.0 ** 63), ib)
    return np.abs(ia - ib)


def _cmp(hip, sym, c):
    a = H.run(H.oracle(), sym, c)
    b = H.run(hip.api, sym, c, device="cuda")
    hip.check()
    if sym in SUMS_2D:
        H.assert_same(a, b, sym, out_abs=1e-13)
    elif sym == POW_2D:
        assert _ulps(a[0]["kappa"], b[0]["kappa"]).max() <= 1
    else:
        H.assert_same(a, b, sym, skip=("D1", "D2", "D3"))


REDUCTIONS_2D = sorted(SUMS_2D) + ["lsm2dmaxnormdiff_", "lsm2dcomputestableadvectiondt_",
                                   "lsm2dcomputestablenormalveldt_", "lsm2dcomputestableconstnormalveldt_"]


@pytest.mark.parametrize("sym", REDUCTIONS_2D)
def test_reductions_2d_on_long_rows(hip, sym):
    """Rows of 300 cells (one full 4 x 64 pass of a wave and a partial one),
    5 rows, against the oracle."""
    c = H.grid_case2d(7, gb=(-3, 302, -3, 7), fb=(0, 299, 0, 4))
    prepare2d(c, sym)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", WHOLE_GRID_2D)
@pytest.mark.parametrize("seed", [1, 2])
def test_whole_grid_2d_bitwise(hip, sym, seed):
    c = H.grid_case2d(seed)
    prepare2d(c, sym)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", ["lsm2daddconstnormalveltermtolserhs_", "lsm2dcomputereinitializationeqnrhs_"])
@pytest.mark.parametrize("variant", [0, 1])
def test_sign_variants_2d(hip, sym, variant):
    c = H.grid_case2d(7)
    c.scalars["vn"] = -0.6 if variant else 1e-13
    c.scalars["use_phi0"] = variant
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", QUIRKY_BOUNDS_2D)
def test_quirky_loop_bounds_2d(hip, sym):
    c = H.grid_case2d(3, mixed_boxes=False)
    c.boxes["rhs"] = (-1, 21, -2, 17)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", BCS_2D)
@pytest.mark.parametrize("idx", range(4))
def test_ghost_fill_faces_2d(hip, sym, idx):
    c = H.grid_case2d(11, mixed_boxes=False)
    c.boxes["gb"] = c.gb
    c.scalars["idx"] = idx
    _cmp(hip, sym, c)


def _smooth2d(name, box, h=0.1):
    """Smooth 2D fields per argument name, away from the FP32 zero tolerance."""
    ilo, ihi, jlo, jhi = box
    j, i = np.meshgrid(np.arange(jlo, jhi + 1), np.arange(ilo, ihi + 1), indexing="ij")
    x, y = i * h, j * h
    seed = sum(map(ord, name))
    a, b, ph = 0.3 + (seed % 7) * 0.1, 0.2 + (seed % 5) * 0.13, seed * 0.37
    wave = np.sin(a * x + b * y + ph)
    if name in ("phi", "phi0", "uc", "un", "u1", "u2", "F", "f1", "f2"):
        return np.sqrt((x - 0.9) ** 2 + (y - 0.7) ** 2) - 0.5 + 0.05 * wave
    if name == "gmag":
        return 1.0 + 0.3 * wave
    if name.startswith("p") or name in ("vx", "vy", "vn", "kappa", "rhs", "D1", "D2", "D3"):
        return 0.5 + wave
    return wave


@pytest.mark.parametrize("sym", WHOLE_GRID_2D + BCS_2D)
def test_f32_2d_within_tolerance(hip32, sym):
    c = H.grid_case2d(1, mixed_boxes=False)
    if sym in ORDER4_2D:
        c.boxes["g"] = c.gb
    c.boxes["gb"] = c.gb
    boxes = H.box_of_array(sym)
    for kind, name in capi.SPECS[sym]:
        if kind == "R":
            c.arrays[name] = _smooth2d(name, c.box(boxes[name]), c.scalars["dx"])
    ref = H.run(H.oracle(), sym, c)
    got = H.run(hip32.api, sym, c, device="cuda", real32=True)
    hip32.check()
    _cmp32(ref, got, sym)


@pytest.mark.parametrize("boxes", [((-3, 143, -3, 23), (0, 140, 0, 20)), ((-3, 131, -3, 11), (0, 128, 0, 8))])
def test_f32_2d_weno5_over_many_tiles(hip32, boxes):
    """FP32 2D HJ-WENO5 within the FP32 tolerance of the FP64 oracle on fill
    boxes of three 64 x 4 tiles along i and six along j, and of a single-cell
    last tile on both axes, the outputs on ghost boxes of their own."""
    sym = "lsm2dhjweno5_"
    gb, fb = boxes
    c = H.grid_case2d(2, gb=gb, fb=fb)  # outputs on ghost boxes of their own
    bx = H.box_of_array(sym)
    for kind, name in capi.SPECS[sym]:
        if kind == "R":
            c.arrays[name] = _smooth2d(name, c.box(bx[name]), c.scalars["dx"])
    ref = H.run(H.oracle(), sym, c)
    got = H.run(hip32.api, sym, c, device="cuda", real32=True)
    hip32.check()
    _cmp32(ref, got, sym)


@pytest.mark.parametrize("sym,box,bad", [
    ("lsm2dcomputemeancurvatureorder4_", "g", (-3, 22, -1, 18)),   # phi_x(j-2) outside
    ("lsm2dhjweno5_", "phi", (-2, 22, -3, 18)),                    # phi(i-3) outside
    ("lsm2dcomputemeancurvatureorder2_", "phi", (-3, 22, -3, 15)),  # phi(j+1) outside
])
def test_box_validation_2d(hip, sym, box, bad):
    c = H.grid_case2d(5, mixed_boxes=False)
    c.boxes[box] = bad
    c.arrays.pop(box, None)
    out = [n for k, n in capi.SPECS[sym] if k == "R"][0]
    res, _ = H.run(hip.api, sym, c, device="cuda")
    code, _msg = hip.last_error()
    hip.clear_error()
    assert code == 1
    assert np.array_equal(res[out], c.arrays[out])


def test_circle_closed_forms_4096(hip):
    """phi = |x - c| - R on a 4096^2 grid (h = 1/4096, R = 0.3), 3 ghost cells:
    HJ-WENO5 plus/minus x-derivatives equal x/r to 1e-10 away from the centre
    (the oracle: 4.5e-12); the second-order mean curvature from central
    gradients is 1/r to 1e-6 within two cells of the circle (oracle 1.7e-7);
    with eps = 1.5 h the smoothed area of {phi < 0} is pi R^2 to 1e-6 (oracle
    2.1e-7) and the perimeter 2 pi R to 2e-5 (oracle 6.1e-6)."""
    import torch
    n, R, h, g = 4096, 0.3, 1.0 / 4096, 3
    gb = (-g, n - 1 + g, -g, n - 1 + g)
    fb = (0, n - 1, 0, n - 1)
    ax = (torch.arange(-g, n + g, dtype=torch.float64, device="cuda") + 0.5) * h - 0.5
    y, x = torch.meshgrid(ax, ax, indexing="ij")
    r = torch.sqrt(x * x + y * y)
    phi = r - R
    api = hip.api
    e = lambda: torch.zeros_like(phi)  # noqa: E731
    xp, yp, xm, ym, D1 = e(), e(), e(), e(), e()
    api.lsm2dhjweno5(xp, yp, gb, xm, ym, gb, phi, gb, D1, gb, fb, h, h)
    far = (r > 0.05) & (r < 0.45)
    inner = torch.zeros_like(far)
    inner[g:-g, g:-g] = True
    sel = far & inner
    assert float((xp - x / r)[sel].abs().max()) < 1e-10 and float((xm - x / r)[sel].abs().max()) < 1e-10
    px, py, gm, kap = e(), e(), e(), e()
    api.lsm2dcentralgradorder2(px, py, gb, phi, gb, fb, h, h)
    api.lsm2dgradientmagnitude(px, py, gm, gb, fb)
    api.lsm2dcomputemeancurvatureorder2(kap, gb, phi, gb, px, py, gm, gb, fb, h, h)
    near = (phi.abs() < 2 * h) & inner
    e_k = float((kap[near] * r[near] - 1).abs().max())
    area = api.lsm2darearegionphilessthanzero(None, phi, gb, fb, h, h, 1.5 * h)
    perim = api.lsm2dperimeterzerolevelset(None, phi, gb, px, py, gb, fb, h, h, 1.5 * h)
    hip.check()
    print("curvature rel %.2e, area rel %.2e, perimeter rel %.2e" % (
        e_k, abs(area / (np.pi * R * R) - 1), abs(perim / (2 * np.pi * R) - 1)))
    assert e_k < 1e-6
    assert abs(area / (np.pi * R * R) - 1) < 1e-6
    assert abs(perim / (2 * np.pi * R) - 1) < 2e-5


@pytest.mark.parametrize("scheme", ["eno1", "eno2", "eno3", "weno5"])
def test_neumann_zero_normal_derivative_2d_hip(hip, scheme):
    """tests/boundary_conditions/test_neumann_bc_2d.cc on the device: bitwise
    the oracle's, and the face-normal one-sided derivative below 1e-6 on every
    boundary line."""
    from test_oracle_pinning import neumann_fill_and_derivatives2d
    ref = neumann_fill_and_derivatives2d(H.oracle(), scheme)
    got = neumann_fill_and_derivatives2d(hip.api, scheme, device="cuda")
    hip.check()
    H.assert_same(ref, got, scheme, skip=("D1", "D2", "D3"))
    a = got[0]
    I = (slice(3, -3),) * 2
    assert np.abs(a["pxm"][I][:, 0]).max() < 1e-6 and np.abs(a["pxp"][I][:, -1]).max() < 1e-6
    assert np.abs(a["pym"][I][0, :]).max() < 1e-6 and np.abs(a["pyp"][I][-1, :]).max() < 1e-6
