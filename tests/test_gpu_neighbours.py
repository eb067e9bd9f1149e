"""Normals, curvatures, orthogonalization and field extension (the kernels
either side of the curvature / reinitialization path) through the C-ABI with
device arrays: FP64 against the oracle bit for bit, FP32 within tolerance,
box validation, and a sphere at 256^3 against its closed forms."""
import numpy as np
import pytest

import harness as H
from lsmlib_amd import capi
from test_gpu_reference_cases import F32_RTOL, _cmp32, _smooth_case
from test_oracle_vs_reference import NEIGHBOURS, ORDER4, _prepare, prepare_neighbours

pytestmark = pytest.mark.gpu

POW = "lsm3dcomputemeancurvature_"  # divides by (|grad phi|^2)^1.5


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
    """Distance in units in the last place between two float64 arrays."""
    ia = a.view(np.int64).astype(np.float64)
    ib = b.view(np.int64).astype(np.float64)
    ia = np.where(ia < 0, -(ia + 2.0 ** 63), ia)
    ib = np.where(ib < 0, -(ib + 2.0 ** 63), ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("sym", NEIGHBOURS)
@pytest.mark.parametrize("seed", [1, 2])
def test_neighbour_kernels_bitwise(hip, sym, seed):
    c = H.grid_case(seed)
    prepare_neighbours(c, sym)
    want = H.run(H.oracle(), sym, c)
    got = H.run(hip.api, sym, c, device="cuda")
    hip.check()
    if sym == POW:
        # correctly rounded x^1.5 against the C library's pow: equal except
        # within 1 ulp where pow itself is not correctly rounded
        assert _ulps(want[0]["kappa"], got[0]["kappa"]).max() <= 1
    else:
        H.assert_same(want, got, sym)


@pytest.mark.parametrize("boxes", [((-3, 143, -3, 23, -3, 40), (0, 140, 0, 20, 0, 37)),
                                   ((-3, 131, -3, 11, -3, 19), (0, 128, 0, 8, 0, 16))])
@pytest.mark.parametrize("sym", sorted(ORDER4) + ["lsm3dcomputegaussiancurvatureorder2_"])
def test_plane_march_over_many_tiles_and_chunks(hip, sym, boxes):
    """The curvatures on the plane march (64 x 8 tiles, chunks of 16 planes).
    On a fill box of 141 x 21 x 38 cells: three tiles along i (the last one
    13 wide), three along j (the last 5), three chunks along k (the last 6).
    On 129 x 9 x 17 cells every last tile and chunk is a single cell.  phi and
    the gradient on ghost boxes three cells wider."""
    gb, fb = boxes
    c = H.grid_case(8, gb=gb, fb=fb)
    prepare_neighbours(c, sym)
    want = H.run(H.oracle(), sym, c)
    got = H.run(hip.api, sym, c, device="cuda")
    hip.check()
    H.assert_same(want, got, sym)


def test_three_halves_power_over_a_wide_range(hip):
    """The mean curvature's (|grad phi|^2)^1.5 over 4M cells whose |grad phi|^2
    spans 1e-11 .. 1e220 (past the overflow of the power at 1.8e205), against the oracle's
    C-library pow: at most 1 ulp apart, and the count of differing cells is
    printed."""
    n = 128
    c = H.Case((0, n - 1, 0, n - 1, 0, 2 * n - 1), (0, n - 1, 0, n - 1, 0, 2 * n - 1), seed=3)
    rng = c.rng
    shp = H.shape_of(c.gb)
    mag = 10.0 ** rng.uniform(-5.4, 110.0, size=shp)
    u = rng.standard_normal((3,) + shp)
    u /= np.linalg.norm(u, axis=0)
    for n_, a in zip(("px", "py", "pz"), u):
        c.arrays[n_] = a * mag
    for n_ in ("pxx", "pxy", "pxz", "pyy", "pyz", "pzz"):
        c.arrays[n_] = rng.standard_normal(shp)
    for b in ("g", "g2", "kappa"):
        c.boxes[b] = c.gb
    want = H.run(H.oracle(), POW, c)[0]["kappa"]
    got = H.run(hip.api, POW, c, device="cuda")[0]["kappa"]
    hip.check()
    fin = np.isfinite(want)
    assert np.array_equal(fin, np.isfinite(got))
    d = _ulps(want[fin], got[fin])
    print("cells %d, differing %d, max ulps %d" % (want.size, int((d > 0).sum()), int(d.max())))
    assert d.max() <= 1


@pytest.mark.parametrize("sym", NEIGHBOURS)
def test_f32_within_tolerance(hip32, sym):
    c = H.grid_case(1, mixed_boxes=False)
    if sym in ORDER4:
        c.boxes["g"] = c.gb
    _smooth_case(sym, c)
    _prepare(c, sym)
    ref = H.run(H.oracle(), sym, c)
    got = H.run(hip32.api, sym, c, device="cuda", real32=True)
    hip32.check()
    worst = _cmp32(ref, got, sym)
    print("%s: worst relative error %.2e (bound %.0e)" % (sym, worst, F32_RTOL))


@pytest.mark.parametrize("sym,box,bad", [
    ("lsm3dcomputemeancurvatureorder4_", "g", (-3, 20, -1, 16, -3, 12)),   # phi_x(j-2) outside
    ("lsm3dcomputegaussiancurvatureorder4_", "phi", (-1, 20, -3, 16, -3, 12)),  # phi(i-2) outside
    ("lsm3dcomputegaussiancurvatureorder2_", "phi", (-3, 20, -3, 16, -3, 9)),  # phi(k+1) outside
    ("lsm3dcomputefieldextensioneqnrhs_", "phi", (0, 20, -3, 16, -3, 12)),  # phi(i-1) outside
])
def test_box_validation(hip, sym, box, bad):
    """A ghost box too small for the stencil is refused and nothing is written."""
    c = H.grid_case(5, mixed_boxes=False)
    c.boxes["fb"] = (0, 17, 0, 13, 0, 9)
    c.boxes[box] = bad
    c.arrays.pop(box, None)  # regenerated with the shape of its (too small) box
    out = [n for k, n in capi.SPECS[sym] if k == "R"][0]
    res, _ = H.run(hip.api, sym, c, device="cuda")
    code, msg = hip.last_error()
    hip.clear_error()
    assert code == 1
    assert np.array_equal(res[out], c.arrays[out])


def test_sphere_closed_forms_256(hip):
    """phi = |x - c| - R on 256^3 (h = 1/256, R = 0.3) with exact first
    derivatives.  Within two cells of the sphere, against the level set through
    each cell (radius r): the unit normal is the radial direction (1e-12), the
    fourth-order mean curvature 2/r and Gaussian curvature 1/r^2 (relative 1e-7
    and 2e-7; the oracle's errors are 2.1e-8 and 4.2e-8 here and fall 16x per
    halving of h), the second-order Gaussian curvature 1/r^2 (3e-4; oracle
    1.2e-4, 4x per halving)."""
    import torch
    n, R, h = 256, 0.3, 1.0 / 256
    g = 3
    gb = (-g, n - 1 + g) * 3
    fb = (0, n - 1) * 3
    ax = (torch.arange(-g, n + g, dtype=torch.float64, device="cuda") + 0.5) * h - 0.5
    z, y, x = torch.meshgrid(ax, ax, ax, indexing="ij")
    r = torch.sqrt(x * x + y * y + z * z)
    phi = r - R
    px, py, pz = x / r, y / r, z / r
    gmag = torch.ones_like(phi)
    api = hip.api
    out = [torch.empty_like(phi) for _ in range(3)]
    api.lsm3dcomputeunitnormal(*out, gb, px, py, pz, gb, fb)
    near = phi.abs() < 2 * h
    rn = r[near]
    for o, e in zip(out, (px, py, pz)):
        assert float((o - e)[near].abs().max()) < 1e-12
    kappa = torch.empty_like(phi)
    api.lsm3dcomputemeancurvatureorder4(kappa, gb, phi, gb, px, py, pz, gmag, gb, fb, h, h, h)
    e_m4 = float((kappa[near] * rn / 2 - 1).abs().max())
    api.lsm3dcomputegaussiancurvatureorder4(kappa, gb, phi, gb, px, py, pz, gmag, gb, fb, h, h, h)
    e_g4 = float((kappa[near] * rn * rn - 1).abs().max())
    api.lsm3dcomputegaussiancurvatureorder2(kappa, gb, phi, gb, px, py, pz, gmag, gb, fb, h, h, h)
    e_g2 = float((kappa[near] * rn * rn - 1).abs().max())
    hip.check()
    print("relative errors: mean order 4 %.2e, Gaussian order 4 %.2e, Gaussian order 2 %.2e" % (e_m4, e_g4, e_g2))
    assert e_m4 < 1e-7 and e_g4 < 2e-7 and e_g2 < 3e-4
