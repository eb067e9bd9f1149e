"""The reference's own test cases and closed forms run through the HIP
libraries on the device, and the FP32 library (liblsmhip_f32.so) against the
FP64 oracle within stated tolerances."""
import numpy as np
import pytest

import harness as H
from lsmlib_amd import capi
from test_oracle_pinning import _full_band_case, _grid, neumann_fixture
from test_oracle_vs_reference import WHOLE_GRID, _prepare
from test_oracle_vs_reference_local import LOCAL

pytestmark = pytest.mark.gpu

# FP32 against FP64: every real output agrees to F32_RTOL of the output's own
# largest magnitude (plus F32_ATOL), scalar outputs to F32_RTOL relative.
F32_RTOL = 2e-5
F32_ATOL = 1e-6


@pytest.fixture(scope="module")
def hip64():
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


def _neumann_case():
    gb, fb, d, phi = neumann_fixture()
    c = H.Case(gb, fb)
    c.arrays["phi"] = phi
    c.boxes["gb"] = gb
    c.scalars.update(dx=d[0], dy=d[1], dz=d[2])
    return c


@pytest.mark.parametrize("scheme", ["eno1", "eno2", "eno3", "weno5"])
def test_neumann_zero_normal_derivative_hip(hip64, scheme):
    """tests/boundary_conditions/test_neumann_bc_3d.cc on the device: Neumann
    fill of the six faces, then HJ-ENO/WENO; bitwise the oracle's, and the
    face-normal one-sided derivative below 1e-6 on every boundary face."""
    res = {}
    for impl, dev in ((H.oracle(), None), (hip64.api, "cuda")):
        c = _neumann_case()
        for idx in range(6):
            c.scalars["idx"] = idx
            arrs, _ = H.run(impl, "lsm3dhomogeneousneumann%s_" % scheme, c, device=dev)
            c.arrays["phi"] = arrs["phi"]
        res[dev] = H.run(impl, "lsm3dhj%s_" % scheme, c, device=dev)
        hip64.check()
    H.assert_same(res[None], res["cuda"], scheme, skip=("D1", "D2", "D3"))
    a = res["cuda"][0]
    I = (slice(3, -3),) * 3
    assert np.abs(a["pxm"][I][:, :, 0]).max() < 1e-6 and np.abs(a["pxp"][I][:, :, -1]).max() < 1e-6
    assert np.abs(a["pym"][I][:, 0, :]).max() < 1e-6 and np.abs(a["pyp"][I][:, -1, :]).max() < 1e-6
    assert np.abs(a["pzm"][I][0, :, :]).max() < 1e-6 and np.abs(a["pzp"][I][-1, :, :]).max() < 1e-6


@pytest.mark.parametrize("scheme", ["eno1", "eno2", "eno3", "weno5"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_linear_field_exact_gradient_hip(hip64, hip32, scheme, prec):
    gb, fb, i, j, k = _grid()
    c = H.Case(gb, fb)
    c.scalars.update(dx=0.1, dy=0.2, dz=0.05)
    grad = (0.7, -1.3, 2.1)
    c.arrays["phi"] = grad[0] * i * 0.1 + grad[1] * j * 0.2 + grad[2] * k * 0.05 + 0.3
    lib = hip64 if prec == "f64" else hip32
    arrs, _ = H.run(lib.api, "lsm3dhj%s_" % scheme, c, device="cuda", real32=prec == "f32")
    lib.check()
    tol = 1e-12 if prec == "f64" else 2e-4
    I = (slice(3, -3),) * 3
    for n, name in enumerate(("pxp", "pyp", "pzp", "pxm", "pym", "pzm")):
        assert np.abs(arrs[name][I] - grad[n % 3]).max() < tol, (name, prec)


# ---------------------------------------------------------------------------
# FP32 library within tolerance of the FP64 oracle, on smooth inputs
# ---------------------------------------------------------------------------
def _smooth(name, box, h=0.1):
    """A smooth field per argument name, away from the tolerance thresholds
    that the single-precision build moves (zero_tol 1e-5)."""
    ilo, ihi, jlo, jhi, klo, khi = box
    k, j, i = np.meshgrid(np.arange(klo, khi + 1), np.arange(jlo, jhi + 1), np.arange(ilo, ihi + 1), indexing="ij")
    x, y, z = i * h, j * h, k * h
    seed = sum(map(ord, name))
    a, b, c, ph = 0.3 + (seed % 7) * 0.1, 0.2 + (seed % 5) * 0.13, 0.25 + (seed % 3) * 0.17, seed * 0.37
    wave = np.sin(a * x + b * y + c * z + ph)
    if name in ("phi", "phi0", "uc", "un", "u1", "u2", "F", "f1", "f2", "mask"):
        return np.sqrt((x - 0.9) ** 2 + (y - 0.7) ** 2 + (z - 0.6) ** 2) - 0.5 + 0.05 * wave
    if name in ("gmag",):
        return 1.0 + 0.3 * wave
    if name.startswith("p") or name in ("vx", "vy", "vz", "vn", "kappa", "rhs", "D1", "D2", "D3"):
        return 0.5 + wave
    return wave


def _smooth_case(sym, case):
    boxes = H.box_of_array(sym)
    for kind, name in capi.SPECS[sym]:
        if kind == "R":
            case.arrays[name] = _smooth(name, case.box(boxes[name]), case.scalars["dx"])


def _cmp32(ref, got, sym):
    arrs_r, out_r = ref
    arrs_g, out_g = got
    worst = 0.0
    for name, a in arrs_r.items():
        if a.dtype.kind != "f" or name in ("D1", "D2", "D3"):
            continue
        b = arrs_g[name].astype(np.float64)
        scale = float(np.abs(a).max())
        err = float(np.abs(a - b).max())
        assert err <= F32_RTOL * scale + F32_ATOL, (sym, name, err, scale)
        worst = max(worst, err / max(scale, 1e-30))
    if out_r is not None:
        r = np.atleast_1d(np.array(out_r, dtype=np.float64))
        g = np.atleast_1d(np.array(out_g, dtype=np.float64))
        assert np.all(np.abs(r - g) <= F32_RTOL * np.abs(r) + 1e-30), (sym, out_r, out_g)
    return worst


@pytest.mark.parametrize("sym", WHOLE_GRID)
def test_f32_whole_grid_within_tolerance(hip32, sym):
    c = H.grid_case(1, mixed_boxes=False)
    _smooth_case(sym, c)
    _prepare(c, sym)
    ref = H.run(H.oracle(), sym, c)
    got = H.run(hip32.api, sym, c, device="cuda", real32=True)
    hip32.check()
    _cmp32(ref, got, sym)


@pytest.mark.parametrize("boxes", [((-3, 143, -3, 23, -3, 40), (0, 140, 0, 20, 0, 37)),
                                   ((-3, 131, -3, 11, -3, 19), (0, 128, 0, 8, 0, 16))])
def test_f32_weno5_over_many_tiles(hip32, boxes):
    """FP32 HJ-WENO5 forms each window once and passes the plus derivative to
    the neighbouring cell (through LDS within a 64 x 8 tile, from the cells one
    past it at its edges, along k from plane to plane): within the FP32
    tolerance of the FP64 oracle on fill boxes of three tiles per axis and of
    a single-cell last tile and chunk on every axis."""
    sym = "lsm3dhjweno5_"
    gb, fb = boxes
    c = H.grid_case(4, gb=gb, fb=fb)  # outputs on ghost boxes of their own
    _smooth_case(sym, c)
    _prepare(c, sym)
    ref = H.run(H.oracle(), sym, c)
    got = H.run(hip32.api, sym, c, device="cuda", real32=True)
    hip32.check()
    worst = _cmp32(ref, got, sym)
    print("worst relative error %.2e (bound %.0e)" % (worst, F32_RTOL))


@pytest.mark.parametrize("sym", LOCAL)
def test_f32_local_within_tolerance(hip32, sym):
    c = _full_band_case()
    _smooth_case(sym, c)
    ref = H.run(H.oracle(), sym, c)
    got = H.run(hip32.api, sym, c, device="cuda", real32=True)
    hip32.check()
    _cmp32(ref, got, sym)
