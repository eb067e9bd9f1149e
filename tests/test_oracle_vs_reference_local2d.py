"""Pin the 2D narrow-band restatement (oracle/oracle_local2d.c) bitwise against
the reference's own 2D local Fortran and 2D band construction, compiled by
oracle/Makefile into oracle/_ref/ (CPU; skipped when the build is absent).

lsm2dHJENO3LOCAL cannot be called through its C prototype: the Fortran
(src/toolbox/lsm_spatial_derivatives2d_local.f.in:498-522) leaves index_x and
index_y out of its dummy argument list while the header
(src/toolbox/lsm_spatial_derivatives2d_local.h:220-273) passes them, so every
later argument lands one or two places off.  The oracle follows the header and
is pinned instead against the whole-grid lsm2dhjeno3_ (itself bitwise to the
reference) on band points whose stencil lies in the inner levels, where the
two must agree exactly; the other narrow-band derivatives get the same check."""
import numpy as np
import pytest

import harness as H
from lsmlib_amd import capi

pytestmark = pytest.mark.skipif(H.reference() is None, reason="oracle/_ref/liblsm_ref.so not built")

LOCAL_2D = [
    "lsm2dhjeno1local_", "lsm2dhjeno2local_", "lsm2dhjweno5local_", "lsm2dupwindhjeno2local_",
    "lsm2dcentralgradorder2local_", "lsm2dcentralgradorder4local_", "lsm2dlaplacianorder2local_",
    "lsm2dcomputeavegradphilocal_", "lsm2dgradientmagnitudelocal_", "lsm2dzerooutlevelseteqnrhslocal_",
    "lsm2daddadvectiontermtolserhslocal_", "lsm2daddnormalveltermtolserhslocal_",
    "lsm2daddconstnormalveltermtolserhslocal_", "lsm2daddconstcurvtermtolserhslocal_",
    "lsm2daddconstprecomputedcurvtermtolserhslocal_", "lsm2daddexternalandnormalveltermtolserhslocal_",
    "lsm2drk1steplocal_", "lsm2dtvdrk2stage1local_", "lsm2dtvdrk2stage2local_", "lsm2dtvdrk3stage1local_",
    "lsm2dtvdrk3stage2local_", "lsm2dtvdrk3stage3local_", "lsm2dmaxnormdifflocal_",
    "lsm2dcomputestableadvectiondtlocal_", "lsm2dcomputestablenormalveldtlocal_",
    "lsm2dcomputestableconstnormalveldtlocal_", "lsm2dcomputereinitializationeqnrhslocal_",
    "lsm2dcomputeorthogonalizationeqnrhslocal_", "lsm2dcomputemeancurvatureorder2local_",
    "lsm2dcomputemeancurvatureorder4local_", "lsm2dcomputeunitnormallocal_", "lsm2dcomputesignedunitnormallocal_",
    "lsm2dmultiplycutofflserhslocal_", "lsm2dimposemasklocal_", "lsm2dcopydatalocal_",
]
BAND_OUT = ("nb", "ix", "iy", "outer", "n_lo", "n_hi")
OUTER = ("nlo_outer_plus", "nhi_outer_plus", "nlo_outer_minus", "nhi_outer_minus")


def prepare_local2d(c, sym):
    """Inputs that take every branch on the band: gradient components exactly
    zero or of size 1e-6 / 3e-5 at some points, phi and phi0 exactly +-1e-11
    (the zero tolerance, where >= and > differ) at some band points, zero
    gradient magnitudes, psi with zeros and tiny values, velocities with exact
    zeros and 1e-12 values, and a mask for the mask helpers."""
    boxes = H.box_of_array(sym)
    names = [n for k, n in capi.SPECS[sym] if k == "R"]
    c.arrays.setdefault("mask", c.rng.standard_normal(c.arrays["phi"].shape))
    c.boxes.setdefault("mask", c.gb)
    for group in (("px", "py"), ("qxp", "qyp"), ("qxm", "qym"), ("pxp", "pyp"), ("pxm", "pym")):
        if group[0] in names:
            for n in group:
                f = c.array(n, boxes[n]).reshape(-1)
                f[::13] = 0.0
                f[5::29] *= 1e-6
                f[9::31] *= 3e-5
    if "px" in names:  # |grad phi| exactly zero_tol: >= and > differ
        c.array("px", boxes["px"]).reshape(-1)[11::43] = 1e-11
        c.array("py", boxes["py"]).reshape(-1)[11::43] = 0.0
    nlo, nhi = c.scalars["nlo0"], c.scalars["nhi0"]
    pts = list(zip(c.arrays["ix"][: nhi - nlo + 1], c.arrays["iy"][: nhi - nlo + 1]))
    for n in ("phi0",):
        if n in names:
            c.array(n, boxes[n])
    for n in ("phi", "phi0"):
        if n in names and n in c.arrays and c.box(boxes[n]) == c.gb:
            a = c.arrays[n]
            for q, (i, j) in enumerate(pts[4::37]):
                a[j + 3, i + 3] = 1e-11 if q % 2 == 0 else -1e-11
            for q, (i, j) in enumerate(pts[6::41]):  # |phi| exactly beta or gamma (the cut-off's <=)
                a[j + 3, i + 3] = (1, -1)[q % 2] * c.scalars[("beta", "gamma")[(q // 2) % 2]]
    if "psi" in names:
        f = c.array("psi", boxes["psi"]).reshape(-1)
        f[3::11] = 0.0
        f[7::23] *= 1e-12
    if "gmag" in names:
        c.array("gmag", boxes["gmag"]).reshape(-1)[::17] = 0.0
    for v in ("vx", "vy", "vn"):
        if v in names:
            f = c.array(v, boxes[v]).reshape(-1)
            f[::11] = 0.0
            f[6::19] *= 1e-12


def flat_patch(c, start=100, n=60):
    """phi nearly constant over n consecutive level-0 list entries (steps of
    1e-19): there every WENO5 difference is tiny and its additive 1e-35 term
    decides the weights."""
    phi = c.arrays["phi"]
    for q, l in enumerate(range(c.scalars["nlo0"] + start, c.scalars["nlo0"] + start + n)):
        i, j = int(c.arrays["ix"][l]), int(c.arrays["iy"][l])
        phi[j + 3, i + 3] = 0.01 + 1e-19 * q


def test_reference_exports_the_2d_local_table():
    syms = [s for s in capi.SPECS if s.startswith("lsm2d") and ("local" in s or "narrowband" in s or "outer" in s)]
    assert len(syms) == 40
    assert [s for s in syms if s in H.reference().missing] == []  # exported, ENO3 included
    assert [s for s in syms if s in H.oracle().missing] == []


@pytest.mark.parametrize("level", [0, 1, 3, 4])
def test_determine_narrow_band_2d(level):
    c = H.band_case2d(level=level)
    r = H.band_case2d(H.reference(), level=level)
    for k in BAND_OUT:
        assert np.array_equal(c.arrays[k], r.arrays[k]), k
    for k in OUTER:
        assert c.scalars[k] == r.scalars[k], k


def test_determine_narrow_band_2d_single_point_does_not_grow():
    """Level 0 with one point: the 2D routine grows no level (:231) and leaves
    n_lo/n_hi(1..level) as the caller set them."""
    for impl in (H.oracle(), H.reference()):
        c = H.Case((-3, 12, -3, 10), (0, 9, 0, 7))
        phi = np.full(H.shape_of(c.gb), 5.0)
        phi[6, 7] = 0.01
        c.arrays.update(phi=phi, nb=np.zeros(phi.shape, np.uint8), ix=np.full(phi.size, -7, np.int32),
                        iy=np.full(phi.size, -7, np.int32), outer=np.full(phi.size, -7, np.int32),
                        n_lo=np.full(4, -5, np.int32), n_hi=np.full(4, -5, np.int32))
        c.boxes.update(nb=c.gb)
        c.scalars.update(nlo=0, nhi=phi.size - 1, nlo_outer=0, nhi_outer=phi.size - 1, level=3, width=0.3,
                         width_inner=0.005)
        arrs, outs = H.run(impl, "lsm2ddeterminenarrowband_", c)
        assert arrs["n_lo"].tolist() == [0, -5, -5, -5] and arrs["n_hi"].tolist() == [0, -5, -5, -5]
        assert (arrs["ix"][0], arrs["iy"][0]) == (4, 3)
        assert outs == (phi.size - 1, phi.size - 1, 0, -1)


@pytest.mark.parametrize("use_sign", [0, 1])
def test_determine_narrow_band_from_mask_2d(use_sign):
    c = H.band_case2d(level=3)
    mask = c.arrays["phi"] - 0.1
    mask.reshape(-1)[::7] = 0.0  # mask == 0 is in the band for use_mask_sign > 0, out of it otherwise
    c.arrays["mask"] = mask
    c.boxes["mask"] = c.gb
    c.scalars["use_mask_sign"] = use_sign
    c.scalars.update(nlo=0, nhi=c.arrays["ix"].size - 1)
    H.assert_same(H.run(H.reference(), "lsm2ddeterminenarrowbandfrommask_", c),
                  H.run(H.oracle(), "lsm2ddeterminenarrowbandfrommask_", c), "frommask2d")


@pytest.mark.parametrize("sym", LOCAL_2D)
@pytest.mark.parametrize("seed", [0, 1])
def test_local_2d_bitwise(sym, seed):
    c = H.band_case2d(seed=seed, level=4)
    c.scalars["dy"] = 0.8 * c.scalars["dx"]
    flat_patch(c)
    if sym == "lsm2dcomputemeancurvatureorder4local_":
        c.boxes["g"] = c.gb
    prepare_local2d(c, sym)
    if sym.startswith("lsm2dcomputestableconstnormalveldt"):
        c.scalars["vn"] = -0.75
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


@pytest.mark.parametrize("sym", ["lsm2daddconstnormalveltermtolserhslocal_", "lsm2dcomputereinitializationeqnrhslocal_"])
@pytest.mark.parametrize("variant", [0, 1])
def test_sign_variants_local_2d(sym, variant):
    c = H.band_case2d(seed=3, level=4)
    prepare_local2d(c, sym)
    c.scalars["vn"] = -0.6 if variant else 1e-13
    c.scalars["use_phi0"] = variant
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


def test_check_outer_layer_2d():
    c = H.band_case2d(level=3)
    sym = "lsm2dcheckouternarrowbandlayer_"
    a = H.run(H.reference(), sym, c)
    H.assert_same(a, H.run(H.oracle(), sym, c), sym)
    assert a[1] == 0
    for lo_key in ("nlo_outer_plus", "nlo_outer_minus"):
        d = H.band_case2d(level=3)
        l = d.arrays["outer"][d.scalars[lo_key]]
        i, j = d.arrays["ix"][l], d.arrays["iy"][l]
        d.arrays["phi"][j + 3, i + 3] = -d.arrays["phi"][j + 3, i + 3] if lo_key.endswith("plus") else 0.5
        a = H.run(H.reference(), sym, d)
        H.assert_same(a, H.run(H.oracle(), sym, d), sym)
        assert a[1] == 1


WHOLE_TWIN = {"lsm2dhjeno1local_": "lsm2dhjeno1_", "lsm2dhjeno2local_": "lsm2dhjeno2_",
              "lsm2dhjeno3local_": "lsm2dhjeno3_", "lsm2dhjweno5local_": "lsm2dhjweno5_",
              "lsm2dupwindhjeno2local_": "lsm2dupwindhjeno2_"}


def inner_points(c, reach=3):
    """Level-0 points at least `reach` cells inside the fill box whose axis
    neighbours up to `reach` away are all band points of levels 0 and 1
    (marks 1 and 2): there every difference the narrow-band stencil reads has
    been computed from phi."""
    nb = c.arrays["nb"].astype(np.int16)
    ilo, ihi, jlo, jhi = c.fb
    pts = []
    for l in range(c.scalars["nlo0"], c.scalars["nhi0"] + 1):
        i, j = int(c.arrays["ix"][l]), int(c.arrays["iy"][l])
        if not (ilo + reach <= i <= ihi - reach and jlo + reach <= j <= jhi - reach):
            continue
        ok = all(1 <= nb[j + 3, i + s + 3] <= 2 and 1 <= nb[j + s + 3, i + 3] <= 2 for s in range(-reach, reach + 1))
        if ok:
            pts.append((i, j))
    return pts


@pytest.mark.parametrize("sym", sorted(WHOLE_TWIN))
@pytest.mark.parametrize("field", ["band", "quadratic", "cubic", "zigzag"])
def test_local_derivatives_equal_whole_grid_on_inner_band(sym, field):
    """field: the band's own phi, or after building the band an integer
    polynomial of degree 2 or 3 in (i, j), whose second or third differences
    are constant along each axis, or a +-1 checkerboard whose differences
    alternate in sign with equal magnitude, so that every ENO comparison is a
    tie."""
    c = H.band_case2d(seed=1, level=4, width=12.0 / 40, width_inner=6.0 / 40)
    c.scalars.update(mark_D1=126, mark_D2=125, mark_D3=124)
    if field == "zigzag":  # |D_n| equal with alternating signs: every ENO tie picks a different value
        ilo, ihi, jlo, jhi = c.gb
        j, i = np.meshgrid(np.arange(jlo, jhi + 1), np.arange(ilo, ihi + 1), indexing="ij")
        c.arrays["phi"] = np.where((i + j) % 2 == 0, 1.0, -1.0) + 0.25 * np.where(j % 3 == 0, 1.0, 0.0)
    elif field != "band":
        ilo, ihi, jlo, jhi = c.gb
        j, i = np.meshgrid(np.arange(jlo, jhi + 1.0), np.arange(ilo, ihi + 1.0), indexing="ij")
        c.arrays["phi"] = (i * i - 3 * i * j + 2 * j * j) if field == "quadratic" else (i ** 3 - 2 * j ** 3 + i * j)
    pts = inner_points(c)
    assert len(pts) > 200
    loc, _ = H.run(H.oracle(), sym, c)
    whole, _ = H.run(H.oracle(), WHOLE_TWIN[sym], c)
    outs = [n for k, n in capi.SPECS[sym] if k == "R" and n[0] == "p" and n != "phi"]
    ii = np.array([p[0] for p in pts]) + 3
    jj = np.array([p[1] for p in pts]) + 3
    for n in outs:
        a, b = loc[n][jj, ii], whole[n][jj, ii]
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), (sym, n, np.abs(a - b).max())
    if sym != "lsm2dhjeno3local_":
        ref, _ = H.run(H.reference(), sym, c)
        for n in outs:
            assert np.array_equal(ref[n], loc[n]), (sym, n)


def test_empty_lists_2d():
    """nhi < nlo: the maximum-norm difference is 0 and the CFL steps divide by
    -1 + tiny, as the reference's loops that never run leave them."""
    for sym in ("lsm2dmaxnormdifflocal_", "lsm2dcomputestableadvectiondtlocal_",
                "lsm2dcomputestablenormalveldtlocal_", "lsm2dcomputestableconstnormalveldtlocal_",
                "lsm2dcomputeavegradphilocal_"):
        c = H.band_case2d(level=1)
        c.scalars.update(nlo=5, nhi=4)
        a, b = H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c)
        H.assert_same(a, b, sym)
        assert a[1] == (0.0 if sym in ("lsm2dmaxnormdifflocal_", "lsm2dcomputeavegradphilocal_")
                        else c.scalars["cfl"] / (-1.0 + 1e-36)), (sym, a[1])
