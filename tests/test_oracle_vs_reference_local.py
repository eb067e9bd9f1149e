"""Pin the oracle's narrow-band kernels and band construction bitwise against
the reference build (CPU; skipped when oracle/_ref is absent)."""
import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.skipif(H.reference() is None, reason="oracle/_ref/liblsm_ref.so not built")

LOCAL = [
    "lsm3dhjeno1local_", "lsm3dhjeno2local_", "lsm3dcentralgradorder2local_", "lsm3dcentralgradorder4local_",
    "lsm3dlaplacianorder2local_", "lsm3dcomputeavegradphilocal_", "lsm3dgradientmagnitudelocal_",
    "lsm3dzerooutlevelseteqnrhslocal_", "lsm3daddadvectiontermtolserhslocal_",
    "lsm3daddnormalveltermtolserhslocal_", "lsm3daddconstnormalveltermtolserhslocal_",
    "lsm3daddconstcurvtermtolserhslocal_", "lsm3daddconstprecomputedcurvtermtolserhslocal_",
    "lsm3daddexternalandnormalveltermtolserhslocal_", "lsm3drk1steplocal_", "lsm3dtvdrk2stage1local_",
    "lsm3dtvdrk2stage2local_", "lsm3dtvdrk3stage1local_", "lsm3dtvdrk3stage2local_", "lsm3dtvdrk3stage3local_",
    "lsm3dmaxnormdifflocal_", "lsm3dcomputestableadvectiondtlocal_", "lsm3dcomputestablenormalveldtlocal_",
    "lsm3dcomputereinitializationeqnrhslocal_", "lsm3dcomputemeancurvatureorder2local_",
    "lsm3dmultiplycutofflserhslocal_", "lsm3dimposemasklocal_", "lsm3dcopydatalocal_",
    "lsm3dcomputegaussiancurvatureorder2local_", "lsm3dcomputemeancurvatureorder4local_",
    "lsm3dcomputegaussiancurvatureorder4local_", "lsm3dcomputeorthogonalizationeqnrhslocal_",
]


@pytest.mark.parametrize("level", [1, 3, 4])
def test_determine_narrow_band(level):
    c = H.band_case(level=level)
    for impl_case in (H.band_case(H.reference(), level=level),):
        for k in ("nb", "ix", "iy", "iz", "outer", "n_lo", "n_hi"):
            assert np.array_equal(c.arrays[k], impl_case.arrays[k]), k
        for k in ("nlo_outer_plus", "nhi_outer_plus", "nlo_outer_minus", "nhi_outer_minus"):
            assert c.scalars[k] == impl_case.scalars[k], k


def test_determine_narrow_band_from_mask():
    for use_sign in (0, 1):
        c = H.band_case(level=3)
        c.arrays["mask"] = c.arrays["phi"] - 0.1
        c.boxes["mask"] = c.gb
        c.scalars["use_mask_sign"] = use_sign
        c.scalars.update(nlo=0, nhi=c.arrays["ix"].size - 1)
        H.assert_same(H.run(H.reference(), "lsm3ddeterminenarrowbandfrommask_", c),
                      H.run(H.oracle(), "lsm3ddeterminenarrowbandfrommask_", c), "frommask")


@pytest.mark.parametrize("sym", LOCAL)
def test_local_bitwise(sym):
    c = H.band_case(level=4)
    c.arrays["mask"] = c.rng.standard_normal(c.arrays["phi"].shape)
    c.boxes["mask"] = c.gb
    if "curvature" in sym:
        g = c.array("gmag", "g")
        g.reshape(-1)[::13] = 0.0
    H.assert_same(H.run(H.reference(), sym, c), H.run(H.oracle(), sym, c), sym)


def test_check_outer_layer():
    c = H.band_case(level=3)
    sym = "lsm3dcheckouternarrowbandlayer_"
    a = H.run(H.reference(), sym, c)
    H.assert_same(a, H.run(H.oracle(), sym, c), sym)
    assert a[1] == 0
    # flip the sign of one outer plus point: the check must report a sign change
    l = c.arrays["outer"][c.scalars["nlo_outer_plus"]]
    i, j, k = c.arrays["ix"][l], c.arrays["iy"][l], c.arrays["iz"][l]
    c.arrays["phi"][k + 3, j + 3, i + 3] = -abs(c.arrays["phi"][k + 3, j + 3, i + 3])
    a = H.run(H.reference(), sym, c)
    H.assert_same(a, H.run(H.oracle(), sym, c), sym)
    assert a[1] == 1
