"""HIP narrow-band kernels and band construction (liblsmhip.so, FP64) against
the oracle on the same inputs: bitwise for values, index lists and masks."""
import numpy as np
import pytest

import harness as H
from test_oracle_vs_reference_local import LOCAL

pytestmark = pytest.mark.gpu

SUMS = {"lsm3dcomputeavegradphilocal_"}


@pytest.fixture(scope="module")
def hip():
    import lsmlib_amd
    lib = lsmlib_amd.load("f64")
    lib.set_stream(0, True)
    return lib


def _local_case(level=4, seed=0):
    c = H.band_case(level=level, seed=seed)
    c.arrays["mask"] = c.rng.standard_normal(c.arrays["phi"].shape)
    c.boxes["mask"] = c.gb
    return c


@pytest.mark.parametrize("sym", LOCAL + ["lsm3dhjweno5local_", "lsm3dcomputestableconstnormalveldtlocal_"])
def test_local_bitwise(hip, sym):
    c = _local_case()
    if "curvature" in sym:
        c.array("gmag", "g").reshape(-1)[::13] = 0.0
    a = H.run(H.oracle(), sym, c)
    b = H.run(hip.api, sym, c, device="cuda")
    hip.check()
    H.assert_same(a, b, sym, out_abs=1e-13 if sym in SUMS else None)


@pytest.mark.parametrize("level", [1, 3, 4])
@pytest.mark.parametrize("seed", [0, 1])
def test_determine_narrow_band_bitwise(hip, level, seed):
    ref = H.band_case(level=level, seed=seed)
    got = H.band_case(hip.api, level=level, seed=seed, device="cuda")
    for k in ("nb", "ix", "iy", "iz", "outer", "n_lo", "n_hi"):
        assert np.array_equal(ref.arrays[k], got.arrays[k]), k
    for k in ("nlo_outer_plus", "nhi_outer_plus", "nlo_outer_minus", "nhi_outer_minus"):
        assert ref.scalars[k] == got.scalars[k], k


@pytest.mark.parametrize("level", [1, 3])
def test_determine_narrow_band_with_a_wider_mark_array(hip, level):
    """narrow_band on a ghost box wider than phi's (two cells more on some
    sides, one on others): the marks are addressed through their own box and
    the cells outside phi's box keep their values; lists and marks bitwise the
    oracle's."""
    c = H.band_case(level=level, seed=5)
    gb = c.boxes["gb"]
    wide = (gb[0] - 2, gb[1] + 1, gb[2] - 1, gb[3] + 2, gb[4] - 2, gb[5] + 1)
    c.boxes["nb"] = wide
    c.arrays["nb"] = np.full(H.shape_of(wide), 9, np.uint8)
    npts = c.arrays["phi"].size
    for nm in ("ix", "iy", "iz", "outer"):
        c.arrays[nm] = np.full(npts, -7, np.int32)
    c.arrays["n_lo"] = np.full(level + 1, -5, np.int32)
    c.arrays["n_hi"] = np.full(level + 1, -5, np.int32)
    c.scalars.update(nlo=0, nhi=npts - 1, level=level)
    sym = "lsm3ddeterminenarrowband_"
    a = H.run(H.oracle(), sym, c)
    b = H.run(hip.api, sym, c, device="cuda")
    hip.check()
    for k in ("nb", "ix", "iy", "iz", "outer", "n_lo", "n_hi"):
        assert np.array_equal(a[0][k], b[0][k]), k
    assert a[1] == b[1]
    assert (a[0]["nb"] == 9).sum() > 0


def test_band_from_mask_and_outer_check(hip):
    for use_sign in (0, 1):
        c = _local_case(level=3)
        c.arrays["mask"] = c.arrays["phi"] - 0.05
        c.scalars["use_mask_sign"] = use_sign
        c.scalars.update(nlo=0, nhi=c.arrays["ix"].size - 1)  # index-array bounds, not a level range
        sym = "lsm3ddeterminenarrowbandfrommask_"
        H.assert_same(H.run(H.oracle(), sym, c), _run_mixed(hip, sym, c), sym)
    c = _local_case(level=3)
    sym = "lsm3dcheckouternarrowbandlayer_"
    a = H.run(H.oracle(), sym, c)
    assert a[1] == _run_mixed(hip, sym, c)[1] == 0
    l = c.arrays["outer"][c.scalars["nlo_outer_plus"]]
    i, j, k = c.arrays["ix"][l], c.arrays["iy"][l], c.arrays["iz"][l]
    c.arrays["phi"][k + 3, j + 3, i + 3] = -abs(c.arrays["phi"][k + 3, j + 3, i + 3])
    assert _run_mixed(hip, sym, c)[1] == H.run(H.oracle(), sym, c)[1] == 1


def test_bad_index_is_reported_not_read(hip):
    c = _local_case()
    c.arrays["ix"] = c.arrays["ix"].copy()
    c.arrays["ix"][0] = 10_000  # far outside every ghost box
    H.run(hip.api, "lsm3drk1steplocal_", c, device="cuda")
    code, msg = hip.last_error()
    hip.clear_error()
    assert code == 4, (code, msg)


class _DeviceBand:
    """Adapter: run band construction through the HIP library inside band_case."""

    def __init__(self, hip):
        self.hip = hip

    def __getattr__(self, name):
        fn = getattr(self.hip.api, name)

        def call(*args):
            return fn(*args)
        return call


def _run_mixed(hip, sym, case):
    r = H.run(hip.api, sym, case, device="cuda")
    hip.check()
    return r
