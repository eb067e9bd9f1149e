"""The oracle's per-cell addends of the smoothed volume and area sums and of
the band mean of |grad phi| (what tests/test_gpu_integrals.py sums exactly)
are the addends of the oracle's own sums -- which are bitwise the reference's
(tests/test_oracle_vs_reference*.py): summed serially in loop order they give
the oracle's result bit for bit.  On these >10^6-cell inputs the serial sum
is many ulps off the exactly rounded sum (math.fsum), so the GPU pin (2 ulp of
math.fsum) separates a compensated sum from a naive one."""
import math

import numpy as np
import pytest

import harness as H
import integral_cases as IC

SYMS = {0: "lsm3dvolumeregionphilessthanzero", 1: "lsm3dvolumeregionphigreaterthanzero",
        2: "lsm3dsurfaceareazerolevelset"}


@pytest.mark.parametrize("which", [0, 1, 2])
def test_terms_are_the_addends_of_the_oracle_sum(which):
    phi, grads = IC.volume_case()
    terms = IC.integral_terms(which, phi, grads)
    args = [None, phi, IC.GB] + ([grads[0], grads[1], grads[2], IC.GB] if which == 2 else []) + [IC.IB, IC.H_, IC.H_,
                                                                                          IC.H_, IC.EPS]
    want = getattr(H.oracle(), SYMS[which])(*args)
    assert IC.serial(terms) == want
    exact = math.fsum(terms)
    print("%s: serial %.17g exact %.17g (%.0f ulp apart)" % (SYMS[which], want, exact, IC.ulps(want, exact)))
    assert IC.ulps(want, exact) > 20  # the naive sum is measurably off on these inputs


def test_avegrad_terms_are_the_addends_of_the_oracle_mean():
    phi = IC.avegrad_case()
    ix, iy, iz, nb = IC.band_all(phi.shape)
    terms = IC.avegrad_terms(phi, ix, iy, iz, nb)
    assert terms.size == ix.size
    want = H.oracle().lsm3dcomputeavegradphilocal(None, phi, IC.GB, IC.H_, IC.H_, IC.H_, ix, iy, iz, 0, ix.size - 1, nb,
                                                  IC.GB, 1)
    assert IC.serial(terms) / terms.size == want
    exact = math.fsum(terms) / terms.size
    print("ave |grad phi|: serial %.17g exact %.17g (%.0f ulp apart)" % (want, exact, IC.ulps(want, exact)))
    assert IC.ulps(want, exact) > 20


def test_parity_case_serial_sums_are_within_the_absolute_bar():
    """On the small parity cases (tests/test_gpu_parity_*.py) the reference's
    serial sums lie within 5e-14 of the exact sums, so holding the GPU's
    compensated sums to 1e-13 of the serial ones there is meaningful."""
    c = H.grid_case(1)
    gb, fb = c.gb, c.fb
    phi = c.arrays["phi"]
    for which in (0, 1):
        s = getattr(H.oracle(), SYMS[which])(None, phi, gb, fb, 0.1, 0.125, 0.15, 0.3)
        import ctypes
        lib = ctypes.CDLL(H.ORACLE_SO)
        n = (fb[1] - fb[0] + 1) * (fb[3] - fb[2] + 1) * (fb[5] - fb[4] + 1)
        out = np.empty(n)
        k1, g = IC._ints(*gb)
        k2, f = IC._ints(*fb)
        d = [ctypes.c_double(v) for v in (0.1, 0.125, 0.15, 0.3)]
        lib.oracle_integral_terms.restype = None
        lib.oracle_integral_terms(IC._ptr(out), ctypes.c_int(which), IC._ptr(phi), *g, IC._ptr(phi), IC._ptr(phi),
                                  IC._ptr(phi), *g, *f, *[ctypes.byref(x) for x in d])
        assert abs(s - math.fsum(out)) < 5e-14
