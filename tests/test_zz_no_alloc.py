"""lsmhip_volume_region_phi_lt0_noalloc, the allocation-free variant of
lsm3dvolumeregionphilessthanzero_: its partial sums live in a file-scope
device array instead of library workspace.  Through the C-ABI on device
arrays, two calls in a row (the second on a smaller box, so the array holds
the first call's partials when the second starts) give the existing entry
point's result bit for bit (FP64 and FP32), and in FP64 lie within 2 ulp of
the exactly rounded sum (math.fsum) of the oracle's per-cell addends; a box of
2^31 cells or more is refused before anything is launched."""
import ctypes
import math

import numpy as np
import pytest

import harness as H
import integral_cases as IC


def _fns(lib):
    R, RP = lib.real, ctypes.POINTER(lib.real)
    box = [ctypes.POINTER(ctypes.c_int)] * 6
    noalloc = lib.fn("lsmhip_volume_region_phi_lt0_noalloc", [RP, ctypes.c_void_p] + box + box + [RP] * 4)
    ref = lib.cdll.lsm3dvolumeregionphilessthanzero_
    ref.argtypes = [RP, ctypes.c_void_p] + box + box + [RP] * 4
    ref.restype = None
    return noalloc, ref


def _b(v):
    return [ctypes.byref(ctypes.c_int(x)) for x in v]


def _s(real, *v):
    return [ctypes.byref(real(x)) for x in v]


CASES = [((-3, 130, -3, 100, -3, 90), (0, 127, 0, 97, 0, 87), 0),   # 1.1e6 cells, every block of the cap used
         ((-3, 30, -2, 17, -3, 12), (1, 25, 0, 14, -1, 9), 1)]       # a smaller box second


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_volume_equals_the_existing_entry_point_twice(prec):
    import torch
    import lsmlib_amd
    lib = lsmlib_amd.load(prec)
    lib.set_stream(0, True)
    noalloc, ref = _fns(lib)
    dt = torch.float64 if prec == "f64" else torch.float32
    for gb, ib, seed in CASES:
        print("noalloc", prec, gb, ib, flush=True)
        rng = np.random.default_rng(seed)
        phi = 0.2 * rng.standard_normal(H.shape_of(gb)) - 0.05
        d = torch.tensor(phi, dtype=dt, device="cuda")
        h = 0.07
        a, b = lib.real(-1.0), lib.real(-2.0)
        assert noalloc(ctypes.byref(a), d.data_ptr(), *_b(gb), *_b(ib), *_s(lib.real, h, h, h, 1.5 * h)) == 0
        lib.check()
        ref(ctypes.byref(b), d.data_ptr(), *_b(gb), *_b(ib), *_s(lib.real, h, h, h, 1.5 * h))
        lib.check()
        assert a.value == b.value and a.value > 0
        if prec == "f64":
            exact = math.fsum(IC.integral_terms(0, phi, gbox=gb, ibox=ib, h=h, eps=1.5 * h))
            assert IC.ulps(a.value, exact) <= 2, (a.value, exact)


def test_noalloc_volume_refuses_a_box_over_the_limit_before_launching():
    """Checked on the host: the error is LSMHIP_ERR_SIZE (5), not a HIP error,
    so nothing reached the device; the array is never dereferenced."""
    import lsmlib_amd
    lib = lsmlib_amd.load("f64")
    noalloc, _ = _fns(lib)
    huge = (0, 2047, 0, 2047, 0, 1023)  # 2^32 cells
    dummy = np.zeros(8)
    out = lib.real(0.0)
    assert noalloc(ctypes.byref(out), dummy.ctypes.data, *_b(huge), *_b(huge), *_s(lib.real, 1, 1, 1, 1)) == -1
    code, msg = lib.last_error()
    lib.clear_error()
    assert code == 5 and "2^31" in msg
