"""NULL device arrays are refused on the host, before anything is launched.

A kernel that dereferences a NULL array faults the device.  Every exported
entry point therefore tests its device-array arguments first and reports
LSMHIP_ERR_ARG.  These tests call the raw C symbols through ctypes (not the
binder, which refuses NULL itself), one NULL array at a time, with every other
pointer argument non-NULL.

They run only where no GPU is visible: the other device arrays are stand-in
addresses that must never reach a kernel, and without a device nothing can be
launched even if a check were missing (the call would then report a HIP error,
not LSMHIP_ERR_ARG, and the test would fail).
"""
import ctypes
import os

import numpy as np
import pytest

from lsmlib_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBS = {"f64": (os.path.join(ROOT, "lsmlib_amd", "liblsmhip.so"), ctypes.c_double),
        "f32": (os.path.join(ROOT, "lsmlib_amd", "liblsmhip_f32.so"), ctypes.c_float)}
ERR_ARG = 3
FAKE = 0x7F0000001000  # a stand-in device address: never dereferenced when a NULL is refused first


def _no_gpu():
    import torch
    return not torch.cuda.is_available()


pytestmark = pytest.mark.skipif(not _no_gpu(), reason="stand-in device addresses must not reach a GPU")


def _lib(prec):
    path, real = LIBS[prec]
    if not os.path.exists(path):
        pytest.skip("not built")
    import torch  # noqa: F401  (the library shares torch's HIP runtime)
    lib = ctypes.CDLL(path)
    lib.lsmhip_get_last_error.restype = ctypes.c_int
    lib.lsmhip_get_last_error_string.restype = ctypes.c_char_p
    return lib, real


def _args(sym, real, null_at):
    """ctypes arguments of sym: host scalars and boxes valid, device arrays
    stand-ins, the device array at position null_at NULL."""
    keep, out = [], []
    box = (0, 9, 0, 9, 0, 9)
    for q, (kind, _name) in enumerate(capi.SPECS[sym]):
        if kind in ("B", "Q"):
            for v in box[:6 if kind == "B" else 4]:
                c = ctypes.c_int(v)
                keep.append(c)
                out.append(ctypes.byref(c))
        elif kind in ("S", "SO"):
            c = real(1.0)
            keep.append(c)
            out.append(ctypes.byref(c))
        elif kind in ("I", "IO"):
            c = ctypes.c_int(0)
            keep.append(c)
            out.append(ctypes.byref(c))
        elif kind == "U8":
            c = ctypes.c_ubyte(1)
            keep.append(c)
            out.append(ctypes.byref(c))
        elif kind == "IH":
            c = (ctypes.c_int * 8)(*([0] * 8))
            keep.append(c)
            out.append(c)
        else:  # R, IA, M: device arrays
            out.append(ctypes.c_void_p(None if q == null_at else FAKE))
    return keep, out


def _device_positions(sym):
    """Device arrays the call dereferences.  The whole-grid derivatives form
    their differences in registers and never touch the caller's D1/D2/D3
    scratch, so those may be anything, NULL included.  phi0 of the
    reinitialization is read only under its flag (the next test).  The
    first-order delta function never reads norm_phi_z (its z terms use
    norm_phi_x, as the reference's do)."""
    unused = ("phi0",) + (("D1", "D2", "D3") if not sym.endswith("local_") else ())
    if sym == "lsm3ddeltafunctionorder1_":
        unused += ("nz",)
    return [q for q, (kind, name) in enumerate(capi.SPECS[sym]) if kind in ("R", "IA", "M") and name not in unused]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_every_drop_in_refuses_each_null_array(prec):
    lib, real = _lib(prec)
    bad = []
    checked = 0
    for sym in capi.SPECS:
        for q in _device_positions(sym):
            lib.lsmhip_clear_error()
            keep, args = _args(sym, real, q)
            f = getattr(lib, sym)
            f.restype = None
            f(*args)
            code = lib.lsmhip_get_last_error()
            msg = lib.lsmhip_get_last_error_string().decode()
            name = capi.SPECS[sym][q][1]
            if code != ERR_ARG or "NULL" not in msg:
                bad.append((sym, name, code, msg))
            checked += 1
    lib.lsmhip_clear_error()
    assert not bad, bad[:10]
    assert checked > 600


def test_reinitialization_phi0_is_optional_without_the_flag():
    """phi0 is read only when use_phi0_for_sgn == 1 (the reference's routine
    does not touch it otherwise), so a NULL phi0 is refused only then."""
    lib, real = _lib("f64")
    for sym in ("lsm3dcomputereinitializationeqnrhs_", "lsm2dcomputereinitializationeqnrhs_",
                "lsm3dcomputereinitializationeqnrhslocal_", "lsm2dcomputereinitializationeqnrhslocal_"):
        spec = capi.SPECS[sym]
        q = [n for _, n in spec].index("phi0")
        flag = [n for _, n in spec].index("use_phi0")
        for use, refused in ((1, True), (0, False)):
            lib.lsmhip_clear_error()
            keep, args = _args(sym, real, q)
            # the flag's argument position after box expansion
            pos = sum({"B": 6, "Q": 4}.get(k, 1) for k, _ in spec[:flag])
            c = ctypes.c_int(use)
            keep.append(c)
            args[pos] = ctypes.byref(c)
            if not refused:
                # the call passes the argument checks: give it an empty fill box (or an empty index
                # range) so that nothing can be launched whatever happens next
                names = [n for _, n in spec]
                lo, hi = ctypes.c_int(5), ctypes.c_int(4)
                keep += [lo, hi]
                if "fb" in names:
                    p0 = sum({"B": 6, "Q": 4}.get(k, 1) for k, _ in spec[:names.index("fb")])
                    args[p0], args[p0 + 1] = ctypes.byref(lo), ctypes.byref(hi)
                else:
                    for n, c in (("nlo", lo), ("nhi", hi)):
                        args[sum({"B": 6, "Q": 4}.get(k, 1) for k, _ in spec[:names.index(n)])] = ctypes.byref(c)
            f = getattr(lib, sym)
            f.restype = None
            f(*args)
            code = lib.lsmhip_get_last_error()
            assert (code == ERR_ARG) == refused, (sym, use, code, lib.lsmhip_get_last_error_string())
    lib.lsmhip_clear_error()


def _call_int(lib, name, *args):
    f = getattr(lib, name)
    f.restype = ctypes.c_int
    lib.lsmhip_clear_error()
    rc = f(*args)
    return rc, lib.lsmhip_get_last_error(), lib.lsmhip_get_last_error_string().decode()


def _boxp(*v):
    cs = [ctypes.c_int(x) for x in v]
    return cs, [ctypes.byref(c) for c in cs]


def test_fused_and_control_entry_points_refuse_null_arrays():
    lib, real = _lib("f64")
    keep, gb = _boxp(0, 15, 0, 15, 0, 15)
    keep2, fb = _boxp(3, 12, 3, 12, 3, 12)
    fake = ctypes.c_void_p(FAKE)
    nul = ctypes.c_void_p(None)
    d = real(0.1)
    for out, us in ((nul, fake), (fake, nul)):
        rc, code, msg = _call_int(lib, "lsmhip_weno5_stage_normal", out, us, fake, *gb, *fb, ctypes.c_int(1),
                                  real(1.0), nul, d, d, d, d)
        assert (rc, code) == (-1, ERR_ARG), msg
        rc, code, msg = _call_int(lib, "lsmhip_weno5_stage_reinit", out, us, fake, *gb, *fb, ctypes.c_int(1),
                                  nul, ctypes.c_int(0), d, d, d, d)
        assert (rc, code) == (-1, ERR_ARG), msg
    for q in range(3):
        v = [fake, fake, fake]
        v[q] = nul
        rc, code, msg = _call_int(lib, "lsmhip_weno5_stage_advect", fake, fake, fake, *gb, *fb, ctypes.c_int(1),
                                  *v, d, d, d, d)
        assert (rc, code) == (-1, ERR_ARG) and "NULL" in msg, msg
    # one NULL entry of the deformation tables
    tabs = (ctypes.c_void_p * 6)(*([FAKE] * 5 + [0]))
    rc, code, msg = _call_int(lib, "lsmhip_weno5_stage_advect_deform", fake, fake, fake, *gb, *fb, ctypes.c_int(1),
                              tabs, d, d, d, d, d)
    assert (rc, code) == (-1, ERR_ARG) and "tabs[5]" in msg, msg
    rc, code, msg = _call_int(lib, "lsmhip_fill_ghosts", nul, *gb, *fb, ctypes.c_int(9), ctypes.c_int(0))
    assert (rc, code) == (-1, ERR_ARG) and "phi" in msg, msg
    dd = ctypes.c_double(0.1)
    rc, code, msg = _call_int(lib, "lsmhip_volume_phi_lt0_f64", fake, nul, *gb, *fb, ctypes.byref(dd),
                              ctypes.byref(dd), ctypes.byref(dd), ctypes.byref(dd))
    assert (rc, code) == (-1, ERR_ARG) and "phi" in msg, msg
    lib.lsmhip_clear_error()


def test_binder_refuses_a_null_array():
    """The ctypes binder refuses None for an array before it calls anything."""
    import lsmlib_amd
    lib = lsmlib_amd.load("f64")
    a = np.zeros(8)
    with pytest.raises(ValueError, match="NULL"):
        lib.api.lsm3dcopydatalocal(None, (0, 1, 0, 1, 0, 1), a, (0, 1, 0, 1, 0, 1), a, a, a, 0, 0)
