"""Loader of the in-tree HIP libraries.

``load("f64")`` returns the FP64 drop-in (liblsmhip.so), ``load("f32")`` the
FP32 build (liblsmhip_f32.so).  There is no CPU fallback: a missing library
raises, and every entry point runs on the GPU through the C-ABI.
"""
from __future__ import annotations

import ctypes
import os

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"f64": "liblsmhip.so", "f32": "liblsmhip_f32.so"}


class LsmHipError(RuntimeError):
    pass


class HipLib:
    """One precision of the HIP library: drop-in entry points as attributes
    (``lib.api.lsm3dhjweno5(...)``), plus the library control calls."""

    def __init__(self, precision: str = "f64"):
        if precision not in LIBS:
            raise ValueError("precision must be 'f64' or 'f32'")
        self.precision = precision
        self.path = os.path.join(HERE, LIBS[precision])
        if not os.path.exists(self.path):
            raise LsmHipError("%s is not built; run `python -m lsmlib_amd.build` (hipcc, gfx950)" % self.path)
        # One HIP runtime per process: torch loads its bundled libamdhip64 under
        # the same SONAME (libamdhip64.so.7) that this library depends on, so
        # importing torch first makes the dynamic linker reuse it.  Device
        # pointers and streams are then shared between torch and this library.
        import torch  # noqa: F401
        self.cdll = ctypes.CDLL(self.path)
        self.real = ctypes.c_double if precision == "f64" else ctypes.c_float
        self.api = capi.Bound(self.cdll, real=self.real)
        c = self.cdll
        c.lsmhip_get_last_error.restype = ctypes.c_int
        c.lsmhip_get_last_error_string.restype = ctypes.c_char_p
        c.lsmhip_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_int]
        c.lsmhip_set_stream.restype = None
        c.lsmhip_real_size.restype = ctypes.c_int
        c.lsmhip_check_device_errors.restype = ctypes.c_int
        c.lsmhip_synchronize.restype = ctypes.c_int
        assert c.lsmhip_real_size() == ctypes.sizeof(self.real)

    def last_error(self):
        return self.cdll.lsmhip_get_last_error(), self.cdll.lsmhip_get_last_error_string().decode()

    def clear_error(self):
        self.cdll.lsmhip_clear_error()

    def check(self):
        code, msg = self.last_error()
        if code:
            self.clear_error()
            raise LsmHipError("lsmhip error %d: %s" % (code, msg))

    def set_stream(self, stream_ptr: int = 0, synchronous: bool = True):
        self.cdll.lsmhip_set_stream(ctypes.c_void_p(stream_ptr), int(synchronous))

    def synchronize(self):
        if self.cdll.lsmhip_synchronize() != 0:
            self.check()

    def fn(self, name, argtypes, restype=ctypes.c_int):
        f = getattr(self.cdll, name)
        f.argtypes = argtypes
        f.restype = restype
        return f


_loaded = {}


def load(precision: str = "f64") -> HipLib:
    if precision not in _loaded:
        _loaded[precision] = HipLib(precision)
    return _loaded[precision]
