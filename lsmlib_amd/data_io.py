"""Python side of the reference's data-array and grid files (include/lsm_data_io.h).

write_data_array / read_data_array move a field between a file in the
reference's format (three ints, then the reals, column-major) and a numpy
array or a torch tensor, on the host or the GPU; make_grid / write_grid /
read_grid build and store the reference's Grid record
(src/utils/lsm_grid.h:43-95).  All work is done by the HIP library.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import LsmHipError, load

ACCURACY = {"eno1": 0, "eno2": 1, "eno3": 2, "weno5": 3}


_GRID_TYPES = {}


def grid_struct(real):
    """ctypes mirror of LsmhipGrid (the reference's Grid) for LSMLIB_REAL = real
    (c_double or c_float); one class per real type."""
    if real in _GRID_TYPES:
        return _GRID_TYPES[real]
    fields = [("num_dims", ctypes.c_int)]
    fields += [(k, real * 3) for k in ("x_lo", "x_hi", "x_lo_ghostbox", "x_hi_ghostbox")]
    fields += [("grid_dims", ctypes.c_int * 3), ("grid_dims_ghostbox", ctypes.c_int * 3), ("dx", real * 3),
               ("num_gridpts", ctypes.c_int)]
    for s in ("gb", "fb", "D1_fb", "D2_fb", "D3_fb"):
        fields += [(p + s, ctypes.c_int) for p in ("ilo_", "ihi_", "jlo_", "jhi_", "klo_", "khi_")]
    fields += [("num_nb_levels", ctypes.c_int)]
    fields += [(k, ctypes.c_ubyte) for k in ("mark_gb", "mark_D1", "mark_D2", "mark_D3", "mark_fb")]
    fields += [("beta", real), ("gamma", real)]
    _GRID_TYPES[real] = type("LsmhipGrid_%s" % real.__name__, (ctypes.Structure,), {"_fields_": fields})
    return _GRID_TYPES[real]


_STRUCTS = {}


def _lib(precision):
    lib = load(precision)
    if precision not in _STRUCTS:
        G = grid_struct(lib.real)
        P, I, L, C = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_longlong, ctypes.c_char_p
        R, RP, GP = lib.real, ctypes.POINTER(lib.real), ctypes.POINTER(G)
        _STRUCTS[precision] = {
            "G": G,
            "write": lib.fn("lsmhip_write_data_array", [P, I, C, ctypes.c_int]),
            "dims": lib.fn("lsmhip_read_data_array_dims", [I, C]),
            "read": lib.fn("lsmhip_read_data_array", [P, L, I, C, ctypes.c_int]),
            "set_dims": lib.fn("lsmhip_grid_set_grid_dims", [GP, ctypes.c_int, I, RP, RP, ctypes.c_int]),
            "set_dx": lib.fn("lsmhip_grid_set_dx", [GP, ctypes.c_int, R, RP, RP, ctypes.c_int]),
            "wgrid": lib.fn("lsmhip_write_grid", [GP, C]),
            "rgrid": lib.fn("lsmhip_read_grid", [GP, C]),
        }
    return lib, _STRUCTS[precision]


def _check(lib, rc):
    if rc != 0:
        lib.check()
        raise LsmHipError("file entry point failed")


def _ptr_and_device(a):
    if isinstance(a, np.ndarray):
        return a.ctypes.data, 0
    return a.data_ptr(), int(a.is_cuda)


def write_data_array(a, path, precision="f64"):
    """a: (nz, ny, nx) C-ordered numpy array or torch tensor of LSMLIB_REAL."""
    lib, F = _lib(precision)
    if isinstance(a, np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64 if precision == "f64" else np.float32)
    ptr, dev = _ptr_and_device(a)
    dims = (ctypes.c_int * 3)(a.shape[2], a.shape[1], a.shape[0])
    _check(lib, F["write"](ptr, dims, str(path).encode(), dev))


def read_data_array(path, precision="f64", device=None):
    """Returns the array as (nz, ny, nx): numpy on the host, or a torch tensor on `device`."""
    lib, F = _lib(precision)
    dims = (ctypes.c_int * 3)()
    _check(lib, F["dims"](dims, str(path).encode()))
    shape = (dims[2], dims[1], dims[0])
    if device is None:
        out = np.empty(shape, dtype=np.float64 if precision == "f64" else np.float32)
    else:
        import torch
        out = torch.empty(shape, dtype=torch.float64 if precision == "f64" else torch.float32, device=device)
    ptr, dev = _ptr_and_device(out)
    _check(lib, F["read"](ptr, int(np.prod(shape)), dims, str(path).encode(), dev))
    return out


def make_grid(num_dims, x_lo, x_hi, accuracy="weno5", grid_dims=None, dx=None, precision="f64"):
    """createGridSetGridDims (grid_dims given) or createGridSetDx (dx given)."""
    lib, F = _lib(precision)
    g = F["G"]()
    R3 = lib.real * 3
    lo, hi = R3(*x_lo), R3(*x_hi)
    acc = ACCURACY[accuracy] if isinstance(accuracy, str) else int(accuracy)
    if grid_dims is not None:
        _check(lib, F["set_dims"](ctypes.byref(g), num_dims, (ctypes.c_int * 3)(*grid_dims), lo, hi, acc))
    else:
        _check(lib, F["set_dx"](ctypes.byref(g), num_dims, dx, lo, hi, acc))
    return g


def write_grid(g, path, precision="f64"):
    lib, F = _lib(precision)
    _check(lib, F["wgrid"](ctypes.byref(g), str(path).encode()))


def read_grid(path, precision="f64"):
    lib, F = _lib(precision)
    g = F["G"]()
    _check(lib, F["rgrid"](ctypes.byref(g), str(path).encode()))
    return g


def grid_dict(g):
    """The grid record as a dict of Python values (arrays as lists)."""
    out = {}
    for name, _ in g._fields_:
        v = getattr(g, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out
