"""The reference-named ghost-fill dispatch over a Grid and the data-array
writers on device arrays, on MI355X.

linearExtrapolationBC / signedLinearExtrapolationBC / copyExtrapolationBC /
homogeneousNeumannBC (src/boundary_conditions/lsm_boundary_conditions.c:19-886)
on a device array equal the oracle's face fills applied in the dispatch order
of oracle/lsm_bc_dispatch.py (which tests/test_reference_utils.py pins to the
reference's compiled dispatch) bit for bit in FP64, and the FP32 library's
own single-face fills in the same order in FP32; writeDataArray given a device
array writes the bytes of the restated format (oracle/lsm_files.py)."""
import ctypes
import os
import sys

import numpy as np
import pytest

import harness as H

sys.path.insert(0, os.path.join(H.ROOT, "oracle"))
import lsm_bc_dispatch as D  # noqa: E402
import lsm_files as LF  # noqa: E402

pytestmark = pytest.mark.gpu

KINDS = ["linearExtrapolationBC", "signedLinearExtrapolationBC", "copyExtrapolationBC", "homogeneousNeumannBC"]


def _lib(prec):
    import lsmlib_amd
    from lsmlib_amd import data_io
    lib = lsmlib_amd.load(prec)
    G = data_io.grid_struct(lib.real)
    GP, RP = ctypes.POINTER(G), ctypes.POINTER(lib.real)
    c = lib.cdll
    c.createGridSetGridDims.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), RP, RP, ctypes.c_int]
    c.createGridSetGridDims.restype = GP
    c.destroyGrid.argtypes = [GP]
    c.destroyGrid.restype = None
    for k in KINDS:
        getattr(c, k).argtypes = [ctypes.c_void_p, GP, ctypes.c_int]
        getattr(c, k).restype = None
    c.writeDataArray.argtypes = [ctypes.c_void_p, GP, ctypes.c_char_p, ctypes.c_int]
    c.writeDataArray.restype = None
    return lib, G


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("num_dims", [3, 2])
@pytest.mark.parametrize("kind", KINDS)
def test_grid_dispatch_equals_faces_in_reference_order(prec, num_dims, kind):
    import torch
    lib, G = _lib(prec)
    R = lib.real
    dims = (13, 10, 7) if num_dims == 3 else (21, 17, 1)
    g = lib.cdll.createGridSetGridDims(num_dims, (ctypes.c_int * 3)(*dims), (R * 3)(0, 0, 0), (R * 3)(1, 1, 1), 3)
    gc = g.contents
    gb6 = (gc.ilo_gb, gc.ihi_gb, gc.jlo_gb, gc.jhi_gb, gc.klo_gb, gc.khi_gb)
    fb6 = (gc.ilo_fb, gc.ihi_fb, gc.jlo_fb, gc.jhi_fb, gc.klo_fb, gc.khi_fb)
    shape = tuple(gc.grid_dims_ghostbox)[::-1]
    dt = torch.float64 if prec == "f64" else torch.float32
    face_sym = (D.FACE_SYMBOL[kind] % num_dims).rstrip("_")
    for idx in range(-1, 11):
        print("dispatch", prec, num_dims, kind, idx, flush=True)
        phi = np.random.default_rng(100 + idx).standard_normal(shape)
        d = torch.tensor(phi, dtype=dt, device="cuda")
        getattr(lib.cdll, kind)(d.data_ptr(), g, idx)
        lib.check()
        got = d.cpu().numpy()
        if prec == "f64":
            want = phi.copy()
            for f in D.faces(num_dims, idx):
                getattr(H.oracle(), face_sym)(want, gb6 if num_dims == 3 else gb6[:4],
                                              fb6 if num_dims == 3 else fb6[:4], f)
        else:
            e = torch.tensor(phi, dtype=dt, device="cuda")
            for f in D.faces(num_dims, idx):
                getattr(lib.api, face_sym)(e, gb6 if num_dims == 3 else gb6[:4], fb6 if num_dims == 3 else fb6[:4], f)
            lib.check()
            want = e.cpu().numpy()
        assert np.array_equal(got, want), (kind, num_dims, idx)
    lib.cdll.destroyGrid(g)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_write_data_array_from_device(prec, tmp_path):
    import torch
    lib, G = _lib(prec)
    R = lib.real
    g = lib.cdll.createGridSetGridDims(3, (ctypes.c_int * 3)(20, 12, 9), (R * 3)(0, 0, 0), (R * 3)(1, 1, 1), 3)
    shape = tuple(g.contents.grid_dims_ghostbox)[::-1]
    dt = torch.float64 if prec == "f64" else torch.float32
    x = torch.randn(shape, dtype=dt, device="cuda")
    path = tmp_path / "d.dat"
    lib.cdll.writeDataArray(x.data_ptr(), g, str(path).encode(), 0)
    lib.check()
    want = LF.data_array_bytes(x.cpu().numpy(), tuple(g.contents.grid_dims_ghostbox),
                               "<f8" if prec == "f64" else "<f4")
    assert path.read_bytes() == want
    lib.cdll.destroyGrid(g)


def test_null_grid_or_array_is_refused():
    """Checked on the host before any launch."""
    import torch
    lib, G = _lib("f64")
    R = lib.real
    g = lib.cdll.createGridSetGridDims(3, (ctypes.c_int * 3)(8, 8, 8), (R * 3)(0, 0, 0), (R * 3)(1, 1, 1), 3)
    d = torch.zeros(tuple(g.contents.grid_dims_ghostbox)[::-1], dtype=torch.float64, device="cuda")
    for k in KINDS:
        getattr(lib.cdll, k)(None, g, 9)
        assert lib.last_error()[0] == 3
        lib.clear_error()
        getattr(lib.cdll, k)(d.data_ptr(), None, 9)
        assert lib.last_error()[0] == 3
        lib.clear_error()
    lib.cdll.destroyGrid(g)
