"""The reference's C utilities on the path, under their own names: the ghost-fill
dispatch over a Grid (src/boundary_conditions/lsm_boundary_conditions.c), the
Grid constructors and grid file (src/utils/lsm_grid.c) and the data-array files
(src/utils/lsm_data_arrays.c).  CPU.

The reference's own C is compiled into oracle/_ref by oracle/Makefile; these
tests pin

  * the oracle's dispatch order (oracle/lsm_bc_dispatch.py with the oracle's
    face fills) against the reference's dispatch, every kind, location index
    and dimension, bit for bit; the GPU test applies the same order to the
    HIP library's dispatch;
  * the HIP library's Grid constructors, grid files and data-array files
    against the reference's, byte for byte, with and without compression;
  * the restated file formats (oracle/lsm_files.py) against the reference's.
"""
import ctypes
import gzip
import bz2
import os
import shutil
import sys

import numpy as np
import pytest

import harness as H

sys.path.insert(0, os.path.join(H.ROOT, "oracle"))
import lsm_bc_dispatch as D  # noqa: E402
import lsm_files as LF  # noqa: E402

from lsmlib_amd import data_io  # noqa: E402

pytestmark = pytest.mark.skipif(not os.path.exists(H.REF_SO), reason="reference build (oracle/_ref) absent")

NO_ZIP, GZIP, BZIP2 = 0, 1, 2


def _bind(lib, real):
    G = data_io.grid_struct(real)
    GP, RP, IP, C = ctypes.POINTER(G), ctypes.POINTER(real), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p
    sig = {
        "createGridSetDx": ([ctypes.c_int, real, RP, RP, ctypes.c_int], GP),
        "createGridSetDxDyDz": ([ctypes.c_int, real, real, real, RP, RP, ctypes.c_int], GP),
        "createGridSetGridDims": ([ctypes.c_int, IP, RP, RP, ctypes.c_int], GP),
        "writeGridToBinaryFile": ([GP, C, ctypes.c_int], None),
        "readGridFromBinaryFile": ([C], GP),
        "writeDataArray": ([ctypes.c_void_p, GP, C, ctypes.c_int], None),
        "writeDataArrayNoGrid": ([ctypes.c_void_p, IP, C, ctypes.c_int], None),
        "readDataArray": ([IP, C], RP),
        "linearExtrapolationBC": ([ctypes.c_void_p, GP, ctypes.c_int], None),
        "signedLinearExtrapolationBC": ([ctypes.c_void_p, GP, ctypes.c_int], None),
        "copyExtrapolationBC": ([ctypes.c_void_p, GP, ctypes.c_int], None),
        "homogeneousNeumannBC": ([ctypes.c_void_p, GP, ctypes.c_int], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib, G


@pytest.fixture(scope="module")
def ref():
    return _bind(ctypes.CDLL(H.REF_SO), ctypes.c_double)


@pytest.fixture(scope="module", params=["f64", "f32"])
def hip(request):
    import lsmlib_amd
    lib = lsmlib_amd.load(request.param)
    lib.cdll.lsmhip_get_last_error.restype = ctypes.c_int
    return request.param, _bind(lib.cdll, lib.real)[0], lib


libc = ctypes.CDLL(None)
libc.free.argtypes = [ctypes.c_void_p]
libc.fflush.argtypes = [ctypes.c_void_p]


def _grid_bytes(gp, G):
    return ctypes.string_at(gp, ctypes.sizeof(G))


def _r3(real, v):
    return (real * 3)(*v)


GRIDS = [(3, (7, 6, 5), (0.0, -0.5, 0.25), (1.0, 0.75, 2.0)), (2, (9, 8, 1), (-1.0, 0.0, 0.0), (1.0, 2.0, 0.0)),
         (4, (6, 5, 4), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))]


@pytest.mark.parametrize("case", range(len(GRIDS)))
@pytest.mark.parametrize("accuracy", [0, 1, 2, 3])
def test_grid_constructors_equal_the_reference(ref, accuracy, case):
    """createGridSetGridDims / SetDx / SetDxDyDz: the whole Grid struct equal
    byte for byte to the reference's (FP64), including a num_dims of 4 that both
    replace by 2."""
    rlib, G = ref
    import lsmlib_amd
    hl = _bind(lsmlib_amd.load("f64").cdll, ctypes.c_double)[0]
    nd, dims, lo, hi = GRIDS[case]
    R = ctypes.c_double
    calls = [("createGridSetGridDims", lambda L: L.createGridSetGridDims(nd, (ctypes.c_int * 3)(*dims), _r3(R, lo),
                                                                          _r3(R, hi), accuracy)),
             ("createGridSetDx", lambda L: L.createGridSetDx(nd, 0.125, _r3(R, lo), _r3(R, hi), accuracy)),
             ("createGridSetDxDyDz", lambda L: L.createGridSetDxDyDz(nd, 0.125, 0.1, 0.25, _r3(R, lo), _r3(R, hi),
                                                                      accuracy))]
    for name, call in calls:
        a, b = call(rlib), call(hl)
        assert _grid_bytes(a, G) == _grid_bytes(b, G), name
        libc.free(ctypes.cast(a, ctypes.c_void_p))
        hl.destroyGrid.argtypes = [ctypes.c_void_p]
        hl.destroyGrid(ctypes.cast(b, ctypes.c_void_p))


def test_grid_constructor_fp32_equals_the_restatement():
    """The FP32 library's grids equal the file-format restatement's, which the
    FP64 test above pins to the reference's constructors."""
    import lsmlib_amd
    hl, G = _bind(lsmlib_amd.load("f32").cdll, ctypes.c_float)
    for nd, dims, lo, hi in GRIDS[:2]:
        for acc in range(4):
            g = hl.createGridSetGridDims(nd, (ctypes.c_int * 3)(*dims), _r3(ctypes.c_float, lo),
                                         _r3(ctypes.c_float, hi), acc)
            want = LF.grid_set_grid_dims(nd, dims[:nd], lo[:nd], hi[:nd], acc, LF.fp32)
            got = data_io.grid_dict(g.contents)
            for k, v in want.items():
                assert got[k] == v, (k, got[k], v)
            libc.free(ctypes.cast(g, ctypes.c_void_p))


def test_invalid_accuracy_gives_null(hip):
    prec, hl, lib = hip
    R = lib.real
    assert not hl.createGridSetDx(3, 0.1, _r3(R, (0, 0, 0)), _r3(R, (1, 1, 1)), 4)
    assert lib.cdll.lsmhip_get_last_error() == 3
    lib.clear_error()


def _unpack(path, zip_status):
    raw = open(path, "rb").read()
    return raw if zip_status == NO_ZIP else (gzip.decompress(raw) if zip_status == GZIP else bz2.decompress(raw))


ZIPS = [(NO_ZIP, ""), (GZIP, ".gz"), (BZIP2, ".bz2")]


@pytest.mark.skipif(shutil.which("gzip") is None or shutil.which("bzip2") is None, reason="needs gzip and bzip2")
@pytest.mark.parametrize("zip_status,suffix", ZIPS)
def test_grid_file_equals_the_reference(ref, tmp_path, zip_status, suffix):
    """writeGridToBinaryFile: the same record as the reference's (after
    decompression: gzip headers carry a time stamp), and each library reads the
    other's file, compressed files included, to the same Grid."""
    rlib, G = ref
    import lsmlib_amd
    hl = _bind(lsmlib_amd.load("f64").cdll, ctypes.c_double)[0]
    R = ctypes.c_double
    g = rlib.createGridSetGridDims(3, (ctypes.c_int * 3)(10, 12, 14), _r3(R, (0, 0, 0)), _r3(R, (1, 1.2, 1.4)), 3)
    pr, ph = str(tmp_path / "r.bin"), str(tmp_path / "h.bin")
    rlib.writeGridToBinaryFile(g, pr.encode(), zip_status)
    hl.writeGridToBinaryFile(g, ph.encode(), zip_status)
    assert _unpack(pr + suffix, zip_status) == _unpack(ph + suffix, zip_status)
    assert _unpack(ph + suffix, zip_status) == LF.grid_bytes(data_io.grid_dict(g.contents))
    for reader in (rlib, hl):
        for path in (pr, ph):
            back = reader.readGridFromBinaryFile((path + suffix).encode())
            assert _grid_bytes(back, G) == _grid_bytes(g, G)
            libc.free(ctypes.cast(back, ctypes.c_void_p))
            assert os.path.exists(path + suffix) and (not suffix or not os.path.exists(path))  # zipped again
    libc.free(ctypes.cast(g, ctypes.c_void_p))


@pytest.mark.skipif(shutil.which("gzip") is None or shutil.which("bzip2") is None, reason="needs gzip and bzip2")
@pytest.mark.parametrize("zip_status,suffix", ZIPS)
def test_data_array_files_equal_the_reference(ref, hip, tmp_path, zip_status, suffix):
    """writeDataArray / writeDataArrayNoGrid write the reference's bytes;
    readDataArray reads the reference's files (compressed too) to the same
    values and dimensions, in host memory released with free()."""
    rlib, G = ref
    prec, hl, lib = hip
    R = lib.real
    npr = np.float64 if prec == "f64" else np.float32
    dims = (11, 9, 7)
    data = np.random.default_rng(3).standard_normal(dims[::-1]).astype(npr)
    grid = hl.createGridSetGridDims(3, (ctypes.c_int * 3)(5, 3, 1), _r3(R, (0, 0, 0)), _r3(R, (1, 1, 1)), 1)
    assert tuple(grid.contents.grid_dims_ghostbox) == dims
    ph, pn = str(tmp_path / "h.dat"), str(tmp_path / "n.dat")
    hl.writeDataArray(data.ctypes.data, grid, ph.encode(), zip_status)
    hl.writeDataArrayNoGrid(data.ctypes.data, (ctypes.c_int * 3)(*dims), pn.encode(), zip_status)
    want = LF.data_array_bytes(data, dims, "<f8" if prec == "f64" else "<f4")
    assert _unpack(ph + suffix, zip_status) == want
    assert _unpack(pn + suffix, zip_status) == want
    if prec == "f64":  # the reference build is FP64
        pr = str(tmp_path / "r.dat")
        rlib.writeDataArray(data.ctypes.data, ctypes.cast(grid, ctypes.POINTER(G)), pr.encode(), zip_status)
        assert _unpack(pr + suffix, zip_status) == want
        files = (pr, ph)
    else:
        files = (ph,)
    for path in files:
        got_dims = (ctypes.c_int * 3)()
        p = hl.readDataArray(got_dims, (path + suffix).encode())
        assert tuple(got_dims) == dims
        got = np.ctypeslib.as_array(p, shape=(int(np.prod(dims)),)).copy()
        libc.free(ctypes.cast(p, ctypes.c_void_p))
        assert np.array_equal(got.reshape(dims[::-1]), data)
    libc.free(ctypes.cast(grid, ctypes.c_void_p))


def test_bad_data_array_files(hip, tmp_path, capfd):
    """A missing file gives NULL and the reference's message; a header with a
    negative dimension or a file shorter than its header gives NULL and an
    error, in readDataArray and in lsmhip_read_data_array."""
    prec, hl, lib = hip
    dims = (ctypes.c_int * 3)()
    assert not hl.readDataArray(dims, str(tmp_path / "none.dat").encode())
    libc.fflush(None)  # the message goes through C stdio, as the reference's printf does
    assert "Could not open file" in capfd.readouterr().out
    assert lib.cdll.lsmhip_get_last_error() == 3
    lib.clear_error()
    size = 8 if prec == "f64" else 4
    for head, n in (((-2, -3, 1), 6), ((2, 3, 1), 5)):
        path = tmp_path / "bad.dat"
        path.write_bytes(np.array(head, "<i4").tobytes() + b"\0" * (size * n))
        assert not hl.readDataArray(dims, str(path).encode())
        assert lib.cdll.lsmhip_get_last_error() == 3
        lib.clear_error()
        buf = np.zeros(64, np.float64 if prec == "f64" else np.float32)
        with pytest.raises(Exception):
            data_io.read_data_array(path, prec)
        rd = lib.fn("lsmhip_read_data_array", [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_int),
                                               ctypes.c_char_p, ctypes.c_int])
        assert rd(buf.ctypes.data, 64, dims, str(path).encode(), 0) == -1
        assert lib.cdll.lsmhip_get_last_error() == 3
        lib.clear_error()


KINDS = ["linearExtrapolationBC", "signedLinearExtrapolationBC", "copyExtrapolationBC", "homogeneousNeumannBC"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("num_dims", [2, 3])
def test_dispatch_order_equals_the_reference(ref, kind, num_dims):
    """The oracle's face fills applied in oracle/lsm_bc_dispatch.py's order
    equal the reference's dispatch bit for bit, for every location index
    -1..10 (edge and corner ghosts depend on the order; 4, 5, 8 do nothing
    in 2D, as does an unknown index)."""
    rlib, G = ref
    R = ctypes.c_double
    dims = (7, 6, 5) if num_dims == 3 else (9, 8, 1)
    g = rlib.createGridSetGridDims(num_dims, (ctypes.c_int * 3)(*dims), _r3(R, (0, 0, 0)), _r3(R, (1, 1, 1)), 3)
    gc = g.contents
    gb6 = (gc.ilo_gb, gc.ihi_gb, gc.jlo_gb, gc.jhi_gb, gc.klo_gb, gc.khi_gb)
    fb6 = (gc.ilo_fb, gc.ihi_fb, gc.jlo_fb, gc.jhi_fb, gc.klo_fb, gc.khi_fb)
    shape = tuple(gc.grid_dims_ghostbox)[::-1]
    face = getattr(H.oracle(), (D.FACE_SYMBOL[kind] % num_dims).rstrip("_"))
    for idx in range(-1, 11):
        phi = np.random.default_rng(idx + 5).standard_normal(shape)
        want = phi.copy()
        for f in D.faces(num_dims, idx):
            if num_dims == 3:
                face(want, gb6, fb6, f)
            else:
                face(want, gb6[:4], fb6[:4], f)
        got = phi.copy()
        getattr(rlib, kind)(got.ctypes.data, g, idx)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (kind, num_dims, idx)
        if idx in (9,):
            assert not np.array_equal(got, phi)
    libc.free(ctypes.cast(g, ctypes.c_void_p))
