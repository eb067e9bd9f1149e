"""The reference's data-array and grid files (src/utils/lsm_data_arrays.c:364-437,
src/utils/lsm_grid.c:614-763, the Grid construction of lsm_grid.c) through the
HIP library's host entry points, against the oracle restatement
(oracle/lsm_files.py).  CPU; the device path is in test_gpu_data_files.py."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import lsm_files as LF  # noqa: E402

from lsmlib_amd import data_io  # noqa: E402

REAL = {"f64": ("<f8", "d", float), "f32": ("<f4", "f", LF.fp32)}


def test_oracle_grid_known_answers():
    """Values worked out by hand from lsm_grid.c: WENO5 keeps 4 ghost cells,
    marks 120 / 119 / 118 / 117 and fill mark 116, band widths 3dx / 6dx, and
    a fill box from D1 lo + 2 to D1 hi - 3; ENO2 keeps 3 ghost cells and 2dx /
    4dx; 2D grids carry a unit third dimension."""
    g = LF.grid_set_grid_dims(3, (10, 12, 14), (0.0, 0.0, 0.0), (1.0, 1.2, 1.4), 3)
    assert g["grid_dims_ghostbox"] == [18, 20, 22] and g["num_gridpts"] == 18 * 20 * 22
    assert (g["ilo_fb"], g["ihi_fb"], g["jlo_fb"], g["jhi_fb"], g["klo_fb"], g["khi_fb"]) == (3, 13, 3, 15, 3, 17)
    assert (g["ilo_D1_fb"], g["ihi_D1_fb"], g["ilo_D3_fb"], g["khi_D3_fb"]) == (1, 16, 3, 18)
    assert (g["mark_gb"], g["mark_D1"], g["mark_D3"], g["mark_fb"], g["num_nb_levels"]) == (120, 119, 117, 116, 4)
    assert abs(g["beta"] - 0.3) < 1e-15 and abs(g["gamma"] - 0.6) < 1e-15
    assert abs(g["x_lo_ghostbox"][0] + 0.4) < 1e-15 and abs(g["x_hi_ghostbox"][2] - 1.8) < 1e-15
    e = LF.grid_set_grid_dims(2, (20, 10), (0.0, 0.0), (2.0, 1.0), 1)
    assert e["grid_dims_ghostbox"] == [26, 16, 1] and e["grid_dims"][2] == 1 and e["dx"][2] == 0.0
    assert (e["ilo_fb"], e["ihi_fb"]) == (3, 22) and abs(e["gamma"] - 0.4) < 1e-15
    d = LF.grid_set_dx(3, 0.25, (0.0, 0.0, 0.0), (1.1, 2.0, 0.9), 0)
    assert d["grid_dims"] == [4, 8, 4] and abs(d["x_hi"][0] - 1.0) < 1e-15 and d["grid_dims_ghostbox"] == [8, 12, 8]
    # the grid record: 4 + 12 reals + 6 ints + 3 reals + 1 + 30 + 1 ints + 5 bytes + 2 reals
    assert len(LF.grid_bytes(g)) == 4 * (1 + 6 + 1 + 30 + 1) + 8 * (12 + 3 + 2) + 5


CASES = [(3, (10, 12, 14), (0.0, 0.0, 0.0), (1.0, 1.2, 1.4)), (3, (7, 5, 9), (-1.0, -0.5, 0.25), (1.0, 0.7, 2.0)),
         (2, (16, 24, 1), (0.0, -1.0, 0.0), (1.0, 2.0, 0.0))]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("accuracy", [0, 1, 2, 3])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_grid_construction_and_file(tmp_path, prec, accuracy, case):
    num_dims, dims, lo, hi = CASES[case]
    npt, rf, real = REAL[prec]
    want = LF.grid_set_grid_dims(num_dims, dims, lo, hi, accuracy, real)
    g = data_io.make_grid(num_dims, lo, hi, accuracy, grid_dims=dims, precision=prec)
    got = data_io.grid_dict(g)
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)
    path = tmp_path / "grid.bin"
    data_io.write_grid(g, path, prec)
    raw = path.read_bytes()
    assert raw == LF.grid_bytes(want, rf)
    assert data_io.grid_dict(data_io.read_grid(path, prec)) == got
    # a record written by the oracle reads back to the same grid
    path.write_bytes(LF.grid_bytes(want, rf))
    assert data_io.grid_dict(data_io.read_grid(path, prec)) == got


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("accuracy", [0, 3])
def test_grid_set_dx(prec, accuracy):
    npt, rf, real = REAL[prec]
    for lo, hi, dx in (((0.0, 0.0, 0.0), (1.1, 2.0, 0.9), 0.25), ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 2.0 / 37)):
        want = LF.grid_set_dx(3, dx, lo, hi, accuracy, real)
        got = data_io.grid_dict(data_io.make_grid(3, lo, hi, accuracy, dx=dx, precision=prec))
        for k, v in want.items():
            assert got[k] == v, (k, got[k], v)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_data_array_file_bytes(tmp_path, prec):
    npt, _, _ = REAL[prec]
    rng = np.random.default_rng(1)
    a = rng.standard_normal((9, 7, 11)).astype(npt)
    path = tmp_path / "phi.dat"
    data_io.write_data_array(a, path, prec)
    raw = path.read_bytes()
    assert raw == LF.data_array_bytes(a, (11, 7, 9), npt)
    dims, back = LF.read_data_array(raw, npt)
    assert dims == (11, 7, 9) and np.array_equal(back, a)
    # a file written by the oracle reads back bit for bit
    path.write_bytes(LF.data_array_bytes(a[::-1].copy(), (11, 7, 9), npt))
    assert np.array_equal(data_io.read_data_array(path, prec), a[::-1])


def test_data_array_refuses_short_files(tmp_path):
    a = np.arange(60, dtype=np.float64).reshape(3, 4, 5)
    path = tmp_path / "short.dat"
    path.write_bytes(LF.data_array_bytes(a, (5, 4, 3))[:-8])
    with pytest.raises(Exception):
        data_io.read_data_array(path)
