"""TEST INFRASTRUCTURE ONLY -- never part of the product path.

A plain-Python restatement of the reference's binary files for data arrays
and grids, and of its Grid construction, used by tests/ to check the HIP
library's file entry points byte for byte.

  data array: src/utils/lsm_data_arrays.c:364-437 (writeDataArray,
              writeDataArrayNoGrid, readDataArray); the MATLAB reader
              spikes/matlab/examples/matlab/visualization/readDataArray.m:26-44
  grid file:  src/utils/lsm_grid.c:614-763 (writeGridToBinaryFile,
              readGridFromBinaryFile); readGridFromBinaryFile.m:25-33
  Grid:       src/utils/lsm_grid.c:47 (ghost cells per accuracy), 51-112
              (createGridSetDx), 189-232 (createGridSetGridDims), 789-1031
              (index-space limits, marks, band widths)
"""
from __future__ import annotations

import struct

import numpy as np

GHOST = (2, 3, 5, 4)  # ENO1, ENO2, ENO3, WENO5: src/utils/lsm_grid.c:47

BOX_FIELDS = [p + s for s in ("gb", "fb", "D1_fb", "D2_fb", "D3_fb")
              for p in ("ilo_", "ihi_", "jlo_", "jhi_", "klo_", "khi_")]


def data_array_bytes(data: np.ndarray, dims, real="<f8") -> bytes:
    """writeDataArrayNoGrid: three ints (ghost-box dimensions), then the reals
    in memory order (column-major: i fastest)."""
    return struct.pack("<3i", *dims) + np.ascontiguousarray(data, dtype=real).tobytes()


def read_data_array(raw: bytes, real="<f8"):
    """readDataArray / readDataArray.m: dims then nx*ny*nz reals."""
    nx, ny, nz = struct.unpack_from("<3i", raw, 0)
    n = nx * ny * nz
    data = np.frombuffer(raw, dtype=real, count=n, offset=12)
    return (nx, ny, nz), data.reshape(nz, ny, nx)


def _limits(g, accuracy, real=float):
    """setIndexSpaceLimits and its central / per-scheme helpers, lsm_grid.c:789-1031."""
    n = g["grid_dims_ghostbox"]
    for ax, (lo, hi) in enumerate((("ilo", "ihi"), ("jlo", "jhi"), ("klo", "khi"))):
        g[lo + "_gb"], g[hi + "_gb"] = 0, n[ax] - 1
        for m in (1, 2, 3):
            g["%s_D%d_fb" % (lo, m)], g["%s_D%d_fb" % (hi, m)] = m, n[ax] - 1 - m
    g["mark_gb"] = 120
    g["mark_D1"], g["mark_D2"], g["mark_D3"] = 119, 118, 117
    g["num_nb_levels"] = GHOST[accuracy]
    g["mark_fb"] = 120 - GHOST[accuracy]
    if accuracy > 1:
        g["beta"], g["gamma"] = real(3 * g["dx"][0]), real(6 * g["dx"][0])
    else:
        g["beta"], g["gamma"] = real(2 * g["dx"][0]), real(4 * g["dx"][0])
    # (difference box, minus start offset, plus end offset) per scheme, lsm_grid.c:832-985
    box, lo_off, hi_off = {0: ("D1", 0, 1), 1: ("D2", 1, 1), 2: ("D3", 1, 2), 3: ("D1", 2, 3)}[accuracy]
    for lo, hi in (("ilo", "ihi"), ("jlo", "jhi"), ("klo", "khi")):
        g[lo + "_fb"] = g["%s_%s_fb" % (lo, box)] + lo_off
        g[hi + "_fb"] = g["%s_%s_fb" % (hi, box)] - hi_off
    return g


def _dims(g, accuracy, real):
    ng = GHOST[accuracy]
    g["num_gridpts"] = 1
    g["grid_dims_ghostbox"] = [0, 0, 0]
    g["x_lo_ghostbox"] = [0.0] * 3
    g["x_hi_ghostbox"] = [0.0] * 3
    for i in range(g["num_dims"]):
        g["grid_dims_ghostbox"][i] = g["grid_dims"][i] + 2 * ng
        g["x_lo_ghostbox"][i] = real(g["x_lo"][i] - real(ng * g["dx"][i]))
        g["x_hi_ghostbox"][i] = real(g["x_hi"][i] + real(ng * g["dx"][i]))
        g["num_gridpts"] *= g["grid_dims_ghostbox"][i]
    if g["num_dims"] == 2:
        g["x_lo"][2] = g["x_hi"][2] = g["x_lo_ghostbox"][2] = g["x_hi_ghostbox"][2] = 0.0
        g["grid_dims"][2] = g["grid_dims_ghostbox"][2] = 1
        g["dx"][2] = 0.0
    return _limits(g, accuracy, real)


def grid_set_grid_dims(num_dims, grid_dims, x_lo, x_hi, accuracy, real=float):
    """createGridSetGridDims, lsm_grid.c:189-232 (real: the rounding of LSMLIB_REAL)."""
    g = {"num_dims": num_dims, "grid_dims": [0, 0, 0], "x_lo": [0.0] * 3, "x_hi": [0.0] * 3, "dx": [0.0] * 3}
    for i in range(num_dims):
        g["grid_dims"][i] = grid_dims[i]
        g["x_hi"][i] = real(x_hi[i])
        g["x_lo"][i] = real(x_lo[i])
        g["dx"][i] = real(real(g["x_hi"][i] - g["x_lo"][i]) / grid_dims[i])
    return _dims(g, accuracy, real)


def grid_set_dx(num_dims, dx, x_lo, x_hi, accuracy, real=float):
    """createGridSetDx, lsm_grid.c:51-112 (rint of (x_hi - x_lo)/dx)."""
    g = {"num_dims": num_dims, "grid_dims": [0, 0, 0], "x_lo": [0.0] * 3, "x_hi": [0.0] * 3, "dx": [0.0] * 3}
    for i in range(num_dims):
        g["dx"][i] = real(dx)
        diff = real(real(x_hi[i]) - real(x_lo[i]))
        g["grid_dims"][i] = int(np.rint(real(diff / real(dx))))
        g["x_lo"][i] = real(x_lo[i])
        g["x_hi"][i] = real(real(x_lo[i]) + real(g["grid_dims"][i] * real(dx)))
    return _dims(g, accuracy, real)


def grid_bytes(g, rf="d") -> bytes:
    """writeGridToBinaryFile, lsm_grid.c:614-678: field by field, no padding."""
    out = struct.pack("<i", g["num_dims"])
    for key in ("x_lo", "x_hi", "x_lo_ghostbox", "x_hi_ghostbox"):
        out += struct.pack("<3" + rf, *g[key])
    out += struct.pack("<3i", *g["grid_dims"]) + struct.pack("<3i", *g["grid_dims_ghostbox"])
    out += struct.pack("<3" + rf, *g["dx"]) + struct.pack("<i", g["num_gridpts"])
    out += struct.pack("<30i", *[g[k] for k in BOX_FIELDS])
    out += struct.pack("<i", g["num_nb_levels"])
    out += struct.pack("<5B", g["mark_gb"], g["mark_D1"], g["mark_D2"], g["mark_D3"], g["mark_fb"])
    out += struct.pack("<2" + rf, g["beta"], g["gamma"])
    return out


def read_grid(raw: bytes, rf="d"):
    """readGridFromBinaryFile, lsm_grid.c:680-763 (and readGridFromBinaryFile.m)."""
    rs = struct.calcsize("<" + rf)
    g, o = {}, 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from("<" + fmt, raw, o)
        o += struct.calcsize("<" + fmt)
        return list(v)

    g["num_dims"] = take("i")[0]
    for key in ("x_lo", "x_hi", "x_lo_ghostbox", "x_hi_ghostbox"):
        g[key] = take("3" + rf)
    g["grid_dims"] = take("3i")
    g["grid_dims_ghostbox"] = take("3i")
    g["dx"] = take("3" + rf)
    g["num_gridpts"] = take("i")[0]
    for k, v in zip(BOX_FIELDS, take("30i")):
        g[k] = v
    g["num_nb_levels"] = take("i")[0]
    g["mark_gb"], g["mark_D1"], g["mark_D2"], g["mark_D3"], g["mark_fb"] = take("5B")
    g["beta"], g["gamma"] = take("2" + rf)
    assert o == len(raw) and rs in (4, 8)
    return g


def fp32(x):
    return float(np.float32(x))


__all__ = ["GHOST", "BOX_FIELDS", "data_array_bytes", "read_data_array", "grid_set_grid_dims", "grid_set_dx",
           "grid_bytes", "read_grid", "fp32"]
