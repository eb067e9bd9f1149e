"""Data-array files written from and read into device memory (the reference's
format, src/utils/lsm_data_arrays.c:364-437): same bytes as the oracle's
writer, and a field survives the round trip bit for bit."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import lsm_files as LF  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_data_array_round_trip(tmp_path, prec):
    import torch
    from lsmlib_amd import data_io
    npt = "<f8" if prec == "f64" else "<f4"
    dt = torch.float64 if prec == "f64" else torch.float32
    # larger than one staging chunk (2^24 reals) to cover the chunked copy
    shape = (70, 260, 1000)
    a = torch.randn(shape, dtype=dt, device="cuda")
    path = tmp_path / "phi.dat"
    data_io.write_data_array(a, path, prec)
    host = a.cpu().numpy()
    raw = path.read_bytes()
    assert raw == LF.data_array_bytes(host, (shape[2], shape[1], shape[0]), npt)
    b = data_io.read_data_array(path, prec, device="cuda")
    assert b.is_cuda and torch.equal(a, b)
