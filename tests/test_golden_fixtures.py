"""The oracle reproduces the full-size golden fixtures that
tests/golden/make_goldens.py wrote from the reference's own kernels (CPU)."""
import os
import sys

import numpy as np

import harness as H

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_goldens as MG  # noqa: E402


def test_c1_oracle_matches_reference_golden():
    import json
    g = json.load(open(os.path.join(os.path.dirname(MG.__file__), "c1_reinit_128.json")))
    grid = MG.Grid((g["n"],) * 3, lo=(-1.0,) * 3, hi=(1.0,) * 3, centred=True)
    phi, _ = MG.phi_perturbed_sphere(grid, g["seed"])
    phi = np.ascontiguousarray(phi)
    ch = MG.RefChain(grid, H.oracle())
    ch.fill_linear(phi)
    p1, p2 = np.zeros(grid.shape), np.zeros(grid.shape)
    for _ in range(g["steps"]):
        ch.rk3(phi, p1, p2, ch.rhs_reinit, ch.fill_linear, 0.5 * grid.dx[0])
    assert MG.digest(phi) == g["sha256_phi_f64"]
    flat = phi.reshape(-1)
    for i, hx in g["samples"]:
        assert flat[i] == float.fromhex(hx)
