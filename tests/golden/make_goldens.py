"""Writes the golden fixtures of the full-size workloads from the reference's
own FP64 kernels (oracle/_ref/liblsm_ref.so, built by `make -C oracle ref`
from the reference sources).  Functions the reference lacks (the periodic
fill) come from the oracle's restatement.

    python tests/golden/make_goldens.py c1          # reinitialization, 128^3, 50 steps

Each fixture records the configuration and the SHA-256 of the final field
(FP64, ghost box, column-major bytes) with a few sampled values, so a GPU run
of the same sequence can be checked bit for bit without the reference.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import harness as H  # noqa: E402
from lsmlib_amd.solver import Grid, phi_perturbed_sphere  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def samples(a: np.ndarray, count=8, seed=0):
    rng = np.random.default_rng(seed)
    flat = a.reshape(-1)
    idx = sorted(int(v) for v in rng.choice(flat.size, count, replace=False))
    return [[i, float(flat[i]).hex()] for i in idx]


class RefChain:
    """The drop-in call sequence of one TVD-RK3 step on numpy arrays through
    the reference build."""

    def __init__(self, grid: Grid, impl):
        self.g, self.f = grid, impl
        z = lambda: np.zeros(grid.shape)  # noqa: E731
        self.G = [z() for _ in range(6)]
        self.rhs, self.d1 = z(), z()

    def fill_linear(self, phi):
        for idx in range(6):
            self.f.lsm3dlinearextrapolation(phi, self.g.gb, self.g.fb, idx)

    def rhs_reinit(self, phi):
        g, G, f = self.g, self.G, self.f
        gb, fb = g.gb, g.fb
        f.lsm3dhjweno5(G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, phi, gb, self.d1, gb, fb, *g.dx)
        f.lsm3dcomputereinitializationeqnrhs(self.rhs, gb, phi, gb, phi, gb, G[0], G[1], G[2], gb, G[3], G[4],
                                             G[5], gb, fb, *g.dx, 0)

    def rk3(self, phi, p1, p2, rhs_fn, fill, dt):
        gb, fb, f = self.g.gb, self.g.fb, self.f
        rhs_fn(phi)
        f.lsm3dtvdrk3stage1(p1, gb, phi, gb, self.rhs, gb, fb, dt)
        fill(p1)
        rhs_fn(p1)
        f.lsm3dtvdrk3stage2(p2, gb, p1, gb, phi, gb, self.rhs, gb, fb, dt)
        fill(p2)
        rhs_fn(p2)
        f.lsm3dtvdrk3stage3(phi, gb, p2, gb, phi, gb, self.rhs, gb, fb, dt)
        fill(phi)


def c1(n=128, seed=1, steps=50):
    ref = H.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/liblsm_ref.so is not built (make -C oracle ref)")
    grid = Grid((n, n, n), lo=(-1.0, -1.0, -1.0), hi=(1.0, 1.0, 1.0), centred=True)
    phi, centre = phi_perturbed_sphere(grid, seed)
    phi = np.ascontiguousarray(phi)
    ch = RefChain(grid, ref)
    ch.fill_linear(phi)
    p1, p2 = np.zeros(grid.shape), np.zeros(grid.shape)
    dt = 0.5 * grid.dx[0]
    t0 = time.time()
    for _ in range(steps):
        ch.rk3(phi, p1, p2, ch.rhs_reinit, ch.fill_linear, dt)
    secs = time.time() - t0
    return {
        "workload": "reinitialization, %d^3 cell centres on [-1,1]^3, phi0 = (|x-c|-0.5)(1+0.1 sin3pi x sin3pi y "
                    "sin3pi z), c ~ U[-0.1,0.1]^3 (splitmix64 seed %d), %d pseudo-steps of HJ-WENO5 -> "
                    "reinitialization RHS (sign from phi) -> TVD-RK3, dt = 0.5 dx, linear extrapolation after "
                    "every stage (faces x_lo, x_hi, y_lo, y_hi, z_lo, z_hi)" % (n, seed, steps),
        "n": n, "seed": seed, "steps": steps, "centre": centre,
        "kernels": "reference (oracle/_ref, flang -fdefault-real-8)",
        "sha256_phi_f64": digest(phi), "samples": samples(phi),
        "interior_abs_max": float(np.abs(grid.interior(phi)).max()),
        "cpu_seconds_1thread": round(secs, 2),
    }


CASES = {"c1": ("c1_reinit_128.json", c1)}

if __name__ == "__main__":
    for name in sys.argv[1:] or list(CASES):
        fname, fn = CASES[name]
        rec = fn()
        with open(os.path.join(HERE, fname), "w") as fh:
            json.dump(rec, fh, indent=1)
        print(fname, rec["sha256_phi_f64"], rec.get("cpu_seconds_1thread"))
