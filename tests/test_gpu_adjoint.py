"""Reverse mode of the normal-motion stage and the shape-matching problem.

The stage adjoint is checked against torch autograd through an independent
FP64 restatement of the stage on periodic arrays (CPU); the whole problem
against central finite differences along 16 random directions (FP64, 64^3)
and the FP32 adjoint against the FP64 one (256^3)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch_stage(us, uc, V, dt, dx, stage):
    """One RK stage of phi_t + V|grad phi| = 0 on periodic (k, j, i) arrays:
    HJ-WENO5 (src/toolbox/lsm_spatial_derivatives3d.f.in:1286-1334), Godunov
    (src/toolbox/lsm_level_set_evolution3d.f.in:210-224), TVD-RK3/RK2 weights."""
    import torch
    a, b = {1: (0.0, 1.0), 2: (0.75, 0.25), 3: (1.0 / 3.0, 2.0 / 3.0), 22: (0.5, 0.5)}[stage]

    def weno(v):
        m = torch.stack([x * x for x in v]).amax(0)
        eps = 1e-6 * m + 1e-36
        v1, v2, v3, v4, v5 = v
        p1 = v1 / 3 - 7 * v2 / 6 + 11 * v3 / 6
        p2 = -v2 / 6 + 5 * v3 / 6 + v4 / 3
        p3 = v3 / 3 + 5 * v4 / 6 - v5 / 6
        S1 = 13 / 12 * (v1 - 2 * v2 + v3) ** 2 + 0.25 * (v1 - 4 * v2 + 3 * v3) ** 2
        S2 = 13 / 12 * (v2 - 2 * v3 + v4) ** 2 + 0.25 * (v2 - v4) ** 2
        S3 = 13 / 12 * (v3 - 2 * v4 + v5) ** 2 + 0.25 * (3 * v3 - 4 * v4 + v5) ** 2
        a1, a2, a3 = 0.1 / (S1 + eps) ** 2, 0.6 / (S2 + eps) ** 2, 0.3 / (S3 + eps) ** 2
        return (a1 * p1 + a2 * p2 + a3 * p3) / (a1 + a2 + a3)

    G = 0
    vpos = V > 0
    for dim, h in ((2, dx[0]), (1, dx[1]), (0, dx[2])):
        f = {s: torch.roll(us, shifts=-s, dims=dim) for s in range(-3, 4)}
        D = {t: f[t] - f[t - 1] for t in range(-2, 4)}
        pp = weno([D[3] / h, D[2] / h, D[1] / h, D[0] / h, D[-1] / h])
        pm = weno([D[-2] / h, D[-1] / h, D[0] / h, D[1] / h, D[2] / h])
        A = torch.where(vpos, pm.clamp(min=0), pm.clamp(max=0)) ** 2
        B = torch.where(vpos, pp.clamp(max=0), pp.clamp(min=0)) ** 2
        G = G + torch.where(A > B, A, B)
    L = -V * torch.sqrt(G)
    return a * uc + b * (us + dt * L)


@pytest.mark.parametrize("stage", [1, 2, 3, 22])
def test_stage_adjoint_matches_torch_autograd(stage):
    import torch
    from lsmlib_amd.adjoint import NormalMotionAdjoint
    from lsmlib_amd.solver import Grid
    n = 14
    g = Grid((n, n, n))
    eng = NormalMotionAdjoint(g, "f64", steps=1)
    s = eng.solver
    rng = np.random.default_rng(stage)
    x, y, z = (g.interior(a) for a in g.mesh())
    us = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.4 * np.sin(2 * np.pi * z) + 0.05 * rng.standard_normal(x.shape)
    uc = rng.standard_normal(x.shape)
    V = np.where(rng.random(x.shape) < 0.5, -1.0, 1.0) * (0.5 + rng.random(x.shape))
    lam_out = rng.standard_normal(x.shape)
    dt = 0.2 * g.dx[0]

    # torch autograd on the CPU
    tus, tuc, tV = (torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in (us, uc, V))
    out = _torch_stage(tus, tuc, tV, dt, g.dx, stage)
    gus, guc, gV = torch.autograd.grad((out * torch.tensor(lam_out)).sum(), (tus, tuc, tV))

    # the HIP adjoint on the ghost-box layout
    def ghost(a):
        full = np.zeros(g.shape)
        g.interior(full)[...] = a
        return s.field(full)

    d_us, d_uc, d_V, d_lam = ghost(us), ghost(uc), ghost(V), ghost(lam_out)
    s.fill_ghosts(d_us)
    lam_us, lam_uc, lam_v = s.field(), s.field(), s.field()
    eng._stage_adjoint(lam_us, lam_uc, lam_v, d_lam, d_us, stage, d_V, dt)
    eng._fold_ghosts(lam_us)
    s.lib.check()

    # the forward stage agrees with the restatement
    o = s.field()
    s.stage_normal(o, d_us, d_uc, stage, dt, vn=0.0, vn_field=d_V)
    assert np.abs(g.interior(s.host(o)) - out.detach().numpy()).max() < 1e-12

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))

    assert rel(g.interior(s.host(lam_us)), gus.numpy()) < 1e-10
    assert rel(g.interior(s.host(lam_v)), gV.numpy()) < 1e-10
    if stage != 1:
        assert rel(g.interior(s.host(lam_uc)), guc.numpy()) < 1e-12


def test_two_forward_runs_on_one_engine_keep_their_own_gradients():
    """Two Evolve nodes on one engine before a single backward (a summed loss)
    give, for each speed and phi0, the gradients of separate runs: each node
    keeps its own checkpoints and speed.  A run without checkpoints in between
    (ShapeMatching.value) does not disturb them."""
    import torch
    from lsmlib_amd.adjoint import ShapeMatching, autograd_functions
    P = ShapeMatching(n=32, precision="f64", steps=12, checkpoint_every=4)
    Evolve, Mismatch = autograd_functions()
    a1 = P.a.detach().clone()
    a2 = (P.a * 1.7 + 0.05).detach().clone()
    want = []
    for a in (a1, a2):
        ai = a.clone().requires_grad_(True)
        pi = P.phi0.detach().clone().requires_grad_(True)
        L = Mismatch.apply(Evolve.apply(pi, P.speed(ai), P.engine), P.target, P.engine.solver, P.eps)
        want.append(torch.autograd.grad(L, (ai, pi)))
    b1, b2 = a1.clone().requires_grad_(True), a2.clone().requires_grad_(True)
    p1, p2 = (P.phi0.detach().clone().requires_grad_(True) for _ in range(2))
    L1 = Mismatch.apply(Evolve.apply(p1, P.speed(b1), P.engine), P.target, P.engine.solver, P.eps)
    L2 = Mismatch.apply(Evolve.apply(p2, P.speed(b2), P.engine), P.target, P.engine.solver, P.eps)
    P.value(a1, P.phi0)  # a forward run without checkpoints between the nodes and the backward
    got = torch.autograd.grad(L1 + L2, (b1, p1, b2, p2))
    for (wa, wp), ga, gp in ((want[0], got[0], got[1]), (want[1], got[2], got[3])):
        assert torch.equal(ga, wa) and torch.equal(gp, wp)
    assert not torch.equal(got[0], got[2])


def _smooth_direction(grid, rng, modes=3):
    """A smooth random field on the ghost box (a few random low Fourier modes)."""
    x, y, z = grid.mesh()
    f = np.zeros(grid.shape)
    for _ in range(modes):
        k = rng.integers(1, 4, size=3)
        f += rng.standard_normal() * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + 2 * np.pi * rng.random())
    return f


def test_shape_matching_gradient_matches_central_differences():
    """64^3 FP64, 50 RK3 steps (the 256^3 x 200-step problem's physical time),
    checkpoints every 10: the adjoint directional derivative along 8 random
    directions in the speed parameters and 8 in phi0 matches central finite
    differences to 1e-3 relative."""
    import torch
    from lsmlib_amd.adjoint import ShapeMatching
    P = ShapeMatching(n=64, precision="f64", steps=50, checkpoint_every=10)
    L0, ga, gphi = P.value_and_grads()
    assert L0 > 0
    rng = np.random.default_rng(11)
    worst = 0.0
    for d in range(16):
        if d < 8:
            da = torch.tensor(rng.standard_normal(P.a.numel()), dtype=torch.float64, device=P.a.device)
            dphi = torch.zeros_like(P.phi0)
        else:
            da = torch.zeros_like(P.a)
            dphi = torch.tensor(_smooth_direction(P.grid, rng), dtype=torch.float64, device=P.phi0.device)
        ad = float((ga * da).sum() + (gphi * dphi).sum())
        fds = {}
        for h in (1e-5, 1e-6):
            fds[h] = (P.value(P.a + h * da, P.phi0 + h * dphi) - P.value(P.a - h * da, P.phi0 - h * dphi)) / (2 * h)
        err = abs(fds[1e-6] - ad) / max(abs(ad), 1e-300)
        print("direction %2d: adjoint %+.10e  central difference h=1e-5 %+.10e h=1e-6 %+.10e  rel %.2e"
              % (d, ad, fds[1e-5], fds[1e-6], err))
        worst = max(worst, err)
    assert worst < 1e-3


def test_shape_matching_fp32_gradient_close_to_fp64():
    """256^3, 200 RK3 steps, checkpoints every 20: the FP32 adjoint against the
    FP64 one.  Speed-parameter gradient: relative L2 < 1e-3.  phi0 gradient:
    its directional derivatives along 8 smooth random directions (what the
    finite-difference check tests) within 1e-2; the pointwise field is printed
    only, as adjoint mass gathers in thin layers where FP32 and FP64 may take
    different upwind branches."""
    from lsmlib_amd.adjoint import ShapeMatching
    from lsmlib_amd.solver import Grid
    grid = Grid((256,) * 3)
    rng = np.random.default_rng(7)
    dirs = [grid.interior(_smooth_direction(grid, rng)) for _ in range(8)]
    res = {}
    for prec in ("f64", "f32"):
        P = ShapeMatching(n=256, precision=prec, steps=200, checkpoint_every=20)
        L, ga, gphi = P.value_and_grads()
        gp = grid.interior(gphi.cpu().numpy())
        res[prec] = (L, ga.cpu().numpy(), [float((gp * d).sum()) for d in dirs], gp)
        del P
    (L64, a64, d64, p64), (L32, a32, d32, p32) = res["f64"], res["f32"]
    ra = np.linalg.norm(a32 - a64) / np.linalg.norm(a64)
    rd = max(abs(x - y) / abs(y) for x, y in zip(d32, d64))
    rp = np.linalg.norm(p32 - p64) / np.linalg.norm(p64)
    print("loss f64 %.10e f32 %.10e; speed-parameter gradient rel L2 %.2e; phi0 directional derivatives max rel "
          "%.2e (pointwise rel L2 %.2e)" % (L64, L32, ra, rd, rp))
    for x, y in zip(d32, d64):
        print("  phi0 direction: f64 %+.8e f32 %+.8e" % (y, x))
    assert abs(L32 - L64) / L64 < 1e-4
    assert ra < 1e-3
    assert rd < 1e-2
