"""Fused WENO5 + RHS + RK stage kernels against the drop-in chain (GPU) and
the oracle chain (CPU).  FP64: bitwise.  FP32: error bounds against FP64."""
import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solvers():
    from lsmlib_amd.solver import Grid, Solver
    g = Grid((22, 19, 17), lo=(0, 0, 0), hi=(1.1, 0.95, 0.85))
    return g, Solver(g, "f64", bc="signed_linear"), Solver(g, "f32", bc="signed_linear")


def _phi(g, seed=3):
    from lsmlib_amd.solver import phi_spheres_min
    phi, _ = phi_spheres_min(g, seed, 4, (0.2, 0.8), (0.1, 0.25))
    rng = np.random.default_rng(seed)
    return phi + 0.02 * rng.standard_normal(phi.shape)  # rough enough to switch WENO weights


def _oracle_stage(g, phi_np, phin_np, stage, dt, kind, vel=None, vn=0.8, phi0=None):
    """Oracle drop-in chain for one stage on host arrays."""
    o = H.oracle()
    gb, fb = g.gb, g.fb
    G = [np.zeros(g.shape) for _ in range(6)]
    d1, rhs, out = np.zeros(g.shape), np.zeros(g.shape), np.zeros(g.shape)
    if kind == "advect":
        o.lsm3dupwindhjweno5(G[0], G[1], G[2], gb, phi_np, gb, vel[0], vel[1], vel[2], gb, d1, gb, fb, *g.dx)
        o.lsm3dzerooutlevelseteqnrhs(rhs, gb)
        o.lsm3daddadvectiontermtolserhs(rhs, gb, G[0], G[1], G[2], gb, vel[0], vel[1], vel[2], gb, fb)
    else:
        o.lsm3dhjweno5(G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, phi_np, gb, d1, gb, fb, *g.dx)
        if kind == "normal":
            o.lsm3dzerooutlevelseteqnrhs(rhs, gb)
            o.lsm3daddconstnormalveltermtolserhs(rhs, gb, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, vn, fb)
        else:
            o.lsm3dcomputereinitializationeqnrhs(rhs, gb, phi_np, gb, phi0 if phi0 is not None else phi_np, gb,
                                                 G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, fb, *g.dx,
                                                 1 if phi0 is not None else 0)
    out[...] = 0.0
    if stage == 1:
        o.lsm3dtvdrk3stage1(out, gb, phi_np, gb, rhs, gb, fb, dt)
    elif stage == 2:
        o.lsm3dtvdrk3stage2(out, gb, phi_np, gb, phin_np, gb, rhs, gb, fb, dt)
    else:
        o.lsm3dtvdrk3stage3(out, gb, phi_np, gb, phin_np, gb, rhs, gb, fb, dt)
    return out


@pytest.mark.parametrize("kind", ["normal", "reinit", "reinit_phi0", "advect"])
@pytest.mark.parametrize("stage", [1, 2, 3])
def test_fused_stage_fp64_bitwise_vs_oracle(solvers, kind, stage):
    g, s64, _ = solvers
    phi_np = _phi(g)
    phin_np = _phi(g, 5)
    dt = 0.3 * g.dx[0]
    us, uc = s64.field(phi_np), s64.field(phin_np)
    out = s64.field(np.zeros(g.shape))
    vel = phi0 = None
    if kind == "normal":
        s64.stage_normal(out, us, uc, stage, dt, vn=-0.8)
        ref = _oracle_stage(g, phi_np, phin_np, stage, dt, "normal", vn=-0.8)
    elif kind.startswith("reinit"):
        phi0 = s64.field(phin_np) if kind == "reinit_phi0" else None
        s64.stage_reinit(out, us, uc, stage, dt, phi0=phi0)
        ref = _oracle_stage(g, phi_np, phin_np, stage, dt, "reinit", phi0=phin_np if phi0 is not None else None)
    else:
        c = 0.7
        vel = s64.deform_velocity(c)  # the same device function the fused kernel evaluates
        s64.stage_deform(out, us, uc, stage, dt, c)
        ref = _oracle_stage(g, phi_np, phin_np, stage, dt, "advect", vel=[s64.host(v) for v in vel])
    got = s64.host(out)
    gi, ri = g.interior(got), g.interior(ref)
    bad = np.argwhere(gi != ri)
    assert len(bad) == 0, "differs at %d cells, first %s: %r vs %r" % (
        len(bad), bad[0].tolist(), gi[tuple(bad[0])], ri[tuple(bad[0])])


@pytest.mark.parametrize("kind", ["normal", "advect", "reinit"])
def test_fused_rk3_step_fp64_equals_dropin_chain(solvers, kind):
    from lsmlib_amd.solver import DropInChain
    g, s64, _ = solvers
    phi_np = _phi(g)
    a, b = s64.field(phi_np), s64.field(phi_np)
    s64.fill_ghosts(a)
    s64.fill_ghosts(b)
    wa = (s64.field(), s64.field())
    wb = (s64.field(), s64.field())
    chain = DropInChain(s64)
    dt = 0.4 * g.dx[0]
    for step in range(3):
        if kind == "normal":
            s64.rk3_fused(a, wa, s64.stage_normal, dt, vn=1.0)
            chain.rk3(b, wb, chain.rhs_normal, dt, vn=1.0)
        elif kind == "advect":
            c = np.cos(np.pi * step * dt / 3.0)
            vel = s64.deform_velocity(c)
            s64.rk3_fused(a, wa, s64.stage_deform, dt, c=c)
            chain.rk3(b, wb, chain.rhs_advect, dt, vel=vel)
        else:
            s64.rk3_fused(a, wa, s64.stage_reinit, dt)
            chain.rk3(b, wb, chain.rhs_reinit, dt)
    s64.lib.check()
    ha, hb = s64.host(a), s64.host(b)
    assert np.array_equal(ha, hb), np.abs(ha - hb).max()


@pytest.mark.parametrize("kernels", [1, 2])
@pytest.mark.parametrize("kind", ["normal", "advect", "reinit"])
def test_fused_fp32_close_to_fp64(solvers, kind, kernels):
    """Both FP32 kernel sets (1: a WENO5 window per cell, the default; 2: each
    position's terms shared between cells) within 2e-5 of FP64 after five
    steps."""
    import ctypes
    g, s64, s32 = solvers
    from lsmlib_amd.solver import phi_sphere
    phi_np = phi_sphere(g, (0.5, 0.45, 0.42), 0.3)
    if kind == "reinit":
        phi_np = 1.7 * phi_np  # |grad phi| = 1.7: the reinitialization RHS is O(1)
    select = s32.lib.fn("lsmhip_set_fused_f32_kernels", [ctypes.c_int])
    prev = select(kernels)
    res = []
    for s in (s64, s32):
        a = s.field(phi_np)
        s.fill_ghosts(a)
        w = (s.field(), s.field())
        for step in range(5):
            if kind == "normal":
                s.rk3_fused(a, w, s.stage_normal, 0.4 * g.dx[0], vn=1.0)
            elif kind == "reinit":
                s.rk3_fused(a, w, s.stage_reinit, 0.4 * g.dx[0])
            else:
                s.rk3_fused(a, w, s.stage_deform, 0.4 * g.dx[0], c=1.0)
        s.lib.check()
        res.append(g.interior(s.host(a)))
    assert select(prev) == kernels
    # FP32 storage and arithmetic: relative-to-O(1) errors of a few ulp per step
    assert np.abs(res[0] - res[1]).max() < 2e-5
