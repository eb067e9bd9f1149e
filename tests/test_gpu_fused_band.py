"""The fused narrow-band entry points (lsmhip_weno5_local_stage_normal,
lsmhip_weno5_local_normal_dt) with a normal-speed field, against the drop-in
local chain through the reference-named entry points, FP64 bitwise: the CFL
step and TVD-RK3 stages 1 and 2, with and without the folded copies and the
cut-off."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def band():
    import torch
    from lsmlib_amd.solver import NarrowBandRun
    r = NarrowBandRun(n=64, steps=1, rebuild_every=5, fused=True)
    r.rebuild()
    g = r.grid
    ax = [torch.as_tensor(g.axis(d), dtype=torch.float64, device="cuda") for d in range(3)]
    z, y, x = torch.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    vn = torch.sin(6.0 * x) * torch.cos(5.0 * y) + 0.3 * z - 0.2  # both signs
    vn[(vn.abs() < 0.05)] = 1e-13                                    # below zero_tol
    return r, vn.contiguous()


def _dropin_stage(r, vn, st, src, out, dt, cutoff, copy):
    """The chain for one stage with a speed field; returns dt for stage 0."""
    a, g, G = r.api, r.grid, r.grads
    gb = g.gb
    lo0, hi0, hi3 = int(r.n_lo[0]), int(r.n_hi[0]), int(r.n_hi[r.level])
    band = (r.ix, r.iy, r.iz)
    mfb = r.MARKS[0]
    if copy and out is not None:
        a.lsm3dcopydatalocal(out, gb, r.phi, gb, *band, lo0, hi3)
    r._weno(src)
    if st == 0:
        return a.lsm3dcomputestablenormalveldtlocal(None, vn, gb, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb,
                                                    *g.dx, r.CFL, *band, lo0, hi0, r.nb, gb, mfb)
    a.lsm3dzerooutlevelseteqnrhslocal(r.rhs, gb, *band, lo0, hi0)
    a.lsm3daddnormalveltermtolserhslocal(r.rhs, gb, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, vn, gb, *band,
                                         lo0, hi0, r.nb, gb, mfb)
    if cutoff:
        a.lsm3dmultiplycutofflserhslocal(src, r.rhs, gb, *band, lo0, hi0, r.nb, gb, mfb, r.width_inner, r.width)
    if st == 1:
        a.lsm3dtvdrk3stage1local(out, gb, r.phi, gb, r.rhs, gb, dt, *band, lo0, hi0, r.nb, gb, mfb)
    else:
        a.lsm3dtvdrk3stage2local(out, gb, src, gb, r.phi, gb, r.rhs, gb, dt, *band, lo0, hi0, r.nb, gb, mfb)
    r.solver.lib.check()


def _fused_stage(r, vn, st, src, out, dt, cutoff, copy):
    s, g = r.solver, r.grid
    gbp = s._gb()
    lo0, hi0, hi3 = int(r.n_lo[0]), int(r.n_hi[0]), int(r.n_hi[r.level])
    ptrs = [b.data_ptr() for b in (r.ix, r.iy, r.iz)]
    if st == 0:
        dt = s.real(0)
        s._ok(r._fdt(ctypes.byref(dt), src.data_ptr(), *gbp, 0.0, vn.data_ptr(), *g.dx, r.CFL, *ptrs, lo0, hi0,
                     r.nb.data_ptr(), *gbp, r.MARKS[0], r.MARKS[2]))
        return dt.value
    s._ok(r._fstage(out.data_ptr(), src.data_ptr(), r.phi.data_ptr(), *gbp, st, 0.0, vn.data_ptr(), dt, *g.dx,
                    *ptrs, lo0, hi0, hi3 if copy else -1, r.nb.data_ptr(), *gbp, r.MARKS[0], r.MARKS[2],
                    r.width_inner if cutoff else 0.0, r.width if cutoff else 0.0))
    s.lib.check()


@pytest.mark.parametrize("cutoff", [True, False])
@pytest.mark.parametrize("copy", [True, False])
def test_fused_band_with_speed_field_is_bitwise_the_chain(band, cutoff, copy):
    r, vn = band
    s = r.solver
    dt_ref = _dropin_stage(r, vn, 0, r.phi, None, None, cutoff, copy)
    dt = _fused_stage(r, vn, 0, r.phi, None, None, cutoff, copy)
    assert dt == dt_ref
    p1_ref, p2_ref = s.field(), s.field()
    p1, p2 = s.field(), s.field()
    # distinct stale contents outside the stage's writes, identical in both paths
    for a in (p1_ref, p1):
        a.copy_(r.phi * 0.5 + 1.0)
    for a in (p2_ref, p2):
        a.copy_(r.phi * -0.25 + 2.0)
    _dropin_stage(r, vn, 1, r.phi, p1_ref, dt, cutoff, copy)
    _fused_stage(r, vn, 1, r.phi, p1, dt, cutoff, copy)
    assert np.array_equal(s.host(p1), s.host(p1_ref))
    s.fill_ghosts(p1_ref)
    s.fill_ghosts(p1)
    _dropin_stage(r, vn, 2, p1_ref, p2_ref, dt, cutoff, copy)
    _fused_stage(r, vn, 2, p1, p2, dt, cutoff, copy)
    assert np.array_equal(s.host(p2), s.host(p2_ref))


def _dt_rhs(r, vn, src, cutoff):
    """lsmhip_weno5_local_normal_dt_rhs on src: (dt, rhs per level-0 entry)."""
    s, g = r.solver, r.grid
    gbp = s._gb()
    lo0, hi0 = int(r.n_lo[0]), int(r.n_hi[0])
    ptrs = [b.data_ptr() for b in (r.ix, r.iy, r.iz)]
    rhs = s.torch.full((hi0 - lo0 + 1,), 3.25, dtype=s.dtype, device=s.device)
    dt = s.real(0)
    s._ok(r._fdtr(ctypes.byref(dt), rhs.data_ptr(), src.data_ptr(), *gbp, 0.0,
                  None if vn is None else vn.data_ptr(), *g.dx, r.CFL, *ptrs, lo0, hi0, r.nb.data_ptr(), *gbp,
                  r.MARKS[0], r.MARKS[2], r.width_inner if cutoff else 0.0, r.width if cutoff else 0.0))
    return dt.value, rhs


def _stage_from_rhs(r, st, src, out, rhs, dt, copy):
    s = r.solver
    gbp = s._gb()
    lo0, hi0, hi3 = int(r.n_lo[0]), int(r.n_hi[0]), int(r.n_hi[r.level])
    ptrs = [b.data_ptr() for b in (r.ix, r.iy, r.iz)]
    s._ok(r._fsrhs(out.data_ptr(), src.data_ptr(), r.phi.data_ptr(), *gbp, st, rhs.data_ptr(), dt, *ptrs, lo0, hi0,
                   hi3 if copy else -1, r.nb.data_ptr(), *gbp, r.MARKS[0]))
    s.lib.check()


@pytest.mark.parametrize("cutoff", [True, False])
@pytest.mark.parametrize("copy", [True, False])
def test_stage_from_stored_rhs_is_bitwise_the_chain(band, cutoff, copy):
    """The CFL step that also stores the first stage's RHS, then the stage
    from that RHS: the same dt and the same stage-1 field as the drop-in chain
    (and so as the fused stage), FP64 bitwise, with a speed field; stage 2
    from the RHS stored for phi^1 equals the fused stage 2."""
    r, vn = band
    s = r.solver
    dt_ref = _dropin_stage(r, vn, 0, r.phi, None, None, cutoff, copy)
    dt, rhs = _dt_rhs(r, vn, r.phi, cutoff)
    assert dt == dt_ref
    p1_ref, p1 = s.field(), s.field()
    for a in (p1_ref, p1):
        a.copy_(r.phi * 0.5 + 1.0)
    _dropin_stage(r, vn, 1, r.phi, p1_ref, dt, cutoff, copy)
    _stage_from_rhs(r, 1, r.phi, p1, rhs, dt, copy)
    assert np.array_equal(s.host(p1), s.host(p1_ref))
    s.fill_ghosts(p1)
    p2_ref, p2 = s.field(), s.field()
    for a in (p2_ref, p2):
        a.copy_(r.phi * -0.25 + 2.0)
    _fused_stage(r, vn, 2, p1, p2_ref, dt, cutoff, copy)
    _, rhs1 = _dt_rhs(r, vn, p1, cutoff)
    _stage_from_rhs(r, 2, p1, p2, rhs1, dt, copy)
    assert np.array_equal(s.host(p2), s.host(p2_ref))


def test_stored_rhs_of_constant_speed_matches_the_fused_stage():
    """Constant normal speed (no speed field), the cut-off on: dt and stage 1
    from the stored RHS equal lsmhip_weno5_local_normal_dt and the fused
    stage 1 bit for bit, in FP64 and in FP32."""
    import torch
    from lsmlib_amd.solver import NarrowBandRun
    for prec in ("f64", "f32"):
        r = NarrowBandRun(n=48, steps=1, rebuild_every=5, fused=True, precision=prec)
        r.rebuild()
        s, g = r.solver, r.grid
        gbp = s._gb()
        lo0, hi0, hi3 = int(r.n_lo[0]), int(r.n_hi[0]), int(r.n_hi[r.level])
        ptrs = [b.data_ptr() for b in (r.ix, r.iy, r.iz)]
        dt0 = s.real(0)
        s._ok(r._fdt(ctypes.byref(dt0), r.phi.data_ptr(), *gbp, 1.0, None, *g.dx, r.CFL, *ptrs, lo0, hi0,
                     r.nb.data_ptr(), *gbp, r.MARKS[0], r.MARKS[2]))
        rhs = torch.zeros(hi0 - lo0 + 1, dtype=s.dtype, device=s.device)
        dt1 = s.real(0)
        s._ok(r._fdtr(ctypes.byref(dt1), rhs.data_ptr(), r.phi.data_ptr(), *gbp, 1.0, None, *g.dx, r.CFL, *ptrs,
                      lo0, hi0, r.nb.data_ptr(), *gbp, r.MARKS[0], r.MARKS[2], r.width_inner, r.width))
        assert dt0.value == dt1.value
        a, b = s.field(), s.field()
        s._ok(r._fstage(a.data_ptr(), r.phi.data_ptr(), r.phi.data_ptr(), *gbp, 1, 1.0, None, dt1.value, *g.dx,
                        *ptrs, lo0, hi0, hi3, r.nb.data_ptr(), *gbp, r.MARKS[0], r.MARKS[2], r.width_inner,
                        r.width))
        s._ok(r._fsrhs(b.data_ptr(), r.phi.data_ptr(), r.phi.data_ptr(), *gbp, 1, rhs.data_ptr(), dt1.value, *ptrs,
                       lo0, hi0, hi3, r.nb.data_ptr(), *gbp, r.MARKS[0]))
        s.lib.check()
        assert np.array_equal(s.host(a), s.host(b)), prec


def test_stored_rhs_marks_points_whose_stencil_leaves_the_band():
    """A band of one level: the CFL step reports the level-0 points whose
    stencil reaches outside the band and marks their RHS entries; the stage
    from that RHS skips exactly those points (their output keeps its value)
    and reports them too, while every other point is updated as by the fused
    stage."""
    from lsmlib_amd.lib import LsmHipError
    from lsmlib_amd.solver import NarrowBandRun
    r = NarrowBandRun(n=48, steps=1, rebuild_every=5, level=1, fused=True)
    r.rebuild()
    s, g = r.solver, r.grid
    gbp = s._gb()
    lo0, hi0, hi1 = int(r.n_lo[0]), int(r.n_hi[0]), int(r.n_hi[1])
    ptrs = [b.data_ptr() for b in (r.ix, r.iy, r.iz)]
    rhs = s.torch.zeros(hi0 - lo0 + 1, dtype=s.dtype, device=s.device)
    dt = s.real(0)
    with pytest.raises(LsmHipError, match="error 4"):
        s._ok(r._fdtr(ctypes.byref(dt), rhs.data_ptr(), r.phi.data_ptr(), *gbp, 1.0, None, *g.dx, r.CFL, *ptrs,
                      lo0, hi0, r.nb.data_ptr(), *gbp, r.MARKS[0], r.MARKS[2], r.width_inner, r.width))
    s.lib.clear_error()
    step = 0.25 * g.dx[0]
    a, b = s.field(), s.field()
    for f in (a, b):
        f.fill_(7.0)
    with pytest.raises(LsmHipError, match="error 4"):
        s._ok(r._fstage(a.data_ptr(), r.phi.data_ptr(), r.phi.data_ptr(), *gbp, 1, 1.0, None, step, *g.dx, *ptrs,
                        lo0, hi0, hi1, r.nb.data_ptr(), *gbp, r.MARKS[0], r.MARKS[2], r.width_inner, r.width))
    s.lib.clear_error()
    with pytest.raises(LsmHipError, match="error 4"):
        s._ok(r._fsrhs(b.data_ptr(), r.phi.data_ptr(), r.phi.data_ptr(), *gbp, 1, rhs.data_ptr(), step, *ptrs,
                       lo0, hi0, hi1, r.nb.data_ptr(), *gbp, r.MARKS[0]))
    s.lib.clear_error()
    ha, hb = s.host(a), s.host(b)
    assert np.array_equal(ha, hb)
    assert (ha == 7.0).sum() > 0 and (ha != 7.0).sum() > 0
