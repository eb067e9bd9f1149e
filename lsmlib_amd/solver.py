"""Host side of the time-stepping path: grids, seeded inputs and RK drivers.

Arrays are torch tensors on the GPU (plumbing only: device memory and the
stream); every operation on them is a call into the HIP library through its
C-ABI.  A field on a grid with n = (nx, ny, nz) interior points and g ghost
layers is stored column-major over the ghost box (-g .. n-1+g per axis), i.e.
a contiguous tensor of shape (nz+2g, ny+2g, nx+2g).  Coordinates follow the
reference's Grid (src/utils/lsm_grid.c:189-232, src/toolbox/lsm_initialization3d.c:142-144):
dx = (x_hi - x_lo)/n and x(i) = x_lo + i*dx for index i.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from .lib import LsmHipError, load

BC_KIND = {"linear": 0, "signed_linear": 1, "copy": 2, "neumann": 2, "periodic": 3}
STAGE3 = (1, 2, 3)


# ---------------------------------------------------------------------------
# Seeded inputs: splitmix64 -> doubles in [0, 1)
# ---------------------------------------------------------------------------
class SplitMix64:
    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))


class Grid:
    """n interior points per axis on [lo, hi], `ghost` layers.  Node i sits at
    lo + i*dx (the reference's Grid); centred=True puts it at lo + (i+1/2)*dx."""

    def __init__(self, n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), ghost=3, centred=False):
        self.n = tuple(int(v) for v in n)
        self.lo = tuple(float(v) for v in lo)
        self.hi = tuple(float(v) for v in hi)
        self.g = int(ghost)
        self.centred = bool(centred)
        self.dx = tuple((self.hi[d] - self.lo[d]) / self.n[d] for d in range(3))
        g = self.g
        self.gb = (-g, self.n[0] - 1 + g, -g, self.n[1] - 1 + g, -g, self.n[2] - 1 + g)
        self.fb = (0, self.n[0] - 1, 0, self.n[1] - 1, 0, self.n[2] - 1)
        self.shape = (self.n[2] + 2 * g, self.n[1] + 2 * g, self.n[0] + 2 * g)

    @property
    def cells(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    def axis(self, d):
        """Coordinates of the ghost-box index range along axis d (float64)."""
        i = np.arange(-self.g, self.n[d] + self.g, dtype=np.float64)
        return self.lo[d] + self.dx[d] * (i + 0.5 if self.centred else i)

    def mesh(self):
        z, y, x = np.meshgrid(self.axis(2), self.axis(1), self.axis(0), indexing="ij")
        return x, y, z

    def interior(self, a):
        g = self.g
        return a[g:a.shape[0] - g, g:a.shape[1] - g, g:a.shape[2] - g]


# ---------------------------------------------------------------------------
# Initial level sets of the benchmark configurations
# ---------------------------------------------------------------------------
def phi_perturbed_sphere(grid: Grid, seed: int):
    """(|x-c| - 0.5)(1 + 0.1 sin 3pi x sin 3pi y sin 3pi z), c ~ U[-0.1, 0.1]^3: not a distance function."""
    r = SplitMix64(seed)
    c = [r.uniform(-0.1, 0.1) for _ in range(3)]
    x, y, z = grid.mesh()
    d = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    return (d - 0.5) * (1 + 0.1 * np.sin(3 * np.pi * x) * np.sin(3 * np.pi * y) * np.sin(3 * np.pi * z)), c


def phi_spheres_min(grid: Grid, seed: int, count: int, centre_range, radius_range, chunk=16, threads=8):
    """min over seeded spheres of (|x - c_l| - r_l), in FP64 on the host.

    Work is split into (k, j) tiles of `chunk` x `chunk` lines.  A sphere is
    skipped on a tile when a lower bound of its value there,
    sqrt(dz_min^2 + dy_min^2) - r, is not below the tile's current maximum:
    rounding is monotone in each operation, so a skipped sphere could not
    lower any value and the result is bitwise the unpruned minimum."""
    import concurrent.futures as cf
    r = SplitMix64(seed)
    spheres = []
    for _ in range(count):
        c = [r.uniform(*centre_range) for _ in range(3)]
        spheres.append((c, r.uniform(*radius_range)))
    ax = [grid.axis(d) for d in range(3)]
    out = np.full(grid.shape, np.inf)
    nz, ny = grid.shape[0], grid.shape[1]

    def nearest(a, lo, hi, c):
        seg = a[lo:hi]
        return 0.0 if seg[0] <= c <= seg[-1] else min(abs(seg[0] - c), abs(seg[-1] - c))

    def tile(k0, j0):
        o = out[k0:k0 + chunk, j0:j0 + chunk]
        for (c, rad) in spheres:
            dzm = nearest(ax[2], k0, k0 + chunk, c[2])
            dym = nearest(ax[1], j0, j0 + chunk, c[1])
            if math.sqrt(dzm * dzm + dym * dym) - rad >= o.max():
                continue
            dz2 = (ax[2][k0:k0 + chunk] - c[2]) ** 2
            dy2 = (ax[1][j0:j0 + chunk] - c[1]) ** 2
            dx2 = (ax[0] - c[0]) ** 2
            np.minimum(o, np.sqrt(dz2[:, None, None] + dy2[None, :, None] + dx2[None, None, :]) - rad, out=o)

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda kj: tile(*kj), [(k0, j0) for k0 in range(0, nz, chunk) for j0 in range(0, ny, chunk)]))
    return out, spheres


def phi_sphere(grid: Grid, centre, radius, chunk=32):
    ax = [grid.axis(d) for d in range(3)]
    out = np.empty(grid.shape)
    dx2 = (ax[0] - centre[0]) ** 2
    dy2 = (ax[1] - centre[1]) ** 2
    dz2 = (ax[2] - centre[2]) ** 2
    for k0 in range(0, grid.shape[0], chunk):
        out[k0:k0 + chunk] = np.sqrt(dz2[k0:k0 + chunk, None, None] + dy2[None, :, None] + dx2[None, None, :]) - radius
    return out


def deformation_tables(grid: Grid):
    """Per-axis A = sin^2(pi x), B = sin(2 pi x) over the ghost-box index range."""
    A = [np.sin(np.pi * grid.axis(d)) ** 2 for d in range(3)]
    B = [np.sin(2 * np.pi * grid.axis(d)) for d in range(3)]
    return A + B


# ---------------------------------------------------------------------------
# Solver: RK3 steps through the fused stage kernels or the drop-in chain
# ---------------------------------------------------------------------------
def _box_args(b):
    return [ctypes.byref(ctypes.c_int(v)) for v in b]


class Solver:
    """Level-set time stepping on one grid in one precision ("f64" or "f32")."""

    def __init__(self, grid: Grid, precision="f64", bc="signed_linear", device="cuda"):
        import torch

        self.torch = torch
        self.grid = grid
        self.precision = precision
        self.lib = load(precision)
        self.dtype = torch.float64 if precision == "f64" else torch.float32
        self.real = self.lib.real
        self.device = device
        self.bc = bc
        self._keep = []
        L = self.lib
        P, I, R = ctypes.c_void_p, ctypes.c_int, self.real
        box = [ctypes.POINTER(ctypes.c_int)] * 6
        self._fill = L.fn("lsmhip_fill_ghosts", [P] + box + box + [I, I])
        stage_head = [P, P, P] + box + box + [I]
        self._normal = L.fn("lsmhip_weno5_stage_normal", stage_head + [R, P, R, R, R, R])
        self._advect = L.fn("lsmhip_weno5_stage_advect", stage_head + [P, P, P, R, R, R, R])
        self._deform = L.fn("lsmhip_weno5_stage_advect_deform", stage_head + [P, R, R, R, R, R])
        self._reinit = L.fn("lsmhip_weno5_stage_reinit", stage_head + [P, I, R, R, R, R])
        self._velfill = L.fn("lsmhip_deform_velocity_fill", [P, P, P] + box + box + [P, R])
        self._vol = L.fn("lsmhip_volume_phi_lt0_f64",
                         [P, P] + box + box + [ctypes.POINTER(ctypes.c_double)] * 4)
        self._gbp = [ctypes.c_int(v) for v in grid.gb]
        self._fbp = [ctypes.c_int(v) for v in grid.fb]
        self._tabs = None

    # -- memory ---------------------------------------------------------------
    def field(self, values=None):
        t = self.torch
        if values is None:
            return t.zeros(self.grid.shape, dtype=self.dtype, device=self.device)
        return t.as_tensor(np.ascontiguousarray(values), dtype=self.dtype).to(self.device)

    def host(self, a):
        return a.detach().to("cpu", self.torch.float64).numpy()

    def _gb(self):
        return [ctypes.byref(v) for v in self._gbp]

    def _fb(self):
        return [ctypes.byref(v) for v in self._fbp]

    def _ok(self, rc):
        if rc != 0:
            self.lib.check()
            raise LsmHipError("fused entry point failed")

    # -- ghost cells --------------------------------------------------------------
    def fill_ghosts(self, phi, kind=None):
        kind = BC_KIND[kind or self.bc]
        self._ok(self._fill(phi.data_ptr(), *self._gb(), *self._fb(), 9, kind))

    # -- fused stages ---------------------------------------------------------------
    def stage_normal(self, out, us, uc, stage, dt, vn=1.0, vn_field=None):
        g = self.grid
        self._ok(self._normal(out.data_ptr(), us.data_ptr(), uc.data_ptr() if uc is not None else None,
                              *self._gb(), *self._fb(), stage, vn,
                              vn_field.data_ptr() if vn_field is not None else None, dt, *g.dx))

    def stage_advect(self, out, us, uc, stage, dt, vel):
        g = self.grid
        self._ok(self._advect(out.data_ptr(), us.data_ptr(), uc.data_ptr() if uc is not None else None,
                              *self._gb(), *self._fb(), stage, vel[0].data_ptr(), vel[1].data_ptr(),
                              vel[2].data_ptr(), dt, *g.dx))

    def deform_tables(self):
        if self._tabs is None:
            tabs = [self.torch.as_tensor(a, dtype=self.dtype).to(self.device) for a in deformation_tables(self.grid)]
            arr = (ctypes.c_void_p * 6)(*[a.data_ptr() for a in tabs])
            self._tabs = (tabs, arr)
        return self._tabs[1]

    def stage_deform(self, out, us, uc, stage, dt, c):
        g = self.grid
        self._ok(self._deform(out.data_ptr(), us.data_ptr(), uc.data_ptr() if uc is not None else None,
                              *self._gb(), *self._fb(), stage, self.deform_tables(), c, dt, *g.dx))

    def stage_reinit(self, out, us, uc, stage, dt, phi0=None):
        g = self.grid
        self._ok(self._reinit(out.data_ptr(), us.data_ptr(), uc.data_ptr() if uc is not None else None,
                              *self._gb(), *self._fb(), stage, phi0.data_ptr() if phi0 is not None else None,
                              1 if phi0 is not None else 0, dt, *g.dx))

    def deform_velocity(self, c):
        vel = [self.field(None) for _ in range(3)]
        self._ok(self._velfill(vel[0].data_ptr(), vel[1].data_ptr(), vel[2].data_ptr(), *self._gb(), *self._gb(),
                               self.deform_tables(), c))
        return vel

    # -- whole RK3 steps --------------------------------------------------------------
    def rk3_fused(self, phi, work, stage_fn, dt, **kw):
        """phi <- TVD-RK3 step; work = (phi1, phi2) scratch fields.  stage_fn(out, us, uc, stage, dt, **kw)."""
        p1, p2 = work
        stage_fn(p1, phi, None, 1, dt, **kw)
        self.fill_ghosts(p1)
        stage_fn(p2, p1, phi, 2, dt, **kw)
        self.fill_ghosts(p2)
        stage_fn(phi, p2, phi, 3, dt, **kw)
        self.fill_ghosts(phi)

    def volume_phi_lt0(self, phi, eps):
        out = ctypes.c_double(0.0)
        d = [ctypes.c_double(v) for v in self.grid.dx] + [ctypes.c_double(eps)]
        self._ok(self._vol(ctypes.byref(out), phi.data_ptr(), *self._gb(), *self._fb(), *[ctypes.byref(v) for v in d]))
        return out.value


class DropInChain:
    """The reference's call sequence for one TVD-RK3 step, issued through the
    drop-in entry points (one full-grid pass per derivative, term and stage)."""

    def __init__(self, solver: Solver):
        self.s = solver
        g = solver.grid
        f = solver.field
        self.grads = [f() for _ in range(6)]
        self.rhs = f()
        self.d1 = f()
        self.api = solver.lib.api

    def rhs_normal(self, phi, vn):
        g, a, G = self.s.grid, self.api, self.grads
        gb, fb = g.gb, g.fb
        a.lsm3dhjweno5(G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, phi, gb, self.d1, gb, fb, *g.dx)
        a.lsm3dzerooutlevelseteqnrhs(self.rhs, gb)
        a.lsm3daddconstnormalveltermtolserhs(self.rhs, gb, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, vn, fb)

    def rhs_advect(self, phi, vel):
        g, a, G = self.s.grid, self.api, self.grads
        gb, fb = g.gb, g.fb
        a.lsm3dupwindhjweno5(G[0], G[1], G[2], gb, phi, gb, vel[0], vel[1], vel[2], gb, self.d1, gb, fb, *g.dx)
        a.lsm3dzerooutlevelseteqnrhs(self.rhs, gb)
        a.lsm3daddadvectiontermtolserhs(self.rhs, gb, G[0], G[1], G[2], gb, vel[0], vel[1], vel[2], gb, fb)

    def rhs_reinit(self, phi, phi0=None):
        g, a, G = self.s.grid, self.api, self.grads
        gb, fb = g.gb, g.fb
        a.lsm3dhjweno5(G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, phi, gb, self.d1, gb, fb, *g.dx)
        a.lsm3dcomputereinitializationeqnrhs(self.rhs, gb, phi, gb, phi0 if phi0 is not None else phi, gb,
                                             G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, fb, *g.dx,
                                             1 if phi0 is not None else 0)

    def rk3(self, phi, work, rhs_fn, dt, **kw):
        g, a = self.s.grid, self.api
        gb, fb = g.gb, g.fb
        p1, p2 = work
        rhs_fn(phi, **kw)
        a.lsm3dtvdrk3stage1(p1, gb, phi, gb, self.rhs, gb, fb, dt)
        self.s.fill_ghosts(p1)
        rhs_fn(p1, **kw)
        a.lsm3dtvdrk3stage2(p2, gb, p1, gb, phi, gb, self.rhs, gb, fb, dt)
        self.s.fill_ghosts(p2)
        rhs_fn(p2, **kw)
        a.lsm3dtvdrk3stage3(phi, gb, p2, gb, phi, gb, self.rhs, gb, fb, dt)
        self.s.fill_ghosts(phi)


_CFL_CACHE = {}


def advection_cfl_dt(grid: Grid, cfl: float, c: float) -> float:
    """CFL step of lsm3dComputeStableAdvectionDt for the deformation field
    scaled by c.  max |u|/dx + |v|/dy + |w|/dz over the interior is |c| times
    the maximum of the unscaled field, which is evaluated once per grid on the
    host in FP64 (it is separable per axis)."""
    key = (grid.n, grid.lo, grid.hi)
    if key not in _CFL_CACHE:
        A = [np.sin(np.pi * grid.axis(d)[grid.g:grid.g + grid.n[d]]) ** 2 for d in range(3)]
        B = [np.abs(np.sin(2 * np.pi * grid.axis(d)[grid.g:grid.g + grid.n[d]])) for d in range(3)]
        ix, iy, iz = (1.0 / v for v in grid.dx)
        best = 0.0
        for k in range(grid.n[2]):
            s = (2 * A[0][None, :] * B[1][:, None] * B[2][k] * ix + B[0][None, :] * A[1][:, None] * B[2][k] * iy
                 + B[0][None, :] * B[1][:, None] * A[2][k] * iz)
            best = max(best, float(s.max()))
        _CFL_CACHE[key] = best
    return cfl / (_CFL_CACHE[key] * abs(c) + 1e-36)


def deformation_time_factor(t: float, period: float = 3.0) -> float:
    return math.cos(math.pi * t / period)


# ---------------------------------------------------------------------------
# Deformation-advection workload (sphere through the time-reversing field)
# ---------------------------------------------------------------------------
class DeformationRun:
    """A sphere of radius 0.15 at (0.35, 0.35, 0.35) on [0,1]^3 advected by the
    divergence-free deformation field times cos(pi t / 3) with HJ-WENO5 +
    TVD-RK3; the field is evaluated at the stage times t, t+dt, t+dt/2.  dt is
    the CFL-0.5 advection step of the unscaled field.  Every `reinit_every`
    steps phi is reinitialized with HJ-WENO5 + TVD-RK3 pseudo-steps of
    0.5*dx up to 5*dx (the schedule of
    examples/curvature_driven_motion/curvature_model3d.c:484-624)."""

    def __init__(self, n=512, precision="f32", bc="signed_linear", reinit_every=20, cfl=0.5, dropin=False):
        self.grid = Grid((n, n, n))
        self.solver = Solver(self.grid, precision, bc=bc)
        self.reinit_every = reinit_every
        self.dt = advection_cfl_dt(self.grid, cfl, 1.0)
        self.dt_r = 0.5 * min(self.grid.dx)
        self.reinit_steps = int(round(5.0 / 0.5))
        s = self.solver
        self.phi = s.field(phi_sphere(self.grid, (0.35, 0.35, 0.35), 0.15))
        s.fill_ghosts(self.phi)
        self.work = (s.field(), s.field())
        self.t = 0.0
        self.nsteps = 0
        self.chain = DropInChain(s) if dropin else None
        self.events = None  # list of (stage kind, start event, end event) when instrumented

    def _stage(self, kind, fn, *a, **kw):
        if self.events is None:
            return fn(*a, **kw)
        t = self.solver.torch
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        e0.record()
        fn(*a, **kw)
        e1.record()
        self.events.append((kind, e0, e1))

    def advect_step(self, dt=None):
        s, phi, (p1, p2), t = self.solver, self.phi, self.work, self.t
        dt = self.dt if dt is None else dt
        cs = [deformation_time_factor(t), deformation_time_factor(t + dt), deformation_time_factor(t + 0.5 * dt)]
        if self.chain is None:
            self._stage("advect1", s.stage_deform, p1, phi, None, 1, dt, cs[0])
            s.fill_ghosts(p1)
            self._stage("advect2", s.stage_deform, p2, p1, phi, 2, dt, cs[1])
            s.fill_ghosts(p2)
            self._stage("advect3", s.stage_deform, phi, p2, phi, 3, dt, cs[2])
            s.fill_ghosts(phi)
        else:
            g, a, ch = self.grid, s.lib.api, self.chain
            gb, fb = g.gb, g.fb
            for st, (src, out, c) in enumerate(((phi, p1, cs[0]), (p1, p2, cs[1]), (p2, phi, cs[2])), start=1):
                vel = s.deform_velocity(c)
                ch.rhs_advect(src, vel=vel)
                if st == 1:
                    a.lsm3dtvdrk3stage1(out, gb, phi, gb, ch.rhs, gb, fb, dt)
                elif st == 2:
                    a.lsm3dtvdrk3stage2(out, gb, src, gb, phi, gb, ch.rhs, gb, fb, dt)
                else:
                    a.lsm3dtvdrk3stage3(out, gb, src, gb, phi, gb, ch.rhs, gb, fb, dt)
                s.fill_ghosts(out)
        self.t += dt
        self.nsteps += 1

    def reinitialize(self):
        s, (p1, p2) = self.solver, self.work
        for _ in range(self.reinit_steps):
            if self.chain is None:
                self._stage("reinit1", s.stage_reinit, p1, self.phi, None, 1, self.dt_r)
                s.fill_ghosts(p1)
                self._stage("reinit2", s.stage_reinit, p2, p1, self.phi, 2, self.dt_r)
                s.fill_ghosts(p2)
                self._stage("reinit3", s.stage_reinit, self.phi, p2, self.phi, 3, self.dt_r)
                s.fill_ghosts(self.phi)
            else:
                self.chain.rk3(self.phi, self.work, self.chain.rhs_reinit, self.dt_r)

    def step(self, dt=None):
        """One advection step (of dt, default the CFL step), plus the
        reinitialization block when due.  Returns the number of full-grid
        WENO5-RK3 updates performed."""
        self.advect_step(dt)
        if self.reinit_every and self.nsteps % self.reinit_every == 0:
            self.reinitialize()
            return 1 + self.reinit_steps
        return 1

    def run_to(self, t_end=3.0):
        """Steps of the CFL step until t_end, the last one shortened to land on
        t_end (the field's period: the sphere is back at its start).  Returns
        the number of advection steps."""
        n = 0
        while self.t < t_end * (1.0 - 1e-14):
            self.step(min(self.dt, t_end - self.t))
            n += 1
        return n


def precision_study(n=512, t_end=3.0, log=None):
    """The deformation workload integrated over t in [0, t_end] in FP64 and in
    FP32 through the fused stages, from the same initial field.  Reports the
    FP32 field's L-inf / L1 difference from the FP64 one (all interior cells,
    and cells with |phi_64| < 6 dx), the relative difference of the smoothed
    volume of {phi < 0} (FP64 compensated sum, eps = 1.5 dx), and each
    precision's error against the initial sphere, which the time-reversed field
    brings back at t = 3."""
    import torch
    out = {"grid": n, "t_end": t_end}
    finals = {}
    for prec in ("f64", "f32"):
        run = DeformationRun(n, prec)
        t0 = time.perf_counter()
        steps = run.run_to(t_end)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        run.solver.lib.check()
        eps = 1.5 * run.grid.dx[0]
        vol = run.solver.volume_phi_lt0(run.phi, eps)
        phi = run.grid.interior(run.solver.host(run.phi)).copy()
        finals[prec] = (phi, vol)
        out[prec] = {"advection_steps": steps, "seconds": round(el, 2), "volume": vol}
        if log:
            log("%s: %d advection steps to t = %g in %.1f s, volume %.15g" % (prec, steps, t_end, el, vol))
        if prec == "f64":
            init = run.grid.interior(phi_sphere(run.grid, (0.35, 0.35, 0.35), 0.15))
            ini = run.solver.field(phi_sphere(run.grid, (0.35, 0.35, 0.35), 0.15))
            out["volume_initial"] = run.solver.volume_phi_lt0(ini, eps)
            del ini
        del run
        torch.cuda.empty_cache()
    (p64, v64), (p32, v32) = finals["f64"], finals["f32"]
    h = 1.0 / n
    band = np.abs(p64) < 6 * h
    d = np.abs(p32 - p64)
    out["fp32_vs_fp64"] = {"linf_phi": float(d.max()), "l1_phi": float(d.mean()),
                           "linf_phi_band6dx": float(d[band].max()), "l1_phi_band6dx": float(d[band].mean()),
                           "volume_rel_err": abs(v32 - v64) / abs(v64)}
    vi = out["volume_initial"]
    for prec, (p, v) in finals.items():
        e = np.abs(p - init)
        out[prec].update({"linf_vs_initial_band6dx": float(e[band].max()), "volume_rel_err_vs_initial":
                          abs(v - vi) / abs(vi)})
    return out


# ---------------------------------------------------------------------------
# Reinitialization workload
# ---------------------------------------------------------------------------
class ReinitRun:
    """Plain reinitialization: 128^3 cells on [-1,1]^3 (cell centres), 3
    ghost layers, phi0 = (|x-c| - 0.5)(1 + 0.1 sin 3pi x sin 3pi y sin 3pi z)
    with c ~ U[-0.1,0.1]^3 (splitmix64 seed), `steps` pseudo-steps of
    HJ-WENO5 -> reinitialization RHS (sign from the current phi) -> TVD-RK3
    with dt = 0.5 dx, linear extrapolation into the ghost cells after every
    stage.  dropin=True issues the reference's call sequence through the
    drop-in entry points; otherwise the fused reinitialization stage."""

    def __init__(self, n=128, seed=1, precision="f64", dropin=True, steps=50):
        self.grid = Grid((n, n, n), lo=(-1.0, -1.0, -1.0), hi=(1.0, 1.0, 1.0), centred=True)
        self.solver = Solver(self.grid, precision, bc="linear")
        phi0, self.centre = phi_perturbed_sphere(self.grid, seed)
        s = self.solver
        self.phi = s.field(phi0)
        s.fill_ghosts(self.phi)
        self.work = (s.field(), s.field())
        self.dt = 0.5 * self.grid.dx[0]
        self.steps = steps
        self.chain = DropInChain(s) if dropin else None

    def step(self):
        if self.chain is not None:
            self.chain.rk3(self.phi, self.work, self.chain.rhs_reinit, self.dt)
        else:
            self.solver.rk3_fused(self.phi, self.work, self.solver.stage_reinit, self.dt)

    def run(self):
        for _ in range(self.steps):
            self.step()
        return self.phi


# ---------------------------------------------------------------------------
# Periodic motion workload: advection + normal speed with a curvature term
# ---------------------------------------------------------------------------
def periodic_velocity(grid: Grid):
    """u = 0.5 (sin 2 pi y, sin 2 pi z, sin 2 pi x) over the ghost box."""
    x, y, z = grid.mesh()
    return 0.5 * np.sin(2 * np.pi * y), 0.5 * np.sin(2 * np.pi * z), 0.5 * np.sin(2 * np.pi * x)


def curvature_corrected_dt(dt_hyp: float, grid: Grid, b: float, cfl: float) -> float:
    """dt = cfl / (cfl/dt_hyp + 2 b sum 1/dx^2): the parabolic correction of
    examples/curvature_driven_motion/curvature_model3d.c:95-98, 260-266."""
    corr = 2.0 * b * sum(1.0 / (h * h) for h in grid.dx)
    return cfl / (cfl / dt_hyp + corr)


class PeriodicMotionRun:
    """256^3 on [0,1]^3 with periodic ghost cells.  phi0 = min over 32 seeded
    spheres (centres U[0.15,0.85]^3, radii U[0.04,0.12]).  phi_t + u.grad phi
    + V_n |grad phi| = 0 with u = 0.5 (sin 2pi y, sin 2pi z, sin 2pi x) and
    V_n = 1 - b kappa, b = 0.02, kappa the second-order mean curvature.  Each
    RK stage issues, through the drop-in entry points:
      CENTRAL_GRAD_ORDER2 -> GRADIENT_MAGNITUDE -> COMPUTE_MEAN_CURVATURE_ORDER2,
      UPWIND_HJ_WENO5 -> ZERO_OUT -> ADD_ADVECTION_TERM,
      HJ_WENO5 -> ADD_CONST_NORMAL_VEL_TERM (V = 1),
      ADD_CONST_PRECOMPUTED_CURV_TERM (b), TVD_RK3_STAGEn, periodic fill.
    dt: min of COMPUTE_STABLE_ADVECTION_DT and COMPUTE_STABLE_CONST_NORMAL_VEL_DT
    (from phi^n's WENO5 derivatives) at CFL 0.5, with the parabolic correction
    of the curvature term."""

    B = 0.02
    CFL = 0.5

    def __init__(self, n=256, seed=2, precision="f64", steps=100):
        self.grid = g = Grid((n, n, n))
        self.solver = s = Solver(g, precision, bc="periodic")
        phi0, self.spheres = phi_spheres_min(g, seed, 32, (0.15, 0.85), (0.04, 0.12))
        self.phi = s.field(phi0)
        s.fill_ghosts(self.phi)
        self.vel = [s.field(v) for v in periodic_velocity(g)]
        self.work = (s.field(), s.field())
        self.kappa, self.gmag = s.field(), s.field()
        self.cg = [s.field() for _ in range(3)]
        self.chain = DropInChain(s)
        self.steps = steps
        self.dts = []

    def rhs(self, u):
        g, a, ch = self.grid, self.solver.lib.api, self.chain
        gb, fb, G, C = g.gb, g.fb, ch.grads, self.cg
        a.lsm3dcentralgradorder2(C[0], C[1], C[2], gb, u, gb, fb, *g.dx)
        a.lsm3dgradientmagnitude(C[0], C[1], C[2], self.gmag, gb, fb)
        a.lsm3dcomputemeancurvatureorder2(self.kappa, gb, u, gb, C[0], C[1], C[2], self.gmag, gb, fb, *g.dx)
        ch.rhs_advect(u, vel=self.vel)
        a.lsm3dhjweno5(G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, u, gb, ch.d1, gb, fb, *g.dx)
        a.lsm3daddconstnormalveltermtolserhs(ch.rhs, gb, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, 1.0, fb)
        a.lsm3daddconstprecomputedcurvtermtolserhs(ch.rhs, gb, self.kappa, gb, self.gmag, gb, self.B, fb)

    def stable_dt(self):
        g, a, ch = self.grid, self.solver.lib.api, self.chain
        gb, fb, G = g.gb, g.fb, ch.grads
        dt_adv = a.lsm3dcomputestableadvectiondt(None, self.vel[0], self.vel[1], self.vel[2], gb, fb, *g.dx, self.CFL)
        a.lsm3dhjweno5(G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, self.phi, gb, ch.d1, gb, fb, *g.dx)
        dt_n = a.lsm3dcomputestableconstnormalveldt(None, 1.0, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, fb,
                                                    *g.dx, self.CFL)
        return curvature_corrected_dt(min(dt_adv, dt_n), g, self.B, self.CFL)

    def step(self):
        dt = self.stable_dt()
        self.dts.append(dt)
        self.chain.rk3(self.phi, self.work, lambda u: self.rhs(u), dt)

    def run(self):
        for _ in range(self.steps):
            self.step()
        return self.phi


# ---------------------------------------------------------------------------
# Narrow-band workload
# ---------------------------------------------------------------------------
class NarrowBandRun:
    """Narrow-band normal motion (V_n = 1) on n^3 nodes of [0,1]^3.  phi0 =
    min over 64 seeded spheres (centres U[0.1,0.9]^3, radii U[0.03,0.1]).
    Band: |phi| < width = 6 dx at level 0, outer layer from width_inner = 3 dx,
    levels 0..3, boundary layers marked as the reference's local driver marks
    them (examples/curvature_driven_motion/curvature_model3d_local.c:190-229;
    marks fb 124, D2 125, D1 126, gb 127).  The band is rebuilt every
    `rebuild_every` steps.  One step, all through the drop-in entry points:
      COPY_DATA_LOCAL phi -> both stage fields over levels 0..3 (the stencils
        read frozen phi^n at levels 1..3),
      HJ_WENO5_LOCAL(phi) -> COMPUTE_STABLE_CONST_NORMAL_VEL_DT_LOCAL (CFL 0.5),
      three stages of: HJ_WENO5_LOCAL (stage 1 reuses the derivatives above)
        -> ZERO_OUT_LOCAL -> ADD_CONST_NORMAL_VEL_TERM_LOCAL (V = 1)
        -> MULTIPLY_CUT_OFF_LSE_RHS_LOCAL(beta = width_inner, gamma = width)
        -> TVD_RK3_STAGEn_LOCAL -> signed linear extrapolation fill.
    fused=True runs the same step through lsmhip_weno5_local_normal_dt_rhs
    (the CFL step, which also stores the first stage's RHS per level-0 point
    from the derivatives it computes), lsmhip_local_stage_from_rhs for stage 1
    and lsmhip_weno5_local_stage_normal for stages 2 and 3, the copies folded
    into stages 1 and 2 (FP64: bitwise the same fields)."""

    MARKS = (124, 125, 126, 127)  # fb, D2, D1, gb
    CFL = 0.5

    def __init__(self, n=512, seed=3, precision="f64", steps=200, rebuild_every=10, level=3, phi0=None,
                 fused=False):
        import torch
        self.grid = g = Grid((n, n, n))
        self.solver = s = Solver(g, precision, bc="signed_linear")
        if phi0 is None:
            phi0, self.spheres = phi_spheres_min(g, seed, 64, (0.1, 0.9), (0.03, 0.1))
        self.phi = s.field(phi0)
        del phi0
        s.fill_ghosts(self.phi)
        self.work = (s.field(), s.field())
        self.grads = [s.field() for _ in range(6)]
        self.rhs, self.d1 = s.field(), s.field()
        self.npts = g.shape[0] * g.shape[1] * g.shape[2]
        dev = s.device
        self.nb = torch.zeros(g.shape, dtype=torch.uint8, device=dev)
        self.ix, self.iy, self.iz, self.outer = (torch.zeros(self.npts, dtype=torch.int32, device=dev)
                                                  for _ in range(4))
        self.level = level
        self.n_lo = np.zeros(level + 1, np.int32)
        self.n_hi = np.zeros(level + 1, np.int32)
        self.width, self.width_inner = 6 * g.dx[0], 3 * g.dx[0]
        self.steps, self.rebuild_every = steps, rebuild_every
        self.nsteps = 0
        self.dts, self.band_sizes = [], []
        self.api = s.lib.api
        self.on_rebuild = None  # callable(run) after each rebuild, e.g. to record the band
        self.fused = fused
        if fused:
            L = s.lib
            P, I, R, U8 = ctypes.c_void_p, ctypes.c_int, s.real, ctypes.c_ubyte
            box = [ctypes.POINTER(ctypes.c_int)] * 6
            self._fstage = L.fn("lsmhip_weno5_local_stage_normal",
                                [P, P, P] + box + [I, R, P, R, R, R, R, P, P, P, I, I, I, P] + box + [U8, U8, R, R])
            self._fdt = L.fn("lsmhip_weno5_local_normal_dt",
                             [ctypes.POINTER(R), P] + box + [R, P, R, R, R, R, P, P, P, I, I, P] + box + [U8, U8])
            self._fdtr = L.fn("lsmhip_weno5_local_normal_dt_rhs",
                              [ctypes.POINTER(R), P, P] + box + [R, P, R, R, R, R, P, P, P, I, I, P] + box +
                              [U8, U8, R, R])
            self._fsrhs = L.fn("lsmhip_local_stage_from_rhs",
                               [P, P, P] + box + [I, P, R, P, P, P, I, I, I, P] + box + [U8])
            self._rhs0 = None  # the first stage's RHS per level-0 entry

    def boxes(self):
        n, gb = self.grid.n, self.grid.gb
        d2 = (-1, n[0], -1, n[1], -1, n[2])
        d1 = (-2, n[0] + 1, -2, n[1] + 1, -2, n[2] + 1)
        return d2, d1, gb

    def rebuild(self):
        a, g = self.api, self.grid
        gb = g.gb
        self.outer_bounds = a.lsm3ddeterminenarrowband(
            self.phi, gb, self.nb, gb, self.ix, self.iy, self.iz, 0, self.npts - 1, self.n_lo, self.n_hi,
            self.outer, 0, self.npts - 1, None, None, None, None, self.width, self.width_inner, self.level)
        self.solver.lib.check()
        if (self.n_hi < self.n_lo).any() or (self.n_lo < 0).any():
            raise LsmHipError("empty narrow-band level: n_lo=%s n_hi=%s" % (self.n_lo, self.n_hi))
        for box, mark in zip(self.boxes(), self.MARKS[1:]):
            a.lsm3dmarknarrowbandboundarylayer(self.nb, gb, box, mark)
        self.band_sizes.append(int(self.n_hi[0] - self.n_lo[0] + 1))
        if self.on_rebuild is not None:
            self.on_rebuild(self)

    def _weno(self, u):
        a, g, G = self.api, self.grid, self.grads
        gb = g.gb
        lv = [int(v) for pair in zip(self.n_lo, self.n_hi) for v in pair]
        a.lsm3dhjweno5local(G[0], G[1], G[2], gb, G[3], G[4], G[5], gb, u, gb, self.d1, gb, *g.dx,
                            self.ix, self.iy, self.iz, *lv[:8], self.nb, gb, self.MARKS[0], self.MARKS[2])

    def step(self):
        if self.fused:
            return self._step_fused()
        a, g, G = self.api, self.grid, self.grads
        gb = g.gb
        if self.nsteps % self.rebuild_every == 0:
            self.rebuild()
        lo0, hi0, hi3 = int(self.n_lo[0]), int(self.n_hi[0]), int(self.n_hi[self.level])
        band = (self.ix, self.iy, self.iz)
        mfb = self.MARKS[0]
        phi, (p1, p2) = self.phi, self.work
        for w in (p1, p2):
            a.lsm3dcopydatalocal(w, gb, phi, gb, *band, lo0, hi3)
        self._weno(phi)
        dt = a.lsm3dcomputestableconstnormalveldtlocal(None, 1.0, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb,
                                                       *g.dx, self.CFL, *band, lo0, hi0, self.nb, gb, mfb)
        self.dts.append(dt)
        for st, (src, out) in enumerate(((phi, p1), (p1, p2), (p2, phi)), start=1):
            if st > 1:
                self._weno(src)
            a.lsm3dzerooutlevelseteqnrhslocal(self.rhs, gb, *band, lo0, hi0)
            a.lsm3daddconstnormalveltermtolserhslocal(self.rhs, gb, G[0], G[1], G[2], gb, G[3], G[4], G[5], gb,
                                                      1.0, *band, lo0, hi0, self.nb, gb, mfb)
            a.lsm3dmultiplycutofflserhslocal(src, self.rhs, gb, *band, lo0, hi0, self.nb, gb, mfb,
                                             self.width_inner, self.width)
            if st == 1:
                a.lsm3dtvdrk3stage1local(out, gb, phi, gb, self.rhs, gb, dt, *band, lo0, hi0, self.nb, gb, mfb)
            elif st == 2:
                a.lsm3dtvdrk3stage2local(out, gb, p1, gb, phi, gb, self.rhs, gb, dt, *band, lo0, hi0, self.nb, gb,
                                         mfb)
            else:
                a.lsm3dtvdrk3stage3local(out, gb, p2, gb, phi, gb, self.rhs, gb, dt, *band, lo0, hi0, self.nb, gb,
                                         mfb)
            self.solver.fill_ghosts(out)
        self.solver.lib.check()
        self.nsteps += 1

    def _step_fused(self):
        a, g, s = self.api, self.grid, self.solver
        gb = g.gb
        if self.nsteps % self.rebuild_every == 0:
            self.rebuild()
        lo0, hi0, hi3 = int(self.n_lo[0]), int(self.n_hi[0]), int(self.n_hi[self.level])
        band = (self.ix, self.iy, self.iz)
        ptrs = [b.data_ptr() for b in band]
        mfb, md1 = self.MARKS[0], self.MARKS[2]
        phi, (p1, p2) = self.phi, self.work
        gbp = s._gb()
        n0 = hi0 - lo0 + 1
        if self._rhs0 is None or self._rhs0.numel() < n0:
            self._rhs0 = s.torch.empty(max(n0, 1), dtype=s.dtype, device=s.device)
        dt = s.real(0)
        s._ok(self._fdtr(ctypes.byref(dt), self._rhs0.data_ptr(), phi.data_ptr(), *gbp, 1.0, None, *g.dx, self.CFL,
                         *ptrs, lo0, hi0, self.nb.data_ptr(), *gbp, mfb, md1, self.width_inner, self.width))
        dt = dt.value
        self.dts.append(dt)
        for st, (src, out) in enumerate(((phi, p1), (p1, p2), (p2, phi)), start=1):
            # stages 1 and 2 also carry phi^n into their outputs over levels
            # 0..3 (the drop-in sequence copies phi there before stage 1)
            if st == 1:
                s._ok(self._fsrhs(out.data_ptr(), src.data_ptr(), phi.data_ptr(), *gbp, st, self._rhs0.data_ptr(),
                                  dt, *ptrs, lo0, hi0, hi3, self.nb.data_ptr(), *gbp, mfb))
            else:
                s._ok(self._fstage(out.data_ptr(), src.data_ptr(), phi.data_ptr(), *gbp, st, 1.0, None, dt, *g.dx,
                                   *ptrs, lo0, hi0, hi3 if st < 3 else -1, self.nb.data_ptr(), *gbp, mfb, md1,
                                   self.width_inner, self.width))
            s.fill_ghosts(out)
        s.lib.check()
        self.nsteps += 1

    def run(self):
        for _ in range(self.steps):
            self.step()
        return self.phi
