"""Reverse mode of level-set normal motion through the HIP library, and the
torch.autograd.Function wrappers of a shape-matching inverse problem.

Forward: phi_{n+1} = TVD-RK3 step of phi_t + V(x) |grad phi| = 0 with HJ-WENO5
derivatives and the Godunov Hamiltonian (lsmhip_weno5_stage_normal, the fused
form of HJ_WENO5 -> ADD_NORMAL_VEL_TERM -> TVD_RK3_STAGEn of the reference),
periodic ghost cells, fixed dt.  Reverse: lsmhip_weno5_stage_normal_adjoint per
stage and lsmhip_periodic_fill_adjoint per ghost fill, with the forward
recomputed from checkpoints stored every `checkpoint_every` steps.

Each forward run returns its own Trajectory (checkpoints, speed, step), and
the reverse sweep takes the trajectory it walks back, so that several forward
runs on one engine (several autograd nodes before one backward) keep their
own states.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from .lib import LsmHipError
from .solver import Grid, Solver, SplitMix64, phi_sphere


class Trajectory:
    """What the reverse sweep of one forward run needs: the checkpoints (phi at
    steps 0, k, 2k, ...), the speed field and the step."""

    __slots__ = ("checkpoints", "V", "dt", "steps")

    def __init__(self, checkpoints, V, dt, steps):
        self.checkpoints, self.V, self.dt, self.steps = checkpoints, V, dt, steps


class NormalMotionAdjoint:
    def __init__(self, grid: Grid, precision="f32", steps=200, dt=None, checkpoint_every=20, device="cuda"):
        self.grid = grid
        self.solver = s = Solver(grid, precision, bc="periodic", device=device)
        self.torch = s.torch
        self.steps = int(steps)
        self.k = int(checkpoint_every)
        self.dt = dt if dt is not None else default_dt(grid)
        L = s.lib
        P, I, R = ctypes.c_void_p, ctypes.c_int, s.real
        box = [ctypes.POINTER(ctypes.c_int)] * 6
        self._adj = L.fn("lsmhip_weno5_stage_normal_adjoint",
                         [P, P, P, P, P] + box + box + [I, R, P, R, R, R, R, P])
        self._fold = L.fn("lsmhip_periodic_fill_adjoint", [P] + box + box)
        self._scratch = None
        self.last = None  # the Trajectory of the latest forward run with checkpoints

    # -- forward -------------------------------------------------------------------
    def _step(self, phi, p1, p2, V):
        s, dt = self.solver, self.dt
        s.stage_normal(p1, phi, None, 1, dt, vn=0.0, vn_field=V)
        s.fill_ghosts(p1)
        s.stage_normal(p2, p1, phi, 2, dt, vn=0.0, vn_field=V)
        s.fill_ghosts(p2)
        s.stage_normal(phi, p2, phi, 3, dt, vn=0.0, vn_field=V)
        s.fill_ghosts(phi)

    def run(self, phi0, V, keep_checkpoints=True):
        """phi0, V: tensors on the ghost box (ghost entries of phi0 are
        replaced by the periodic fill; V is read on the fill box).  Returns
        (phi_T, Trajectory); the trajectory holds no checkpoints when
        keep_checkpoints is false."""
        s = self.solver
        phi = phi0.detach().to(s.dtype).clone().contiguous()
        Vs = V.detach().to(s.dtype).contiguous().clone()
        s.fill_ghosts(phi)
        p1, p2 = s.field(), s.field()
        checkpoints = []
        for n in range(self.steps):
            if keep_checkpoints and n % self.k == 0:
                checkpoints.append(phi.clone())
            self._step(phi, p1, p2, Vs)
        return phi, Trajectory(checkpoints, Vs, self.dt, self.steps)

    def forward(self, phi0, V, keep_checkpoints=True):
        """phi_T of run(); with checkpoints, its trajectory becomes the one
        backward() walks back when it is given none."""
        phi, traj = self.run(phi0, V, keep_checkpoints)
        if keep_checkpoints:
            self.last = traj
        return phi

    # -- reverse ----------------------------------------------------------------------
    def _stage_adjoint(self, lam_us, lam_uc, lam_v, lam_out, us, stage, V, dt):
        g, s = self.grid, self.solver
        if self._scratch is None:
            self._scratch = self.torch.empty((21,) + g.shape, dtype=s.dtype, device=s.device)
        rc = self._adj(lam_us.data_ptr(), lam_uc.data_ptr() if lam_uc is not None else None,
                       lam_v.data_ptr() if lam_v is not None else None, lam_out.data_ptr(), us.data_ptr(),
                       *s._gb(), *s._fb(), stage, 0.0, V.data_ptr(), dt, *g.dx, self._scratch.data_ptr())
        if rc != 0:
            s.lib.check()
            raise LsmHipError("stage adjoint failed")

    def _fold_ghosts(self, lam):
        s = self.solver
        if self._fold(lam.data_ptr(), *s._gb(), *s._fb()) != 0:
            s.lib.check()
            raise LsmHipError("periodic fill adjoint failed")

    def backward(self, lam_T, trajectory=None):
        """Given dJ/dphi_T (ghost box) of the run that produced `trajectory`
        (default: the latest forward run with checkpoints), returns
        (dJ/dphi0, dJ/dV) on the ghost box."""
        s, t = self.solver, self.torch
        tr = trajectory if trajectory is not None else self.last
        if tr is None or not tr.checkpoints:
            raise LsmHipError("backward needs a forward run with checkpoints")
        V, dt, steps = tr.V, tr.dt, tr.steps
        lam = lam_T.detach().to(s.dtype).clone().contiguous()
        lam_v = s.field()
        nseg = len(tr.checkpoints)
        states = None
        for seg in range(nseg - 1, -1, -1):
            n0 = seg * self.k
            n1 = min(n0 + self.k, steps)
            # recompute the segment, keeping the three stage inputs of every step
            if states is None:
                states = [[s.field() for _ in range(3)] for _ in range(self.k)]
            phi = tr.checkpoints[seg].clone()
            for m in range(n1 - n0):
                u0, u1, u2 = states[m]
                u0.copy_(phi)
                s.stage_normal(u1, u0, None, 1, dt, vn=0.0, vn_field=V)
                s.fill_ghosts(u1)
                s.stage_normal(u2, u1, u0, 2, dt, vn=0.0, vn_field=V)
                s.fill_ghosts(u2)
                s.stage_normal(phi, u2, u0, 3, dt, vn=0.0, vn_field=V)
                s.fill_ghosts(phi)
            for m in range(n1 - n0 - 1, -1, -1):
                u0, u1, u2 = states[m]
                self._fold_ghosts(lam)  # adjoint of fill(phi_{n+1})
                lam0, lam1, lam2 = s.field(), s.field(), s.field()
                self._stage_adjoint(lam2, lam0, lam_v, lam, u2, 3, V, dt)
                self._fold_ghosts(lam2)
                self._stage_adjoint(lam1, lam0, lam_v, lam2, u1, 2, V, dt)
                self._fold_ghosts(lam1)
                self._stage_adjoint(lam0, None, lam_v, lam1, u0, 1, V, dt)
                lam = lam0
        self._fold_ghosts(lam)  # adjoint of the initial fill
        t.cuda.synchronize()
        return lam, lam_v


def default_dt(grid: Grid, vmax=1.3, cfl=0.5) -> float:
    """CFL step for speeds up to vmax with |grad phi| = 1: cfl dx / (vmax sqrt 3)."""
    return cfl * min(grid.dx) / (vmax * math.sqrt(3.0))


# ---------------------------------------------------------------------------
# torch.autograd wrappers
# ---------------------------------------------------------------------------
def _autograd_classes():
    import torch

    class Evolve(torch.autograd.Function):
        """phi_T = evolve(phi0, V) by engine.steps RK3 steps; backward by the
        hand-written adjoint."""

        @staticmethod
        def forward(ctx, phi0, V, engine):
            # the node keeps its own trajectory: later runs on the same engine do not touch it
            out, ctx.trajectory = engine.run(phi0, V)
            ctx.engine = engine
            return out.to(phi0.dtype)

        @staticmethod
        def backward(ctx, grad_out):
            lam0, lam_v = ctx.engine.backward(grad_out.contiguous(), ctx.trajectory)
            ctx.trajectory = None
            return lam0.to(grad_out.dtype), lam_v.to(grad_out.dtype), None

    class Mismatch(torch.autograd.Function):
        """sum over the fill box of (H_eps(phi) - H_eps(target))^2 dV (FP64)."""

        @staticmethod
        def forward(ctx, phi, target, solver, eps):
            ctx.solver, ctx.eps = solver, eps
            phi_c = phi.detach().to(solver.dtype).contiguous()
            tgt = target.detach().to(solver.dtype).contiguous()
            ctx.save_for_backward(phi_c, tgt)
            loss, _ = heaviside_mismatch(solver, phi_c, tgt, eps, want_grad=False)
            return torch.tensor(loss, dtype=torch.float64, device=phi.device)

        @staticmethod
        def backward(ctx, g):
            phi_c, tgt = ctx.saved_tensors
            _, grad = heaviside_mismatch(ctx.solver, phi_c, tgt, ctx.eps, want_grad=True)
            return grad.to(torch.float64) * g, None, None, None

    return Evolve, Mismatch


_CLASSES = None


def autograd_functions():
    global _CLASSES
    if _CLASSES is None:
        _CLASSES = _autograd_classes()
    return _CLASSES


def heaviside_mismatch(solver: Solver, phi, target, eps, want_grad=True):
    L = solver.lib
    box = [ctypes.POINTER(ctypes.c_int)] * 6
    fn = L.fn("lsmhip_heaviside_mismatch", [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
              + box + box + [ctypes.c_double] * 4)
    loss = ctypes.c_double(0.0)
    grad = solver.field() if want_grad else None
    rc = fn(ctypes.byref(loss), grad.data_ptr() if grad is not None else None, phi.data_ptr(), target.data_ptr(),
            *solver._gb(), *solver._fb(), *solver.grid.dx, eps)
    if rc != 0:
        L.check()
        raise LsmHipError("heaviside mismatch failed")
    return loss.value, grad


# ---------------------------------------------------------------------------
# The shape-matching problem: V(x) = 1 + 0.3 tanh(g(x)), g with 8 Fourier
# modes per axis
# ---------------------------------------------------------------------------
def fourier_modes(seed: int, modes=8):
    """Amplitudes a ~ N(0,1) scaled by 1/|k| and phases theta ~ U[0, 2pi) for
    wave vectors k in {1..modes}^3 (splitmix64 + Box-Muller)."""
    r = SplitMix64(seed)
    n = modes ** 3
    a = np.empty(n)
    th = np.empty(n)
    for m in range(n):
        u1, u2 = max(r.uniform(), 1e-300), r.uniform()
        a[m] = math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2)
        th[m] = 2 * math.pi * r.uniform()
    kk = np.arange(1, modes + 1, dtype=np.float64)
    kz, ky, kx = np.meshgrid(kk, kk, kk, indexing="ij")
    norm = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2).reshape(-1)
    return a / norm, th


class SpeedField:
    """V(x) = 1 + 0.3 tanh(c g(x)), g(x) = sum_k a_k cos(2 pi k.x + theta_k)
    over k in {1..8}^3; the phases theta are fixed by the seed, the amplitudes a
    are the parameters (initially the seed's), and c = 1/sqrt(sum a0^2 / 2)
    scales the seed's g to unit RMS.  Evaluated in torch so that autograd
    carries dJ/dV to dJ/da."""

    def __init__(self, grid: Grid, seed: int, modes=8, device="cuda"):
        import torch
        self.torch = torch
        a, th = fourier_modes(seed, modes)
        self.scale = 1.0 / math.sqrt(0.5 * float(np.sum(a ** 2)))
        self.a0 = torch.tensor(a, dtype=torch.float64, device=device)
        self.theta = torch.tensor(th, dtype=torch.float64, device=device).reshape(modes, modes, modes)
        k = torch.arange(1, modes + 1, dtype=torch.float64, device=device)
        ax = [torch.tensor(grid.axis(d), dtype=torch.float64, device=device) for d in range(3)]
        # E_d[k, i] = exp(2 pi i k x_i)
        self.E = [torch.exp(2j * math.pi * k[:, None] * x[None, :]) for x in ax]
        self.modes = modes

    def __call__(self, a):
        t = self.torch
        m = self.modes
        c = (a.reshape(m, m, m).to(t.complex128) * t.exp(1j * self.theta.to(t.complex128)))
        Ex, Ey, Ez = self.E
        t1 = t.einsum("zyx,xi->zyi", c, Ex)
        t2 = t.einsum("zyi,yj->zji", t1, Ey)
        g = t.einsum("zji,zk->kji", t2, Ez).real
        return 1.0 + 0.3 * t.tanh(self.scale * g)


class ShapeMatching:
    """Loss(a, phi0) = sum (H_eps(phi_T) - H_eps(phi_target))^2 dV, phi_T the
    level set after `steps` RK3 steps from phi0 with speed V(a); phi_target
    from the same phi0 under the speed of another seed.  eps = 1.5 dx."""

    def __init__(self, n=256, precision="f32", steps=200, checkpoint_every=20, seed=5, target_seed=6,
                 radius=0.2, device="cuda"):
        import torch
        self.torch = torch
        self.grid = g = Grid((n, n, n))
        self.engine = NormalMotionAdjoint(g, precision, steps, checkpoint_every=checkpoint_every, device=device)
        self.speed = SpeedField(g, seed, device=device)
        self.a = self.speed.a0.clone()
        self.phi0 = torch.as_tensor(phi_sphere(g, (0.5, 0.5, 0.5), radius), dtype=torch.float64).to(device)
        tgt_speed = SpeedField(g, target_seed, device=device)
        V_t = tgt_speed(tgt_speed.a0)
        self.target = self.engine.forward(self.phi0, V_t, keep_checkpoints=False).clone()
        self.eps = 1.5 * g.dx[0]

    def loss(self, a, phi0):
        Evolve, Mismatch = autograd_functions()
        V = self.speed(a)
        phiT = Evolve.apply(phi0, V, self.engine)
        return Mismatch.apply(phiT, self.target, self.engine.solver, self.eps)

    def value_and_grads(self, a=None, phi0=None):
        t = self.torch
        a = (self.a if a is None else a).detach().clone().requires_grad_(True)
        phi0 = (self.phi0 if phi0 is None else phi0).detach().clone().requires_grad_(True)
        L = self.loss(a, phi0)
        ga, gphi = t.autograd.grad(L, (a, phi0))
        return float(L.detach()), ga, gphi

    def value(self, a, phi0):
        Evolve, Mismatch = autograd_functions()
        with self.torch.no_grad():
            V = self.speed(a)
            phiT = self.engine.forward(phi0, V, keep_checkpoints=False)
            loss, _ = heaviside_mismatch(self.engine.solver, phiT, self.target.to(self.engine.solver.dtype),
                                         self.eps, want_grad=False)
        return loss
