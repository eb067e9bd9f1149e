"""Shared machinery for the parity tests.

A ``Case`` holds named inputs (arrays keyed by the argument names of
``lsmlib_amd.capi.SPECS``, boxes, scalars).  ``run(impl, sym, case)`` calls one
implementation of ``sym`` on private copies of the case's arrays and returns
the arrays after the call plus the scalar outputs, so that the reference
build, the oracle and the HIP library can be compared on identical inputs.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from lsmlib_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.environ.get("LSMHIP_ORACLE_SO", os.path.join(ROOT, "oracle", "liblsm_oracle.so"))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "liblsm_ref.so")

_cache = {}


def oracle():
    if "oracle" not in _cache:
        _cache["oracle"] = capi.Bound(ctypes.CDLL(ORACLE_SO), prefix="oracle_", strip_underscore=True)
    return _cache["oracle"]


def reference():
    """The reference's own FP64 kernels built by oracle/Makefile, or None."""
    if "ref" not in _cache:
        _cache["ref"] = capi.Bound(ctypes.CDLL(REF_SO)) if os.path.exists(REF_SO) else None
    return _cache["ref"]


def shape_of(box):
    if len(box) == 4:  # 2D box (ilo, ihi, jlo, jhi)
        ilo, ihi, jlo, jhi = box
        return (jhi - jlo + 1, ihi - ilo + 1)
    ilo, ihi, jlo, jhi, klo, khi = box
    return (khi - klo + 1, jhi - jlo + 1, ihi - ilo + 1)  # C order: i fastest (column-major)


def box_of_array(sym):
    """Map each array argument of sym to the box argument that describes it."""
    spec = capi.SPECS[sym]
    out = {}
    pending = []
    for kind, name in spec:
        if kind in ("R", "M"):
            pending.append(name)
        elif kind in ("B", "Q"):
            for p in pending:
                out[p] = name
            pending = []
    return out


class Case:
    def __init__(self, gb, fb, seed=0):
        self.gb = tuple(gb)
        self.fb = tuple(fb)
        self.rng = np.random.default_rng(seed)
        self.boxes = {"fb": self.fb, "ib": self.fb}
        self.arrays = {}
        self.scalars = {"dx": 0.1, "dy": 0.125, "dz": 0.15, "dt": 0.01, "cfl": 0.5, "eps": 0.3, "b": 0.7,
                        "vn": 0.8, "use_phi0": 0, "idx": 0, "x": 0.1, "width": 0.3, "width_inner": 0.15,
                        "beta": 0.2, "gamma": 0.4, "cvs": 1}

    def box(self, name):
        return self.boxes.get(name, self.gb)

    def array(self, name, boxname, kind="R"):
        if name not in self.arrays:
            shp = shape_of(self.box(boxname))
            if kind == "M":
                self.arrays[name] = np.zeros(shp, np.uint8)
            else:
                self.arrays[name] = self.rng.standard_normal(shp)
        return self.arrays[name]


def _impl_copy(a, device, real32=False):
    if device is None:
        return np.array(a, copy=True)
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if real32 and t.dtype == torch.float64:
        t = t.to(torch.float32)
    return t.to(device)


def _back(a):
    if isinstance(a, np.ndarray):
        return a
    return a.cpu().numpy()


def run(impl, sym, case, device=None, skip_arrays=(), real32=False):
    """Call impl.<sym> on copies of the case's inputs; return (arrays, outputs).
    real32: real arrays go to the device as float32 (liblsmhip_f32.so)."""
    spec = capi.SPECS[sym]
    boxes = box_of_array(sym)
    arrs = {}
    args = []
    for kind, name in spec:
        if kind in ("R", "M"):
            if name not in arrs:
                arrs[name] = _impl_copy(case.array(name, boxes[name], kind), device, real32 and kind == "R")
            args.append(arrs[name])
        elif kind == "IA":
            if name not in arrs:
                arrs[name] = _impl_copy(case.arrays[name], device)
            args.append(arrs[name])
        elif kind == "IH":
            if name not in arrs:
                arrs[name] = np.array(case.arrays[name], copy=True)
            args.append(arrs[name])
        elif kind in ("B", "Q"):
            args.append(case.box(name))
        elif kind in ("S", "I", "U8"):
            args.append(case.scalars[name])
        else:  # SO, IO
            args.append(case.scalars.get(name))
    out = getattr(impl, sym.rstrip("_"))(*args)
    if device is not None:
        import torch
        torch.cuda.synchronize()
    res = {k: _back(v) for k, v in arrs.items() if k not in skip_arrays}
    return res, out


def assert_same(a, b, sym, tol=0.0, skip=(), out_abs=None):
    """Bitwise (tol == 0) or max-abs comparison of two run() results.
    out_abs: compare the scalar outputs to that absolute tolerance instead (the
    arrays still by tol)."""
    arrs_a, out_a = a
    arrs_b, out_b = b
    for name in arrs_a:
        if name in skip:
            continue
        x, y = arrs_a[name], arrs_b[name]
        if tol == 0.0:
            same = np.array_equal(x.view(np.uint8), y.view(np.uint8)) if x.dtype == y.dtype else False
            if not same:
                bad = np.argwhere(x != y)
                raise AssertionError("%s: array %s differs at %d entries, first %s: %r vs %r" % (
                    sym, name, len(bad), bad[:1].tolist(), x[tuple(bad[0])] if len(bad) else None,
                    y[tuple(bad[0])] if len(bad) else None))
        else:
            d = np.max(np.abs(x.astype(np.float64) - y.astype(np.float64))) if x.size else 0.0
            assert d <= tol, "%s: array %s max-abs diff %g > %g" % (sym, name, d, tol)
    if out_abs is not None:
        oa = out_a if isinstance(out_a, tuple) else (out_a,)
        ob = out_b if isinstance(out_b, tuple) else (out_b,)
        for p, q in zip(oa, ob):
            assert abs(p - q) <= out_abs, "%s: output %r vs %r (|diff| %.3g > %.3g)" % (sym, p, q, abs(p - q), out_abs)
    elif tol == 0.0:
        assert out_a == out_b, "%s: outputs differ %r vs %r" % (sym, out_a, out_b)
    elif out_a is not None:
        oa = out_a if isinstance(out_a, tuple) else (out_a,)
        ob = out_b if isinstance(out_b, tuple) else (out_b,)
        for p, q in zip(oa, ob):
            assert abs(p - q) <= tol * max(1.0, abs(p)), "%s: output %r vs %r" % (sym, p, q)


# ---------------------------------------------------------------------------
# Case builders
# ---------------------------------------------------------------------------

def grid_case(seed=0, gb=(-3, 20, -3, 16, -3, 12), fb=(0, 17, 0, 13, 0, 9), mixed_boxes=True):
    """Random fields on a ghost box with negative lower bounds; when
    mixed_boxes, output and velocity arrays get ghost boxes of their own so
    that each array's addressing is exercised separately."""
    c = Case(gb, fb, seed)
    if mixed_boxes:
        c.boxes["gp"] = (gb[0] + 1, gb[1] - 1, gb[2] + 2, gb[3], gb[4], gb[5] - 1)
        c.boxes["gm"] = (gb[0], gb[1] - 2, gb[2] + 1, gb[3] - 1, gb[4] + 2, gb[5])
        c.boxes["g"] = (gb[0] + 2, gb[1], gb[2], gb[3] - 2, gb[4] + 1, gb[5])
        c.boxes["vel"] = (gb[0] + 1, gb[1], gb[2] + 1, gb[3], gb[4] + 1, gb[5])
        c.boxes["rhs"] = (gb[0] + 2, gb[1] - 1, gb[2] + 1, gb[3] - 2, gb[4], gb[5] - 2)
    rng = c.rng
    # a field with exact zeros and tiny values so that every threshold branch is taken
    phi = rng.standard_normal(shape_of(gb))
    flat = phi.reshape(-1)
    sel = rng.choice(flat.size, flat.size // 20, replace=False)
    flat[sel[: len(sel) // 2]] = 0.0
    flat[sel[len(sel) // 2:]] *= 1e-12
    c.arrays["phi"] = phi
    for v in ("vx", "vy", "vz", "vn"):
        a = rng.standard_normal(shape_of(c.box("vel")))
        f = a.reshape(-1)
        s = rng.choice(f.size, f.size // 10, replace=False)
        f[s[: len(s) // 2]] = 0.0
        f[s[len(s) // 2:]] *= 1e-12
        c.arrays[v] = a
    return c


def grid_case2d(seed=0, gb=(-3, 22, -3, 18), fb=(0, 19, 0, 15), mixed_boxes=True):
    """The 2D counterpart of grid_case: random fields over (i, j) with exact
    zeros and tiny values, output and velocity arrays on ghost boxes of their
    own when mixed_boxes."""
    c = Case(gb, fb, seed)
    if mixed_boxes:
        c.boxes["gp"] = (gb[0] + 1, gb[1] - 1, gb[2] + 2, gb[3])
        c.boxes["gm"] = (gb[0], gb[1] - 2, gb[2] + 1, gb[3] - 1)
        c.boxes["g"] = (gb[0] + 2, gb[1], gb[2], gb[3] - 2)
        c.boxes["vel"] = (gb[0] + 1, gb[1], gb[2] + 1, gb[3])
        c.boxes["rhs"] = (gb[0] + 2, gb[1] - 1, gb[2] + 1, gb[3] - 2)
    rng = c.rng
    phi = rng.standard_normal(shape_of(gb))
    flat = phi.reshape(-1)
    sel = rng.choice(flat.size, flat.size // 20, replace=False)
    flat[sel[: len(sel) // 2]] = 0.0
    flat[sel[len(sel) // 2:]] *= 1e-12
    c.arrays["phi"] = phi
    for v in ("vx", "vy", "vn"):
        a = rng.standard_normal(shape_of(c.box("vel")))
        f = a.reshape(-1)
        s = rng.choice(f.size, f.size // 10, replace=False)
        f[s[: len(s) // 2]] = 0.0
        f[s[len(s) // 2:]] *= 1e-12
        c.arrays[v] = a
    return c


def sphere_phi(gb, centre, radius, h):
    ilo, ihi, jlo, jhi, klo, khi = gb
    k, j, i = np.meshgrid(np.arange(klo, khi + 1), np.arange(jlo, jhi + 1), np.arange(ilo, ihi + 1),
                          indexing="ij")
    x, y, z = (i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h
    return np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2) - radius


def band_case(impl=None, seed=0, n=(24, 22, 20), level=4, h=None, width=None, width_inner=None,
              marks=(124, 125, 126, 127), device=None):
    """A narrow band around two spheres, one of them cut by the domain
    boundary.  The band (levels 0..level) is built with ``impl`` (the oracle
    by default) and the boundary layers are marked as the reference's local
    driver does (examples/curvature_driven_motion/curvature_model3d_local.c:206-229):
    fb planes of the D2 box, then of the D1 box, then of the ghost box."""
    impl = impl or oracle()
    nx, ny, nz = n
    h = h or 1.0 / nx
    gb = (-3, nx + 2, -3, ny + 2, -3, nz + 2)
    fb = (0, nx - 1, 0, ny - 1, 0, nz - 1)
    c = Case(gb, fb, seed)
    c.scalars.update(dx=h, dy=h, dz=h)
    phi = np.minimum(sphere_phi(gb, (0.45, 0.5 * ny * h, 0.5 * nz * h), 0.25, h),
                     sphere_phi(gb, (0.05, 0.1, 0.12), 0.2, h))
    phi += 1e-3 * h * c.rng.standard_normal(phi.shape)
    c.arrays["phi"] = phi
    npts = phi.size
    c.arrays["nb"] = np.zeros(phi.shape, np.uint8)
    for nm in ("ix", "iy", "iz", "outer"):
        c.arrays[nm] = np.full(npts, -7, np.int32)
    c.arrays["n_lo"] = np.full(level + 1, -5, np.int32)
    c.arrays["n_hi"] = np.full(level + 1, -5, np.int32)
    c.boxes.update(nb=gb, gb=gb)
    c.scalars.update(nlo=0, nhi=npts - 1, nlo_outer=0, nhi_outer=npts - 1, level=level,
                     width=width or 6 * h, width_inner=width_inner or 3 * h, beta=2 * h, gamma=4 * h)
    arrs, outs = run(impl, "lsm3ddeterminenarrowband_", c, device=device)
    for k in ("nb", "ix", "iy", "iz", "outer", "n_lo", "n_hi"):
        c.arrays[k] = arrs[k]
    (c.scalars["nlo_outer_plus"], c.scalars["nhi_outer_plus"], c.scalars["nlo_outer_minus"],
     c.scalars["nhi_outer_minus"]) = outs
    mark_fb, mark_d2, mark_d1, mark_gb = marks
    for box, mark in (((-1, nx, -1, ny, -1, nz), mark_d2), ((-2, nx + 1, -2, ny + 1, -2, nz + 1), mark_d1),
                      (gb, mark_gb)):
        c.boxes["fb"] = box
        c.scalars["mark"] = mark
        arrs, _ = run(impl, "lsm3dmarknarrowbandboundarylayer_", c, device=device)
        c.arrays["nb"] = arrs["nb"]
    c.boxes["fb"] = fb
    c.scalars.update(mark_fb=mark_fb, mark_D1=mark_d1, mark_D2=mark_d2)
    nl, nh = c.arrays["n_lo"], c.arrays["n_hi"]
    for L in range(level + 1):
        c.scalars["nlo%d" % L] = int(nl[L])
        c.scalars["nhi%d" % L] = int(nh[L])
    c.scalars["nlo"], c.scalars["nhi"] = int(nl[0]), int(nh[0])
    return c


def circle_phi(gb, centre, radius, h):
    ilo, ihi, jlo, jhi = gb
    j, i = np.meshgrid(np.arange(jlo, jhi + 1), np.arange(ilo, ihi + 1), indexing="ij")
    x, y = (i + 0.5) * h, (j + 0.5) * h
    return np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2) - radius


def band_case2d(impl=None, seed=0, n=(40, 36), level=4, h=None, width=None, width_inner=None,
                marks=(124, 125, 126, 127), device=None):
    """The 2D counterpart of band_case: a band around two circles, one cut by
    the domain boundary, built with ``impl`` (the oracle by default), with the
    boundary lines of the D2, D1 and ghost boxes marked in that order."""
    impl = impl or oracle()
    nx, ny = n
    h = h or 1.0 / nx
    gb = (-3, nx + 2, -3, ny + 2)
    fb = (0, nx - 1, 0, ny - 1)
    c = Case(gb, fb, seed)
    c.scalars.update(dx=h, dy=h)
    phi = np.minimum(circle_phi(gb, (0.45, 0.5 * ny * h), 0.25, h), circle_phi(gb, (0.05, 0.1), 0.2, h))
    phi += 1e-3 * h * c.rng.standard_normal(phi.shape)
    c.arrays["phi"] = phi
    npts = phi.size
    c.arrays["nb"] = np.zeros(phi.shape, np.uint8)
    for nm in ("ix", "iy", "outer"):
        c.arrays[nm] = np.full(npts, -7, np.int32)
    c.arrays["n_lo"] = np.full(level + 1, -5, np.int32)
    c.arrays["n_hi"] = np.full(level + 1, -5, np.int32)
    c.boxes.update(nb=gb, gb=gb)
    c.scalars.update(nlo=0, nhi=npts - 1, nlo_outer=0, nhi_outer=npts - 1, level=level,
                     width=width or 6 * h, width_inner=width_inner or 3 * h, beta=2 * h, gamma=4 * h)
    arrs, outs = run(impl, "lsm2ddeterminenarrowband_", c, device=device)
    for k in ("nb", "ix", "iy", "outer", "n_lo", "n_hi"):
        c.arrays[k] = arrs[k]
    (c.scalars["nlo_outer_plus"], c.scalars["nhi_outer_plus"], c.scalars["nlo_outer_minus"],
     c.scalars["nhi_outer_minus"]) = outs
    mark_fb, mark_d2, mark_d1, mark_gb = marks
    for box, mark in (((-1, nx, -1, ny), mark_d2), ((-2, nx + 1, -2, ny + 1), mark_d1), (gb, mark_gb)):
        c.boxes["fb"] = box
        c.scalars["mark"] = mark
        arrs, _ = run(impl, "lsm2dmarknarrowbandboundarylayer_", c, device=device)
        c.arrays["nb"] = arrs["nb"]
    c.boxes["fb"] = fb
    c.scalars.update(mark_fb=mark_fb, mark_D1=mark_d1, mark_D2=mark_d2, mark_D3=mark_d1)
    nl, nh = c.arrays["n_lo"], c.arrays["n_hi"]
    for L in range(level + 1):
        c.scalars["nlo%d" % L] = int(nl[L])
        c.scalars["nhi%d" % L] = int(nh[L])
    c.scalars["nlo"], c.scalars["nhi"] = int(nl[0]), int(nh[0])
    return c
