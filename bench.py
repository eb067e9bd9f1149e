"""Benchmark of the fused HJ-WENO5 + TVD-RK3 level-set step on one MI355X.

Workload: the 512^3 deformation-advection configuration -- a sphere of radius
0.15 at (0.35, 0.35, 0.35) on [0,1]^3 advected with HJ-WENO5 + TVD-RK3 at CFL
0.5 through the divergence-free deformation field times cos(pi t / 3),
reinitialized every 20 steps (10 HJ-WENO5 + TVD-RK3 pseudo-steps of 0.5 dx).
A "step" is one advection step; the reinitialization blocks due within the
timed steps are inside the timed region.  `value` counts every full-grid
WENO5-RK3 update (advection and reinitialization) of the 512^3 interior.

The headline runs FP32 storage and arithmetic through the fused stage kernels.
Beside it the same workload is timed through the FP64 fused kernels and the
FP64 drop-in chain (the reference's call sequence), and the FP32 result is
compared with the FP64 one (L-inf / L1 of phi and relative interface volume,
FP64 Kahan-compensated).  Inputs are resident in HBM when timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--no-cpu] [--no-compare]
    python bench.py --kernels [--n 512] [--steps K]   # normals, curvatures, orthogonalization,
                                                      # field extension: HBM roofline per entry point
    python bench.py --core [--n 512] [--steps K]      # whole-grid derivatives, RHS terms, RK stages,
                                                      # CFL and integrals: HBM roofline per entry point
    python bench.py --kernels2d [--n 8192]            # 2D entry points and a 2D band
    python bench.py --band [--n 512]                  # narrow-band step: drop-in chain and fused stages
    python bench.py --precision-study [--n 512]       # the deformation run over t in [0, 3], FP32 vs FP64
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "WENO5-RK3 cell-updates/s @512³ + % HBM peak; FP32 vs FP64 at stated L∞ err"
HBM_PEAK_GBS = 8000.0


def _sync():
    import torch
    torch.cuda.synchronize()


def _stream_ordered(lib):
    """Calls of this library go to torch's current stream without waiting for
    each launch (lsmhip_set_stream(stream, 0)); results needed on the host
    (CFL steps, band sizes) still synchronise.  Kernels and ordering are the
    same as in the default synchronous mode."""
    import torch
    lib.set_stream(torch.cuda.current_stream().cuda_stream, False)


def _device_errors(lib):
    """Device-side index errors of stream-ordered calls (raised as in the
    synchronous mode)."""
    if lib.cdll.lsmhip_check_device_errors() != 0:
        lib.check()


def run_timed(run, steps):
    _sync()
    t0 = time.perf_counter()
    updates = 0
    for _ in range(steps):
        updates += run.step()
    _sync()
    return time.perf_counter() - t0, updates


def roofline_from_events(run, bytes_per_scalar):
    """Average duration of the fused advection stage launches (HIP events on
    the launch stream) and their algorithmic bytes: stage 1 reads phi_n and
    writes phi_1 (2 passes), stages 2 and 3 read the stage field and phi_n and
    write the next one (3 passes)."""
    passes = {"advect1": 2, "advect2": 3, "advect3": 3}
    cells = run.grid.cells
    tot_ms, tot_bytes, n = 0.0, 0.0, 0
    for kind, e0, e1 in run.events:
        if kind in passes:
            tot_ms += e0.elapsed_time(e1)
            tot_bytes += passes[kind] * bytes_per_scalar * cells
            n += 1
    avg_ms = tot_ms / max(n, 1)
    achieved = tot_bytes / (tot_ms * 1e-3) / 1e9 if tot_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic(), "kernel": "fused_stage_kernel<advect>",
            "avg_launch_ms": round(avg_ms, 4), "launches": n,
            "alg_bytes_per_cell_update": 8 * bytes_per_scalar}


def _traffic():
    """HBM bytes per launch of the FP32 advection stage kernel from the
    committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration), or None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", "fused_advect_f32_512_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(n=80, steps=2):
    """The reference's own FP64 kernels (oracle/_ref, flang-built) -- or the
    oracle restatement when that build is absent -- driven through the same
    call sequence on the host, one thread, on an n^3 sample of the workload."""
    from lsmlib_amd import capi
    from lsmlib_amd.solver import Grid, deformation_tables, deformation_time_factor, phi_sphere

    ref = os.path.join(ROOT, "oracle", "_ref", "liblsm_ref.so")
    if os.path.exists(ref):
        lib, kind = capi.Bound(ctypes.CDLL(ref)), "reference"
    else:
        lib = capi.Bound(ctypes.CDLL(os.path.join(ROOT, "oracle", "liblsm_oracle.so")), prefix="oracle_",
                         strip_underscore=True)
        kind = "port"
    g = Grid((n, n, n))
    gb, fb = g.gb, g.fb
    phi = phi_sphere(g, (0.35, 0.35, 0.35), 0.15)
    A0, A1, A2, B0, B1, B2 = deformation_tables(g)
    ux = 2 * A0[None, None, :] * B1[None, :, None] * B2[:, None, None]
    uy = -(B0[None, None, :] * A1[None, :, None] * B2[:, None, None])
    uz = -(B0[None, None, :] * B1[None, :, None] * A2[:, None, None])
    grads = [np.zeros(g.shape) for _ in range(3)]
    d1, rhs, p1, p2 = (np.zeros(g.shape) for _ in range(4))
    from lsmlib_amd.solver import advection_cfl_dt
    dt = advection_cfl_dt(g, 0.5, 1.0)

    def fill(a):
        for idx in range(6):
            lib.lsm3dsignedlinearextrapolation(a, gb, fb, idx)

    fill(phi)
    t = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        for st, (src, out, tt) in enumerate(((phi, p1, t), (p1, p2, t + dt), (p2, phi, t + 0.5 * dt)), start=1):
            c = deformation_time_factor(tt)
            vel = (ux * c, uy * c, uz * c)
            lib.lsm3dupwindhjweno5(*grads, gb, src, gb, *vel, gb, d1, gb, fb, *g.dx)
            lib.lsm3dzerooutlevelseteqnrhs(rhs, gb)
            lib.lsm3daddadvectiontermtolserhs(rhs, gb, *grads, gb, *vel, gb, fb)
            if st == 1:
                lib.lsm3dtvdrk3stage1(out, gb, phi, gb, rhs, gb, fb, dt)
            elif st == 2:
                lib.lsm3dtvdrk3stage2(out, gb, src, gb, phi, gb, rhs, gb, fb, dt)
            else:
                lib.lsm3dtvdrk3stage3(out, gb, src, gb, phi, gb, rhs, gb, fb, dt)
            fill(out)
        t += dt
    el = time.perf_counter() - t0
    return {"value": round(g.cells * steps / el, 1), "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": "deformation advection, %d^3, %d HJ-WENO5 + TVD-RK3 steps (FP64, serial), %.1f s" % (
                n, steps, el)}


# Entry points either side of the curvature / reinitialization path, timed
# one by one at n^3: (C-ABI name, real arrays read per cell, real arrays
# written per cell).  Algorithmic bytes = (read + written) x cells x scalar size.
NEIGHBOUR_KERNELS = [
    ("lsm3dcomputeunitnormal", 3, 3),
    ("lsm3dcomputesignedunitnormal", 4, 3),
    ("lsm3dcomputemeancurvature", 9, 1),
    ("lsm3dcomputegaussiancurvatureorder2", 5, 1),
    ("lsm3dcomputemeancurvatureorder4", 5, 1),
    ("lsm3dcomputegaussiancurvatureorder4", 5, 1),
    ("lsm3dcomputeorthogonalizationeqnrhs", 13, 1),
    ("lsm3dcomputefieldextensioneqnrhs", 7, 1),
    ("lsm3dcentralhessian", 1, 6),
]


def kernel_table(n, reps, warmup):
    """Average launch time (HIP events on the launch stream, `reps` launches
    back to back) and HBM roofline fraction of each NEIGHBOUR_KERNELS entry
    point at n^3 (3 ghost cells), FP64 and FP32, on a sphere of radius 0.3:
    phi its signed distance, exact gradient, |grad phi| = 1, the Hessian of
    the distance, psi a plane through the centre, S a smooth field."""
    import torch
    import lsmlib_amd
    rows = []
    # the practical bandwidth reference: one device copy of 1 GiB (reads 1 GiB, writes 1 GiB)
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(warmup):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 2 * src.numel() * 4 / (ms * 1e-3) / 1e9
    rows.append({"kernel": "device copy 1 GiB (torch)", "dtype": "f32", "avg_launch_ms": round(ms, 4),
                 "alg_bytes": 2 * src.numel() * 4, "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
    print("[bench] %-38s %s %8.4f ms %7.1f GB/s %5.1f%% of HBM peak" % (
        "device copy 1 GiB", "f32", ms, gbs, 100 * gbs / HBM_PEAK_GBS), file=sys.stderr, flush=True)
    del src, dst
    g = 3
    gb = (-g, n - 1 + g) * 3
    fb = (0, n - 1) * 3
    h = 1.0 / n
    for prec, dt in (("f64", torch.float64), ("f32", torch.float32)):
        lib = lsmlib_amd.load(prec)
        stream = torch.cuda.current_stream()
        lib.set_stream(stream.cuda_stream, False)
        ax = (torch.arange(-g, n + g, dtype=dt, device="cuda") + 0.5) * h - 0.5
        z, y, x = torch.meshgrid(ax, ax, ax, indexing="ij")
        r = torch.sqrt(x * x + y * y + z * z)
        A = {"phi": r - 0.3, "px": x / r, "py": y / r, "pz": z / r}
        del z, y, x
        A["gmag"] = torch.ones_like(r)
        for key, (p, q) in {"xx": (0, 0), "xy": (0, 1), "xz": (0, 2), "yy": (1, 1), "yz": (1, 2),
                            "zz": (2, 2)}.items():
            d = (A["px"], A["py"], A["pz"])
            A["p" + key] = ((1.0 if p == q else 0.0) - d[p] * d[q]) / r
        A["psi"] = A["py"] * r
        A["S"] = torch.sin(4 * A["phi"]) + A["px"]
        A["w"] = A["S"] * A["pz"]
        del r
        out = [torch.zeros_like(A["phi"]) for _ in range(6)]
        P = A
        calls = {
            "lsm3dcomputeunitnormal": (*out[:3], gb, P["px"], P["py"], P["pz"], gb, fb),
            "lsm3dcomputesignedunitnormal": (*out[:3], gb, P["px"], P["py"], P["pz"], gb, P["phi"], gb, fb, h, h, h),
            "lsm3dcomputemeancurvature": (out[0], gb, P["px"], P["py"], P["pz"], gb, P["pxx"], P["pxy"], P["pxz"],
                                          P["pyy"], P["pyz"], P["pzz"], gb, fb),
            "lsm3dcomputegaussiancurvatureorder2": (out[0], gb, P["phi"], gb, P["px"], P["py"], P["pz"], P["gmag"],
                                                    gb, fb, h, h, h),
            "lsm3dcomputemeancurvatureorder4": (out[0], gb, P["phi"], gb, P["px"], P["py"], P["pz"], P["gmag"], gb,
                                                fb, h, h, h),
            "lsm3dcomputegaussiancurvatureorder4": (out[0], gb, P["phi"], gb, P["px"], P["py"], P["pz"], P["gmag"],
                                                    gb, fb, h, h, h),
            # thirteen distinct input arrays: grad phi plus/minus and grad psi plus/minus stand in
            "lsm3dcomputeorthogonalizationeqnrhs": (out[0], gb, P["px"], P["py"], P["pz"], gb, P["pxx"], P["pyy"],
                                                    P["pzz"], gb, P["psi"], gb, P["pxy"], P["pxz"], P["pyz"], gb,
                                                    P["S"], P["gmag"], P["w"], gb, fb, h, h, h),
            "lsm3dcomputefieldextensioneqnrhs": (out[0], gb, P["S"], gb, P["phi"], gb, P["pxx"], P["pyy"], P["pzz"],
                                                 gb, P["px"], P["py"], P["pz"], gb, fb, h, h, h),
            "lsm3dcentralhessian": (*out, gb, P["phi"], gb, fb, h, h, h),
        }
        cells = n ** 3
        size = 8 if prec == "f64" else 4
        for name, rd, wr in NEIGHBOUR_KERNELS:
            fn, args = getattr(lib.api, name), calls[name]
            for _ in range(warmup):
                fn(*args)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn(*args)
            e1.record(stream)
            e1.synchronize()
            lib.check()
            ms = e0.elapsed_time(e1) / reps
            alg = (rd + wr) * cells * size
            gbs = alg / (ms * 1e-3) / 1e9
            rows.append({"kernel": name, "dtype": prec, "avg_launch_ms": round(ms, 4), "alg_bytes": alg,
                         "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
            print("[bench] %-38s %s %8.4f ms %7.1f GB/s %5.1f%% of HBM peak" % (
                name, prec, ms, gbs, 100 * gbs / HBM_PEAK_GBS), file=sys.stderr, flush=True)
        lib.set_stream(0, True)
        del A, P, out, calls
        torch.cuda.empty_cache()
    return rows


# 2D entry points timed one by one on an n^2 grid: (C-ABI name, real arrays
# read per cell, real arrays written per cell).
KERNELS_2D = [
    ("lsm2dhjweno5", 1, 4),
    ("lsm2dupwindhjweno5", 3, 2),
    ("lsm2dcentralgradorder2", 1, 2),
    ("lsm2dgradientmagnitude", 2, 1),
    ("lsm2daddnormalveltermtolserhs", 6, 1),
    ("lsm2dcomputereinitializationeqnrhs", 5, 1),
    ("lsm2dcomputemeancurvatureorder2", 4, 1),
    ("lsm2dtvdrk3stage2", 3, 1),
]


def kernel_table_2d(n, reps, warmup):
    """Average launch time (HIP events, `reps` back-to-back launches) and HBM
    roofline fraction of each KERNELS_2D entry point on an n^2 grid with 3
    ghost cells (circle of radius 0.3: phi its signed distance, exact
    gradients), FP64 and FP32; then the 2D band of half-width 6h around that
    circle: its construction (levels 0..3) and HJ-WENO5 on the band, whose
    algorithmic bytes per level-0 point are 8 B of index, 1 B of mark, phi and
    four outputs."""
    import torch
    import lsmlib_amd
    rows = []
    g = 3
    gb = (-g, n - 1 + g) * 2
    fb = (0, n - 1) * 2
    h = 1.0 / n
    cells = n * n
    for prec, dt in (("f64", torch.float64), ("f32", torch.float32)):
        lib = lsmlib_amd.load(prec)
        stream = torch.cuda.current_stream()
        lib.set_stream(stream.cuda_stream, False)
        ax = (torch.arange(-g, n + g, dtype=dt, device="cuda") + 0.5) * h - 0.5
        y, x = torch.meshgrid(ax, ax, indexing="ij")
        r = torch.sqrt(x * x + y * y)
        P = {"phi": r - 0.3, "px": x / r, "py": y / r}
        P["gmag"] = torch.ones_like(r)
        P["vn"] = 0.5 + 0.2 * torch.sin(8 * x)
        del y, x, r
        o = [torch.zeros_like(P["phi"]) for _ in range(4)]
        D1 = torch.zeros_like(P["phi"])
        size = 8 if prec == "f64" else 4
        calls = {
            "lsm2dhjweno5": (o[0], o[1], gb, o[2], o[3], gb, P["phi"], gb, D1, gb, fb, h, h),
            "lsm2dupwindhjweno5": (o[0], o[1], gb, P["phi"], gb, P["px"], P["py"], gb, D1, gb, fb, h, h),
            "lsm2dcentralgradorder2": (o[0], o[1], gb, P["phi"], gb, fb, h, h),
            "lsm2dgradientmagnitude": (P["px"], P["py"], o[0], gb, fb),
            "lsm2daddnormalveltermtolserhs": (o[0], gb, P["px"], P["py"], gb, P["gmag"], P["vn"], gb, P["phi"], gb,
                                              fb),
            "lsm2dcomputereinitializationeqnrhs": (o[0], gb, P["phi"], gb, P["phi"], gb, P["px"], P["py"], gb,
                                                   P["gmag"], P["vn"], gb, fb, h, h, 0),
            "lsm2dcomputemeancurvatureorder2": (o[0], gb, P["phi"], gb, P["px"], P["py"], P["gmag"], gb, fb, h, h),
            "lsm2dtvdrk3stage2": (o[0], gb, P["px"], gb, P["phi"], gb, P["py"], gb, fb, 0.25 * h),
        }
        for name, rd, wr in KERNELS_2D:
            fn, args = getattr(lib.api, name), calls[name]
            for _ in range(warmup):
                fn(*args)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn(*args)
            e1.record(stream)
            e1.synchronize()
            _device_errors(lib)
            lib.check()
            ms = e0.elapsed_time(e1) / reps
            alg = (rd + wr) * cells * size
            gbs = alg / (ms * 1e-3) / 1e9
            rows.append({"kernel": name, "dtype": prec, "avg_launch_ms": round(ms, 4), "alg_bytes": alg,
                         "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
            print("[bench] %-38s %s %8.4f ms %7.1f GB/s %5.1f%% of HBM peak" % (
                name, prec, ms, gbs, 100 * gbs / HBM_PEAK_GBS), file=sys.stderr, flush=True)
        # the band: construction, then HJ-WENO5 over it
        phi = P["phi"]
        npts = phi.numel()
        nb = torch.zeros(phi.shape, dtype=torch.uint8, device="cuda")
        ix, iy, outer = (torch.zeros(npts, dtype=torch.int32, device="cuda") for _ in range(3))
        n_lo, n_hi = np.zeros(4, np.int32), np.zeros(4, np.int32)
        _sync()
        t0 = time.perf_counter()
        lib.api.lsm2ddeterminenarrowband(phi, gb, nb, gb, ix, iy, 0, npts - 1, n_lo, n_hi, outer, 0, npts - 1,
                                         None, None, None, None, 6 * h, 3 * h, 3)
        for box, mark in (((-1, n, -1, n), 125), ((-2, n + 1, -2, n + 1), 126), (gb, 127)):
            lib.api.lsm2dmarknarrowbandboundarylayer(nb, gb, box, mark)
        _sync()
        t_band = time.perf_counter() - t0
        lv = [int(v) for pair in zip(n_lo, n_hi) for v in pair]
        args = (o[0], o[1], gb, o[2], o[3], gb, phi, gb, D1, gb, h, h, ix, iy, *lv, nb, gb, 124, 126)
        for _ in range(warmup):
            lib.api.lsm2dhjweno5local(*args)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            lib.api.lsm2dhjweno5local(*args)
        e1.record(stream)
        e1.synchronize()
        _device_errors(lib)
        lib.check()
        ms = e0.elapsed_time(e1) / reps
        lv0 = lv[1] - lv[0] + 1
        alg = lv0 * (8 + 1 + 5 * size)
        gbs = alg / (ms * 1e-3) / 1e9
        rows.append({"kernel": "lsm2dhjweno5local (band half-width 6h, levels 0..3)", "dtype": prec,
                     "band_level0_points": lv0, "band_build_ms": round(1e3 * t_band, 3),
                     "avg_launch_ms": round(ms, 4), "alg_bytes": alg, "achieved": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)})
        print("[bench] 2D band %s: build %.3f ms, %d level-0 points, WENO5 local %.4f ms, %.1f GB/s" % (
            prec, 1e3 * t_band, lv0, ms, gbs), file=sys.stderr, flush=True)
        lib.set_stream(0, True)
        del P, o, D1, phi, nb, ix, iy, outer
        torch.cuda.empty_cache()
    return rows


# (entry point, arrays read, arrays written) per fill-box cell; the reductions
# write one host scalar.  The HJ scratch arrays (D1, D2) are neither read nor
# written by these kernels and are not counted; the surface integral reads the
# gradient only where |phi| < eps (under 1 % of the cells here), so only phi
# is counted for it.
CORE_KERNELS = [
    ("lsm3dhjweno5", 1, 6),
    ("lsm3dhjeno2", 1, 6),
    ("lsm3dupwindhjweno5", 4, 3),
    ("lsm3dcentralgradorder2", 1, 3),
    ("lsm3dlaplacianorder2", 1, 1),
    ("lsm3dgradientmagnitude", 3, 1),
    ("lsm3dzerooutlevelseteqnrhs", 0, 1),
    ("lsm3daddadvectiontermtolserhs", 7, 1),
    ("lsm3daddnormalveltermtolserhs", 8, 1),
    ("lsm3daddconstnormalveltermtolserhs", 7, 1),
    ("lsm3daddexternalandnormalveltermtolserhs", 11, 1),
    ("lsm3dcomputereinitializationeqnrhs", 7, 1),
    ("lsm3dcomputemeancurvatureorder2", 5, 1),
    ("lsm3daddconstcurvtermtolserhs", 10, 1),
    ("lsm3daddconstprecomputedcurvtermtolserhs", 3, 1),
    ("lsm3dtvdrk3stage1", 2, 1),
    ("lsm3dtvdrk3stage2", 3, 1),
    ("lsm3dtvdrk3stage3", 3, 1),
    ("lsm3dcomputestableadvectiondt", 3, 0),
    ("lsm3dcomputestablenormalveldt", 7, 0),
    ("lsm3dmaxnormdiff", 2, 0),
    ("lsm3dvolumeregionphilessthanzero", 1, 0),
    ("lsm3dsurfaceareazerolevelset", 1, 0),
]


def core_table(n, reps, warmup):
    """Average launch time (HIP events on the launch stream, `reps` calls back
    to back) and HBM roofline fraction of each CORE_KERNELS entry point -- the
    whole-grid derivatives, RHS terms, RK stages, CFL and integral reductions
    of the time-stepping path -- at n^3 with 3 ghost cells, FP64 and FP32, on
    a sphere of radius 0.3: phi its signed distance, the plus and minus
    gradients exact and in distinct arrays, a rotating velocity field and a
    varying normal speed.  Every array is distinct, so no read is served by
    another array's cache lines.  The reductions return a host scalar and
    their time includes that read-back."""
    import torch
    import lsmlib_amd
    rows = []
    g = 3
    gb = (-g, n - 1 + g) * 3
    fb = (0, n - 1) * 3
    h = 1.0 / n
    cells = n ** 3
    for prec, dt in (("f64", torch.float64), ("f32", torch.float32)):
        lib = lsmlib_amd.load(prec)
        stream = torch.cuda.current_stream()
        lib.set_stream(stream.cuda_stream, False)
        ax = (torch.arange(-g, n + g, dtype=dt, device="cuda") + 0.5) * h - 0.5
        z, y, x = torch.meshgrid(ax, ax, ax, indexing="ij")
        r = torch.sqrt(x * x + y * y + z * z)
        A = {"phi": r - 0.3, "px": x / r, "py": y / r, "pz": z / r}
        A["vx"], A["vy"], A["vz"] = -y + 0.1, x - 0.2, 0.3 * torch.sin(6 * z)
        A["vn"] = 0.5 + 0.2 * torch.sin(8 * x)
        del z, y, x
        for k in ("px", "py", "pz"):
            A["m" + k] = A[k].clone()
        A["gmag"] = torch.ones_like(r)
        d = (A["px"], A["py"], A["pz"])
        for key, (p, q) in {"xx": (0, 0), "xy": (0, 1), "xz": (0, 2), "yy": (1, 1), "yz": (1, 2),
                            "zz": (2, 2)}.items():
            A["p" + key] = ((1.0 if p == q else 0.0) - d[p] * d[q]) / r
        A["kappa"] = 2.0 / r
        A["rhs0"] = 0.1 * A["vn"]
        del r, d
        o = [torch.zeros_like(A["phi"]) for _ in range(6)]
        D1 = torch.zeros_like(A["phi"])
        P = A
        plus = (P["px"], P["py"], P["pz"], gb)
        minus = (P["mpx"], P["mpy"], P["mpz"], gb)
        vel = (P["vx"], P["vy"], P["vz"], gb)
        calls = {
            "lsm3dhjweno5": (*o[:3], gb, *o[3:], gb, P["phi"], gb, D1, gb, fb, h, h, h),
            "lsm3dhjeno2": (*o[:3], gb, *o[3:], gb, P["phi"], gb, D1, gb, D1, gb, fb, h, h, h),
            "lsm3dupwindhjweno5": (*o[:3], gb, P["phi"], gb, *vel, D1, gb, fb, h, h, h),
            "lsm3dcentralgradorder2": (*o[:3], gb, P["phi"], gb, fb, h, h, h),
            "lsm3dlaplacianorder2": (o[0], gb, P["phi"], gb, fb, h, h, h),
            "lsm3dgradientmagnitude": (P["px"], P["py"], P["pz"], o[0], gb, fb),
            "lsm3dzerooutlevelseteqnrhs": (o[0], gb),
            "lsm3daddadvectiontermtolserhs": (o[0], gb, *plus, *vel, fb),
            "lsm3daddnormalveltermtolserhs": (o[0], gb, *plus, *minus, P["vn"], gb, fb),
            "lsm3daddconstnormalveltermtolserhs": (o[0], gb, *plus, *minus, 0.75, fb),
            "lsm3daddexternalandnormalveltermtolserhs": (o[0], gb, *plus, *minus, P["vn"], *vel, fb),
            "lsm3dcomputereinitializationeqnrhs": (o[0], gb, P["phi"], gb, P["phi"], gb, *plus, *minus, fb, h, h, h,
                                                   0),
            "lsm3dcomputemeancurvatureorder2": (o[0], gb, P["phi"], gb, P["px"], P["py"], P["pz"], P["gmag"], gb, fb,
                                                h, h, h),
            "lsm3daddconstcurvtermtolserhs": (o[0], gb, *plus, P["pxx"], P["pxy"], P["pxz"], P["pyy"], P["pyz"],
                                              P["pzz"], gb, -0.02, fb),
            "lsm3daddconstprecomputedcurvtermtolserhs": (o[0], gb, P["kappa"], gb, P["gmag"], gb, -0.02, fb),
            "lsm3dtvdrk3stage1": (o[0], gb, P["phi"], gb, P["rhs0"], gb, fb, 0.25 * h),
            "lsm3dtvdrk3stage2": (o[0], gb, P["px"], gb, P["phi"], gb, P["rhs0"], gb, fb, 0.25 * h),
            "lsm3dtvdrk3stage3": (o[0], gb, P["py"], gb, P["phi"], gb, P["rhs0"], gb, fb, 0.25 * h),
            "lsm3dcomputestableadvectiondt": (None, *vel, fb, h, h, h, 0.5),
            "lsm3dcomputestablenormalveldt": (None, P["vn"], gb, *plus, *minus, fb, h, h, h, 0.5),
            "lsm3dmaxnormdiff": (None, P["phi"], gb, P["px"], gb, fb),
            "lsm3dvolumeregionphilessthanzero": (None, P["phi"], gb, fb, h, h, h, 1.5 * h),
            "lsm3dsurfaceareazerolevelset": (None, P["phi"], gb, *plus, fb, h, h, h, 1.5 * h),
        }
        size = 8 if prec == "f64" else 4
        for name, rd, wr in CORE_KERNELS:
            fn, args = getattr(lib.api, name), calls[name]
            for _ in range(warmup):
                fn(*args)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn(*args)
            e1.record(stream)
            e1.synchronize()
            _device_errors(lib)
            lib.check()
            ms = e0.elapsed_time(e1) / reps
            alg = (rd + wr) * cells * size
            gbs = alg / (ms * 1e-3) / 1e9
            rows.append({"kernel": name, "dtype": prec, "avg_launch_ms": round(ms, 4), "alg_bytes": alg,
                         "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
            print("[bench] %-42s %s %8.4f ms %7.1f GB/s %5.1f%% of HBM peak" % (
                name, prec, ms, gbs, 100 * gbs / HBM_PEAK_GBS), file=sys.stderr, flush=True)
        lib.set_stream(0, True)
        del A, P, o, D1, plus, minus, vel, calls
        torch.cuda.empty_cache()
    return rows


def band_table(n, steps, warmup):
    """The narrow-band workload (64 spheres at n^3, band 6dx / 3dx, levels
    0..3, local WENO5 + constant normal speed + cut-off + local TVD-RK3, as
    lsmlib_amd.solver.NarrowBandRun) through the drop-in entry points, FP64 and
    FP32: time of one band construction (level 0 compaction, levels 1..3,
    boundary layers) and of a step (CFL, three stages, ghost fills), band
    sizes, and band cell-updates/s = level-0 points x steps / time; the step
    through the drop-in local chain and through the fused band stages; and the
    FP32 fused field against the FP64 fused one at the level-0 points of both
    final bands (the same initial field and steps)."""
    import torch
    from lsmlib_amd.solver import Grid, NarrowBandRun, phi_spheres_min
    g = Grid((n, n, n))
    phi0, _ = phi_spheres_min(g, 3, 64, (0.1, 0.9), (0.03, 0.1))
    rows = []
    finals = {}
    for prec, fused in (("f64", False), ("f32", False), ("f64", True), ("f32", True)):
        run = NarrowBandRun(n, precision=prec, steps=steps, rebuild_every=1 << 30, phi0=phi0, fused=fused)
        _stream_ordered(run.solver.lib)
        _sync()
        t0 = time.perf_counter()
        run.rebuild()
        _sync()
        t_band = time.perf_counter() - t0
        run.nsteps = 1  # keep this band: step() rebuilds only when nsteps % rebuild_every == 0
        for _ in range(warmup):
            run.step()
        _sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            run.step()
        _sync()
        el = time.perf_counter() - t0
        _device_errors(run.solver.lib)
        run.solver.lib.check()
        lv0 = int(run.n_hi[0] - run.n_lo[0] + 1)
        lv3 = int(run.n_hi[run.level] - run.n_lo[0] + 1)
        row = {"workload": "narrow band %d^3, 64 spheres, levels 0..%d" % (n, run.level), "dtype": prec,
               "path": "fused band stages" if fused else "drop-in local chain",
               "band_level0_points": lv0, "band_points_levels_0_3": lv3, "band_build_ms": round(1e3 * t_band, 3),
               "ms_per_step": round(1e3 * el / steps, 3), "band_cell_updates_per_s": round(lv0 * steps / el, 1)}
        if fused:  # final field and level-0 marks, for the FP32 vs FP64 comparison
            finals[prec] = (run.solver.host(run.phi), run.nb.cpu().numpy() == 1)
        rows.append(row)
        print("[bench] band %s %s: build %.2f ms, %.3f ms/step, %d level-0 points, %.3e band cell-updates/s" % (
            prec, row["path"], row["band_build_ms"], row["ms_per_step"], lv0, row["band_cell_updates_per_s"]),
            file=sys.stderr, flush=True)
        del run
        torch.cuda.empty_cache()
    (p64, b64), (p32, b32) = finals["f64"], finals["f32"]
    both = b64 & b32
    d = np.abs(p64 - p32)[both]
    acc = {"steps": warmup + steps, "points": int(both.sum()), "linf_phi_level0": float(d.max()),
           "mean_abs_phi_level0": float(d.mean())}
    print("[bench] band FP32 fused vs FP64 fused after %d steps: L-inf %.3e, mean %.3e" % (
        acc["steps"], acc["linf_phi_level0"], acc["mean_abs_phi_level0"]), file=sys.stderr, flush=True)
    return rows, acc


def adjoint_table(n, steps):
    """Wall time of the shape-matching loss (forward only) and of the loss with
    its gradients (forward with checkpoints every 20 steps, then the reverse
    sweep that recomputes each segment and walks its stages back), n^3, FP32
    and FP64, synchronised around each call."""
    import torch
    from lsmlib_amd.adjoint import ShapeMatching
    out = {}
    for prec in ("f32", "f64"):
        sm = ShapeMatching(n=n, precision=prec, steps=steps)
        sm.value(sm.a, sm.phi0)  # warm-up: loads the kernels, allocates the fields
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = sm.value(sm.a, sm.phi0)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        t0 = time.perf_counter()
        loss2, ga, gphi = sm.value_and_grads()
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        sm.engine.solver.lib.check()
        out[prec] = {"steps": steps, "forward_ms": round(1e3 * tf, 2), "loss_and_gradients_ms": round(1e3 * tb, 2),
                     "gradient_over_forward": round(tb / tf, 2), "loss": loss2,
                     "grad_speed_l2": float(torch.linalg.vector_norm(ga.double()))}
        print("[bench] adjoint %s %d^3 x %d steps: forward %.1f ms, loss + gradients %.1f ms" % (
            prec, n, steps, 1e3 * tf, 1e3 * tb), file=sys.stderr, flush=True)
        del sm, ga, gphi
        torch.cuda.empty_cache()
    return out


def band_traffic_floor(n, line_sizes=(128, 64, 32)):
    """Host-only bound on the HBM bytes of one fused band stage (FP32) on the
    narrow-band workload's band at n^3: the bytes of every distinct cache line
    (or sector) of line_size bytes that the stage touches -- phi_stage on the
    19-point cross, the marks of the 18 neighbours and the point, phi_n, the
    three index lists and the output -- as if each were fetched exactly once.
    The band is level 0 = |phi| < 6 dx and levels 1..3 grown from it along the
    axes, the band NarrowBandRun builds."""
    from lsmlib_amd.solver import Grid, phi_spheres_min
    g = Grid((n, n, n))
    phi, _ = phi_spheres_min(g, 3, 64, (0.1, 0.9), (0.03, 0.1))
    lvl0 = np.abs(phi) < 6.0 / n
    del phi
    p0 = np.flatnonzero(lvl0.reshape(-1))
    del lvl0
    npts = int(p0.size)
    sy, sz = g.shape[2], g.shape[1] * g.shape[2]
    alg = 25 * npts

    def distinct(offsets, elem, line):
        return int(np.unique((offsets * elem) // line).size) * line

    out = {"points": npts, "algorithmic_bytes": alg, "floors": {}}
    for line in line_sizes:
        cross = np.concatenate([p0] + [p0 + s * st for st in (1, sy, sz) for s in (-3, -2, -1, 1, 2, 3)])
        parts = {"phi_stage": distinct(cross, 4, line)}
        del cross
        marks = np.concatenate([p0] + [p0 + s * st for st in (1, sy, sz) for s in (-2, -1, 1, 2, 3)])
        parts["marks"] = distinct(marks, 1, line)
        del marks
        parts["phi_n"] = distinct(p0, 4, line)
        parts["index_lists"] = 3 * distinct(np.arange(npts), 4, line)
        parts["output"] = distinct(p0, 4, line)
        total = sum(parts.values())
        out["floors"][str(line)] = {"bytes": total, "ratio": round(total / alg, 3), "parts": parts}
        print("[bench] %d-byte lines: %.1f MB, %.2fx algorithmic" % (line, total / 1e6, total / alg),
              file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 timings and the accuracy check")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=64)
    ap.add_argument("--kernels", action="store_true",
                    help="time the normal / curvature / orthogonalization / field-extension entry points one by "
                         "one at n^3 (--steps launches each) instead of the level-set step")
    ap.add_argument("--kernels2d", action="store_true",
                    help="time 2D entry points (whole grid, and HJ-WENO5 on a 2D band) one by one on an n^2 grid")
    ap.add_argument("--core", action="store_true",
                    help="time the whole-grid derivative, RHS, RK, CFL and integral entry points one by one at n^3 "
                         "(--steps calls each) instead of the level-set step")
    ap.add_argument("--band", action="store_true",
                    help="time the narrow-band workload (band construction and local steps, FP64 and FP32) at n^3 "
                         "instead of the whole-grid step")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the shape-matching loss and its gradients (adjoint of the normal-motion stages) at "
                         "n^3 over --steps steps (default n 256, steps 200)")
    ap.add_argument("--band-floor", action="store_true",
                    help="host only: the cache-line bound on one fused band stage's HBM bytes at n^3 (no GPU)")
    ap.add_argument("--precision-study", action="store_true",
                    help="integrate the deformation workload over t in [0, 3] in FP64 and FP32 (fused stages) and "
                         "report the FP32 field's L-inf / L1 and volume error against FP64")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one device copy of 1 GiB before timing (the known-size kernel for HBM counters)")
    ap.add_argument("--f32-kernels", type=int, choices=(1, 2), default=1,
                    help="FP32 fused-stage kernel set: 1 (default) a WENO5 window per cell, 2 the terms of each "
                         "position shared between cells (lsmhip_set_fused_f32_kernels)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the level-set path runs on a single MI355X here and is not sharded "
              "across GPUs; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if a.band_floor:
        res = band_traffic_floor(a.n)
        print(json.dumps({"metric": "fused band stage HBM bytes floor @%d^3 (FP32)" % a.n, "data": "synthetic",
                          **res}))
        return

    import torch
    import lsmlib_amd
    from lsmlib_amd.solver import DeformationRun

    lsmlib_amd.load("f32").fn("lsmhip_set_fused_f32_kernels", [ctypes.c_int])(a.f32_kernels)
    if a.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        _sync()
        del src, dst
    if a.kernels:
        rows = kernel_table(a.n, a.steps, a.warmup)
        print(json.dumps({"metric": "drop-in entry point HBM roofline @%d^3" % a.n, "unit": "GB/s",
                          "peak": HBM_PEAK_GBS, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                          "data": "synthetic", "kernels": rows}))
        return
    if a.kernels2d:
        rows = kernel_table_2d(a.n, a.steps, a.warmup)
        print(json.dumps({"metric": "2D drop-in entry point HBM roofline @%d^2" % a.n, "unit": "GB/s",
                          "peak": HBM_PEAK_GBS, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                          "data": "synthetic", "kernels": rows}))
        return
    if a.core:
        rows = core_table(a.n, a.steps, a.warmup)
        print(json.dumps({"metric": "whole-grid time-stepping entry point HBM roofline @%d^3" % a.n,
                          "unit": "GB/s", "peak": HBM_PEAK_GBS, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                          "data": "synthetic", "kernels": rows}))
        return
    if a.band:
        rows, acc = band_table(a.n, a.steps, a.warmup)
        print(json.dumps({"metric": "narrow-band local step @%d^3" % a.n, "unit": "band cell-updates/s",
                          "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "data": "synthetic", "runs": rows,
                          "accuracy_fp32_vs_fp64": acc}))
        return
    if a.adjoint:
        n = a.n if a.n != 512 else 256
        steps = a.steps if a.steps != 40 else 200
        res = adjoint_table(n, steps)
        print(json.dumps({"metric": "adjoint of the level-set evolution @%d^3" % n, "unit": "ms", "n_gpus": 1,
                          "data": "synthetic", "config": {"workload": "shape matching: a sphere r=0.2 under "
                                                          "V = 1 + 0.3 tanh(g), RK3 normal motion, periodic, "
                                                          "checkpoints every 20 steps"}, "runs": res}))
        return
    if a.precision_study:
        from lsmlib_amd.solver import precision_study
        res = precision_study(a.n, 3.0, log=lambda m: print("[bench] precision study " + m, file=sys.stderr,
                                                            flush=True))
        print(json.dumps({"metric": "FP32 vs FP64 over t in [0,3] @%d^3" % a.n, "n_gpus": 1, "data": "synthetic",
                          "config": {"workload": "deformation advection of a sphere r=0.15 at (0.35,0.35,0.35), "
                                                 "HJ-WENO5 + TVD-RK3 at CFL 0.5 (dt of the unscaled field, last "
                                                 "step to t = 3), reinitialization every 20 steps, fused stages"},
                          "result": res}))
        return
    print("[bench] building the %d^3 deformation workload (FP32 fused)" % a.n, file=sys.stderr, flush=True)
    run32 = DeformationRun(a.n, "f32")
    _stream_ordered(run32.solver.lib)
    for _ in range(a.warmup):
        run32.step()
    # warm-up steps advance time; restart the clock of the field from the same state
    run32.events = []
    el, updates = run_timed(run32, a.steps)
    _device_errors(run32.solver.lib)
    run32.solver.lib.check()
    value = run32.grid.cells * updates / el
    roof = roofline_from_events(run32, 4)
    print("[bench] FP32 fused: %.3f ms/step, %.3e cell-updates/s" % (1e3 * el / a.steps, value), file=sys.stderr,
          flush=True)

    out = {"metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": round(1e3 * el / a.steps, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "deformation advection of a sphere r=0.15 at (0.35,0.35,0.35), %d^3 on [0,1]^3, "
                                  "HJ-WENO5 + TVD-RK3 at CFL 0.5 (dt fixed at the CFL step of the unscaled "
                                  "analytic field, evaluated once on the host: no per-step CFL reduction), "
                                  "reinitialization (10 WENO5-RK3 pseudo-steps of 0.5dx) every 20 steps, signed "
                                  "linear extrapolation ghosts" % a.n,
                      "grid": a.n, "full_grid_updates_timed": updates, "parallelism": "single GPU"},
           "roofline": roof}

    if not a.no_compare:
        torch.cuda.empty_cache()
        run64 = DeformationRun(a.n, "f64")
        _stream_ordered(run64.solver.lib)
        for _ in range(a.warmup):
            run64.step()
        el64, up64 = run_timed(run64, a.steps)
        _device_errors(run64.solver.lib)
        run64.solver.lib.check()
        v64 = run64.grid.cells * up64 / el64
        print("[bench] FP64 fused: %.3f ms/step" % (1e3 * el64 / a.steps), file=sys.stderr, flush=True)
        # accuracy: FP32 against FP64 after the same warmup + timed steps from the same initial field
        p32 = run32.grid.interior(run32.solver.host(run32.phi))
        p64 = run64.grid.interior(run64.solver.host(run64.phi))
        eps = 1.5 * run64.grid.dx[0]
        v_32 = run32.solver.volume_phi_lt0(run32.phi, eps)
        v_64 = run64.solver.volume_phi_lt0(run64.phi, eps)
        band = np.abs(p64) < 6 * run64.grid.dx[0]
        acc = {"steps": a.warmup + a.steps, "linf_phi": float(np.abs(p32 - p64).max()),
               "l1_phi": float(np.abs(p32 - p64).mean()),
               "linf_phi_band6dx": float(np.abs(p32 - p64)[band].max()) if band.any() else 0.0,
               "volume_rel_err": abs(v_32 - v_64) / abs(v_64), "volume_f64": v_64, "volume_f32": v_32}
        del p32, p64, band, run64
        torch.cuda.empty_cache()
        chain = DeformationRun(a.n, "f64", dropin=True)
        ksteps = max(1, min(a.steps, 20))
        for _ in range(1):
            chain.step()
        elc, upc = run_timed(chain, ksteps)
        _device_errors(chain.solver.lib)
        chain.solver.lib.check()
        vchain = chain.grid.cells * upc / elc
        print("[bench] FP64 drop-in chain: %.3f ms/step" % (1e3 * elc / ksteps), file=sys.stderr, flush=True)
        del chain
        torch.cuda.empty_cache()
        out["fp64_fused"] = {"value": round(v64, 1), "ms_per_step": round(1e3 * el64 / a.steps, 4)}
        out["fp64_dropin_chain"] = {"value": round(vchain, 1), "ms_per_step": round(1e3 * elc / ksteps, 4),
                                    "steps": ksteps}
        out["speedup_fp32_fused_vs_fp64_dropin"] = round(value / vchain, 2)
        out["speedup_fp32_fused_vs_fp64_fused"] = round(value / v64, 2)
        out["accuracy_fp32_vs_fp64"] = acc

    if not a.no_cpu:
        print("[bench] CPU baseline (reference kernels, 1 thread)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(a.cpu_n, a.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
