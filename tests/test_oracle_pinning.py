"""Pins on the oracle that do not need the reference build: the reference's
own golden tables and assertions, closed forms, and invariants chosen so that
a dropped term, a wrong sign, index or operand order fails one of them.

CPU only.  (tests/test_oracle_vs_reference*.py compare the oracle bitwise with
the reference's own kernels when oracle/_ref is built; tests/golden/ holds
fixtures written by those kernels for when it is not.)
"""
import numpy as np
import pytest

import harness as H
from lsmlib_amd import capi

O = H.oracle


# ---------------------------------------------------------------------------
# Golden tables of the reference's tests/toolbox/test_calculus_toolbox.cc:48-106
# ---------------------------------------------------------------------------
H_EXPECTED = [0, 0, 0, 0, 0, 0, 0, 0, 0.028834442811219, 0.195501109477885, 0.5, 0.804498890522115,
              0.971165557188782, 1, 1, 1, 1, 1, 1, 1]
D_EXPECTED = [0, 0, 0, 0, 0, 0, 0, 0, 0.833333333333334, 2.5, 3.333333333333333, 2.499999999999999,
              0.833333333333331, 0, 0, 0, 0, 0, 0, 0]


def test_heaviside_delta_golden_table():
    c = H.Case((0, 0, 0, 0, 0, 0), (0, 0, 0, 0, 0, 0))
    for i in range(20):
        c.scalars.update(x=-1.0 + i * 0.1, eps=0.3)
        (_, h) = H.run(O(), "lsmheaviside_", c)
        (_, d) = H.run(O(), "lsmdeltafunction_", c)
        assert abs(h - H_EXPECTED[i]) < 1e-6, (i, h)
        assert abs(d - D_EXPECTED[i]) < 1e-6, (i, d)


def test_heaviside_delta_golden_table_hip_host_functions():
    import lsmlib_amd
    lib = lsmlib_amd.load("f64")
    c = H.Case((0, 0, 0, 0, 0, 0), (0, 0, 0, 0, 0, 0))
    for i in range(20):
        c.scalars.update(x=-1.0 + i * 0.1, eps=0.3)
        assert abs(H.run(lib.api, "lsmheaviside_", c)[1] - H_EXPECTED[i]) < 1e-6
        assert abs(H.run(lib.api, "lsmdeltafunction_", c)[1] - D_EXPECTED[i]) < 1e-6


# ---------------------------------------------------------------------------
# The reference's tests/boundary_conditions/test_neumann_bc_3d.cc: after a
# homogeneous Neumann fill of all six faces the face-normal one-sided
# derivative on every boundary face vanishes (tolerance 1e-6, fixture :28-128,
# including its use of j for z at :99).
# ---------------------------------------------------------------------------
def neumann_fixture():
    n = (25, 20, 30)
    g = 3
    gb = (-g, n[0] - 1 + g, -g, n[1] - 1 + g, -g, n[2] - 1 + g)
    fb = (0, n[0] - 1, 0, n[1] - 1, 0, n[2] - 1)
    dx, dy, dz = 1.0 / n[0], 1.0 / n[1], 1.0 / n[2]
    phi = np.zeros(H.shape_of(gb))
    for k in range(n[2]):
        for j in range(n[1]):
            y = (j + 0.5) * dy
            z = (j + 0.5) * dz
            x = (np.arange(n[0]) + 0.5) * dx
            phi[k + g, j + g, g:g + n[0]] = (x - 0.25) ** 2 + 4 * (y - 0.35) ** 2 + 0.25 * (z - 0.45) ** 2 - 0.25
    return gb, fb, (dx, dy, dz), phi


@pytest.mark.parametrize("scheme", ["eno1", "eno2", "eno3", "weno5"])
def test_neumann_zero_normal_derivative(scheme):
    gb, fb, d, phi = neumann_fixture()
    o = O()
    for idx in range(6):
        getattr(o, "lsm3dhomogeneousneumann" + scheme)(phi, gb, fb, idx)
    out = [np.zeros(phi.shape) for _ in range(6)]
    D = [np.zeros(phi.shape) for _ in range(3)]
    fn = getattr(o, "lsm3dhj" + scheme)
    if scheme in ("eno1", "weno5"):
        fn(*out[:3], gb, *out[3:], gb, phi, gb, D[0], gb, fb, *d)
    elif scheme == "eno2":
        fn(*out[:3], gb, *out[3:], gb, phi, gb, D[0], gb, D[1], gb, fb, *d)
    else:
        fn(*out[:3], gb, *out[3:], gb, phi, gb, D[0], gb, D[1], gb, D[2], gb, fb, *d)
    xp, yp, zp, xm, ym, zm = out
    g = 3
    I = lambda a: a[g:-g, g:-g, g:-g]  # noqa: E731
    assert np.abs(I(xm)[:, :, 0]).max() < 1e-6 and np.abs(I(xp)[:, :, -1]).max() < 1e-6
    assert np.abs(I(ym)[:, 0, :]).max() < 1e-6 and np.abs(I(yp)[:, -1, :]).max() < 1e-6
    assert np.abs(I(zm)[0, :, :]).max() < 1e-6 and np.abs(I(zp)[-1, :, :]).max() < 1e-6


def neumann_case2d():
    """The fixture of the reference's tests/boundary_conditions/test_neumann_bc_2d.cc:
    25 x 20 cells of [0,1]^2 (cell centres), 3 ghost cells,
    phi = (x-0.25)^2 + 4 (y-0.35)^2 - 0.25 in the interior (ghosts zero here)."""
    n, g = (25, 20), 3
    c = H.Case((-g, n[0] - 1 + g, -g, n[1] - 1 + g), (0, n[0] - 1, 0, n[1] - 1))
    c.boxes["gb"] = c.gb
    dx, dy = 1.0 / n[0], 1.0 / n[1]
    c.scalars.update(dx=dx, dy=dy)
    phi = np.zeros(H.shape_of(c.gb))
    j, i = np.meshgrid(np.arange(n[1]), np.arange(n[0]), indexing="ij")
    x, y = (i + 0.5) * dx, (j + 0.5) * dy
    phi[g:g + n[1], g:g + n[0]] = (x - 0.25) ** 2 + 4 * (y - 0.35) ** 2 - 0.25
    c.arrays["phi"] = phi
    for nm in ("pxp", "pyp", "pxm", "pym", "D1", "D2", "D3"):
        c.arrays[nm] = np.zeros(H.shape_of(c.gb))
    return c


def neumann_fill_and_derivatives2d(impl, scheme, device=None):
    c = neumann_case2d()
    for idx in range(4):
        c.scalars["idx"] = idx
        arrs, _ = H.run(impl, "lsm2dhomogeneousneumann%s_" % scheme, c, device=device)
        c.arrays["phi"] = arrs["phi"]
    return H.run(impl, "lsm2dhj%s_" % scheme, c, device=device)


@pytest.mark.parametrize("scheme", ["eno1", "eno2", "eno3", "weno5"])
def test_neumann_zero_normal_derivative_2d(scheme):
    """tests/boundary_conditions/test_neumann_bc_2d.cc (:109-600): after a
    homogeneous Neumann fill of the four faces the face-normal one-sided
    derivative vanishes on every boundary line (tolerance 1e-6)."""
    a, _ = neumann_fill_and_derivatives2d(O(), scheme)
    I = (slice(3, -3),) * 2
    assert np.abs(a["pxm"][I][:, 0]).max() < 1e-6 and np.abs(a["pxp"][I][:, -1]).max() < 1e-6
    assert np.abs(a["pym"][I][0, :]).max() < 1e-6 and np.abs(a["pyp"][I][-1, :]).max() < 1e-6


# ---------------------------------------------------------------------------
# Closed forms
# ---------------------------------------------------------------------------
def _grid(n=(14, 12, 10), g=3):
    gb = (-g, n[0] - 1 + g, -g, n[1] - 1 + g, -g, n[2] - 1 + g)
    fb = (0, n[0] - 1, 0, n[1] - 1, 0, n[2] - 1)
    k, j, i = np.meshgrid(np.arange(gb[4], gb[5] + 1), np.arange(gb[2], gb[3] + 1), np.arange(gb[0], gb[1] + 1),
                          indexing="ij")
    return gb, fb, i.astype(float), j.astype(float), k.astype(float)


def _pm(scheme, phi, gb, fb, d):
    out = [np.zeros(phi.shape) for _ in range(6)]
    D = [np.zeros(phi.shape) for _ in range(3)]
    fn = getattr(O(), "lsm3dhj" + scheme)
    if scheme in ("eno1", "weno5"):
        fn(*out[:3], gb, *out[3:], gb, phi, gb, D[0], gb, fb, *d)
    elif scheme == "eno2":
        fn(*out[:3], gb, *out[3:], gb, phi, gb, D[0], gb, D[1], gb, fb, *d)
    else:
        fn(*out[:3], gb, *out[3:], gb, phi, gb, D[0], gb, D[1], gb, D[2], gb, fb, *d)
    return out


def _fb(a, gb, fb):
    return a[fb[4] - gb[4]:fb[5] - gb[4] + 1, fb[2] - gb[2]:fb[3] - gb[2] + 1, fb[0] - gb[0]:fb[1] - gb[0] + 1]


@pytest.mark.parametrize("scheme", ["eno1", "eno2", "eno3", "weno5"])
def test_linear_field_exact_gradient(scheme):
    """Every one-sided derivative of a linear field is its gradient."""
    gb, fb, i, j, k = _grid()
    d = (0.1, 0.2, 0.05)
    grad = (0.7, -1.3, 2.1)
    phi = grad[0] * i * d[0] + grad[1] * j * d[1] + grad[2] * k * d[2] + 0.3
    out = _pm(scheme, phi, gb, fb, d)
    for n, a in enumerate(out):
        assert np.allclose(_fb(a, gb, fb), grad[n % 3], rtol=0, atol=1e-12), (scheme, n)


@pytest.mark.parametrize("scheme", ["eno2", "eno3", "weno5"])
def test_quadratic_field_exact_derivatives(scheme):
    """ENO2/3 and every WENO5 candidate are exact for quadratics, so the
    one-sided derivatives equal the analytic derivative at the cell (the
    plus/minus stencils are biased, but a quadratic's one-sided second-order
    reconstructions are exact at the node)."""
    gb, fb, i, j, k = _grid()
    d = (0.1, 0.1, 0.1)
    x, y, z = i * d[0], j * d[1], k * d[2]
    phi = 1.5 * x * x - 0.5 * y * y + 0.25 * z * z + x * y
    exact = (3.0 * x + y, -1.0 * y + x, 0.5 * z)
    out = _pm(scheme, phi, gb, fb, d)
    for n, a in enumerate(out):
        assert np.allclose(_fb(a, gb, fb), _fb(exact[n % 3], gb, fb), rtol=0, atol=1e-10), (scheme, n)


@pytest.mark.parametrize("scheme,order", [("eno1", 1), ("eno2", 2), ("eno3", 3), ("weno5", 5)])
def test_convergence_order(scheme, order):
    """Observed order of every one-sided derivative on a smooth field without
    critical points, grid spacing halved: ENO-k is k-th order, WENO5 fifth.
    A wrong WENO linear weight or ENO stencil coefficient lowers the order."""
    errs = []
    for n in (12, 24):
        h = 0.5 / n
        gb, fb, i, j, k = _grid((n, n, n))
        x, y, z = i * h, j * h, k * h
        phi = np.exp(0.9 * x - 0.6 * y + 0.4 * z)
        out = _pm(scheme, phi, gb, fb, (h, h, h))
        exact = (0.9 * phi, -0.6 * phi, 0.4 * phi)
        errs.append([np.abs(_fb(a - exact[m % 3], gb, fb)).max() for m, a in enumerate(out)])
    rates = np.log2(np.array(errs[0]) / np.array(errs[1]))
    assert rates.min() > order - 0.35, rates


def test_reflection_maps_plus_to_minus():
    """Mirroring i -> -i maps phi_x_plus to -phi_x_minus of the mirror, bitwise."""
    rng = np.random.default_rng(4)
    gb, fb = (-3, 12, -3, 9, -3, 8), (0, 9, 0, 6, 0, 5)
    phi = rng.standard_normal(H.shape_of(gb))
    mir = phi[:, :, ::-1].copy()
    gbm = (-gb[1], -gb[0], gb[2], gb[3], gb[4], gb[5])
    fbm = (-fb[1], -fb[0], fb[2], fb[3], fb[4], fb[5])
    a = _pm("weno5", phi, gb, fb, (0.1, 0.1, 0.1))
    b = _pm("weno5", mir, gbm, fbm, (0.1, 0.1, 0.1))
    assert np.array_equal(_fb(a[0], gb, fb), -_fb(b[3], gbm, fbm)[:, :, ::-1])
    assert np.array_equal(_fb(a[3], gb, fb), -_fb(b[0], gbm, fbm)[:, :, ::-1])
    # y and z derivatives are untouched by the mirror
    assert np.array_equal(_fb(a[1], gb, fb), _fb(b[1], gbm, fbm)[:, :, ::-1])


def test_godunov_normal_term_on_cone():
    """|x| - 1 in 1D along x: for V > 0 the Godunov upwind picks the outward
    side, rhs = -V exactly away from the kink and 0 at the kink's minimum."""
    gb, fb, i, j, k = _grid((20, 4, 4))
    d = (0.1, 0.1, 0.1)
    phi = np.abs((i - 9.5) * d[0]) - 1.0
    out = _pm("weno5", phi, gb, fb, d)
    rhs = np.zeros(phi.shape)
    O().lsm3daddconstnormalveltermtolserhs(rhs, gb, *out[:3], gb, *out[3:], gb, 2.0, fb)
    r = _fb(rhs, gb, fb)
    assert np.allclose(r[:, :, [0, 1, 2, 17, 18, 19]], -2.0, atol=1e-12)


def _cone(a, sign, n=20):
    """sign * (a|x - c| - r) along x, c halfway between nodes 9 and 10."""
    gb, fb, i, j, k = _grid((n, 4, 4))
    d = (0.1, 0.2, 0.15)
    phi = sign * (a * np.abs((i - 9.5) * d[0]) - 0.42)
    return gb, fb, d, phi


def test_godunov_normal_and_upwind_on_cone():
    """ENO1 derivatives of |x-c| - r are exactly -1 / +1 / 0 (at the kink).
    Godunov's |grad phi| for V>0 is 0 at the two kink nodes (expansion fan)
    and 1 elsewhere; for V<0 it is 1 everywhere.  Upwinding with u picks the
    backward difference for u>0, the forward one for u<0."""
    gb, fb, d, phi = _cone(1.0, 1.0)
    o = O()
    out = _pm("eno1", phi, gb, fb, d)
    for v, expect_kink in ((2.0, 0.0), (-2.0, 2.0)):
        rhs = np.zeros(phi.shape)
        vn = np.full(phi.shape, v)
        o.lsm3daddnormalveltermtolserhs(rhs, gb, *out[:3], gb, *out[3:], gb, vn, gb, fb)
        r = _fb(rhs, gb, fb)
        assert np.allclose(r[:, :, [9, 10]], expect_kink, rtol=1e-14, atol=0), v
        assert np.allclose(np.delete(r, [9, 10], axis=2), -v, rtol=1e-14, atol=0), v
    for u in (1.5, -1.5):
        px, py, pz, D1 = (np.zeros(phi.shape) for _ in range(4))
        vel = np.full(phi.shape, u)
        o.lsm3dupwindhjeno1(px, py, pz, gb, phi, gb, vel, vel, vel, gb, D1, gb, fb, *d)
        want = out[3] if u > 0 else out[0]
        assert np.array_equal(_fb(px, gb, fb), _fb(want, gb, fb)), u
        rhs = np.zeros(phi.shape)
        o.lsm3daddadvectiontermtolserhs(rhs, gb, px, py, pz, gb, vel, vel, vel, gb, fb)
        # -u * phi_x with the upwind value: u>0 sees -1 left of c, 0 at node 10, +1 right
        assert np.allclose(_fb(rhs, gb, fb), -u * _fb(want, gb, fb), rtol=1e-14, atol=0)


@pytest.mark.parametrize("a", [1.0, 2.0])
@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("use_phi0", [0, 1])
def test_reinit_rhs_on_cones(a, sign, use_phi0):
    """On a V (phi < 0 at its minimum) or a tent (phi > 0 at its peak) with
    slope a, Godunov's upwind |grad phi| is a at every node, the kink nodes
    included; reinit RHS is S * (1 - a) with the smoothed sign
    S = phi0 / sqrt(phi0^2 + h^2) (Sussman, Smereka & Osher 1994) for
    use_phi0 = 1 and S = phi / sqrt(phi^2 + |grad phi|^2 h^2) (Peng et al.
    1999) otherwise, h = max(dx, dy, dz)."""
    gb, fb, d, phi = _cone(a, sign)
    h = max(d)
    phi0 = 0.9 * phi
    out = _pm("eno1", phi, gb, fb, d)
    rhs = np.zeros(phi.shape)
    O().lsm3dcomputereinitializationeqnrhs(rhs, gb, phi, gb, phi0, gb, *out[:3], gb, *out[3:], gb, fb, *d,
                                           use_phi0)
    src = phi0 if use_phi0 else phi
    S = src / np.sqrt(src ** 2 + (1.0 if use_phi0 else a * a) * h * h)
    assert np.allclose(_fb(rhs, gb, fb), _fb(S * (1 - a), gb, fb), rtol=1e-13, atol=1e-15)
    if a == 1.0:  # a signed distance is a fixed point
        assert np.abs(_fb(rhs, gb, fb)).max() < 1e-14


def test_ghost_fills_closed_forms():
    """Linear extrapolation reproduces a linear field in the whole ghost box
    (corners included, faces applied in order 0..5); copy extrapolation makes
    the field constant along each face normal; signed linear extrapolation
    never moves a ghost value towards zero."""
    rng = np.random.default_rng(6)
    gb, fb, i, j, k = _grid((9, 7, 6))
    lin = 0.3 * i - 0.7 * j + 0.2 * k + 0.05
    a = np.where(np.isin(i, range(0, 9)) & np.isin(j, range(0, 7)) & np.isin(k, range(0, 6)), lin, 99.0)
    o = O()
    for idx in range(6):
        o.lsm3dlinearextrapolation(a, gb, fb, idx)
    assert np.allclose(a, lin, rtol=0, atol=1e-13)
    c = np.where(np.isin(i, range(0, 9)) & np.isin(j, range(0, 7)) & np.isin(k, range(0, 6)),
                 rng.standard_normal(lin.shape), 99.0)
    for idx in range(6):
        o.lsm3dcopyextrapolation(c, gb, fb, idx)
    ci = np.clip(i, 0, 8).astype(int) + 3
    cj = np.clip(j, 0, 6).astype(int) + 3
    ck = np.clip(k, 0, 5).astype(int) + 3
    assert np.array_equal(c, c[ck, cj, ci])
    s = np.where(np.isin(i, range(0, 9)) & np.isin(j, range(0, 7)) & np.isin(k, range(0, 6)),
                 rng.standard_normal(lin.shape), 99.0)
    inner = s[3:-3, 3:-3, 3:-3].copy()
    for idx in (0, 1):
        o.lsm3dsignedlinearextrapolation(s, gb, fb, idx)
    # along x: ghost values keep the edge sign and grow in magnitude by |edge - inner| per cell
    e0, e1 = inner[:, :, 0], inner[:, :, 1]
    for g in (1, 2, 3):
        assert np.allclose(s[3:-3, 3:-3, 3 - g], e0 + np.sign(e0) * np.abs(e0 - e1) * g, rtol=1e-14, atol=1e-14)
    assert np.all(np.abs(s[3:-3, 3:-3, 0]) >= np.abs(e0))


def test_sphere_curvature_and_volume():
    """Mean curvature (sum of principal curvatures) of a sphere is 2/r and the
    smoothed volume of {phi<0} is 4/3 pi r^3, both to O(dx^2)."""
    n = 48
    h = 1.0 / n
    gb, fb, i, j, k = _grid((n, n, n))
    r = 0.3
    phi = np.sqrt((i * h - 0.5) ** 2 + (j * h - 0.5) ** 2 + (k * h - 0.5) ** 2) - r
    o = O()
    px, py, pz, gm, kap = (np.zeros(phi.shape) for _ in range(5))
    o.lsm3dcentralgradorder2(px, py, pz, gb, phi, gb, fb, h, h, h)
    o.lsm3dgradientmagnitude(px, py, pz, gm, gb, fb)
    o.lsm3dcomputemeancurvatureorder2(kap, gb, phi, gb, px, py, pz, gm, gb, fb, h, h, h)
    shell = np.abs(_fb(phi, gb, fb)) < 0.5 * h
    assert abs(_fb(kap, gb, fb)[shell].mean() - 2 / r) < 0.02 * (2 / r)
    c = H.Case(gb, fb)
    c.arrays["phi"] = phi
    c.scalars.update(dx=h, dy=h, dz=h, eps=1.5 * h)
    _, vol = H.run(o, "lsm3dvolumeregionphilessthanzero_", c)
    assert abs(vol - 4 / 3 * np.pi * r ** 3) < 1e-2 * (4 / 3 * np.pi * r ** 3)


@pytest.mark.parametrize("lam", [-3.0, 1.7])
def test_rk_stages_linear_ode_amplification(lam):
    """For du/dt = lam*u one TVD-RK step multiplies u by the Taylor polynomial
    of exp(z), z = lam*dt, truncated at the scheme's order: RK1 1+z, RK2
    1+z+z^2/2, RK3 1+z+z^2/2+z^3/6.  A wrong stage weight breaks the match."""
    gb, fb = (0, 5, 0, 4, 0, 3), (0, 5, 0, 4, 0, 3)
    rng = np.random.default_rng(3)
    u = 1.0 + rng.random(H.shape_of(gb))
    dt = 0.1
    z = lam * dt
    o = O()
    u1, u2, u3 = (np.zeros(u.shape) for _ in range(3))
    o.lsm3drk1step(u1, gb, u, gb, lam * u, gb, fb, dt)
    assert np.allclose(u1, u * (1 + z), rtol=1e-14, atol=0)
    o.lsm3dtvdrk2stage1(u1, gb, u, gb, lam * u, gb, fb, dt)
    o.lsm3dtvdrk2stage2(u2, gb, u1, gb, u, gb, lam * u1, gb, fb, dt)
    assert np.allclose(u2, u * (1 + z + z * z / 2), rtol=1e-14, atol=0)
    o.lsm3dtvdrk3stage1(u1, gb, u, gb, lam * u, gb, fb, dt)
    o.lsm3dtvdrk3stage2(u2, gb, u1, gb, u, gb, lam * u1, gb, fb, dt)
    o.lsm3dtvdrk3stage3(u3, gb, u2, gb, u, gb, lam * u2, gb, fb, dt)
    assert np.allclose(u3, u * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-14, atol=0)


def test_cfl_dt_closed_form():
    gb, fb = (0, 5, 0, 4, 0, 3), (0, 5, 0, 4, 0, 3)
    c = H.Case(gb, fb)
    for v, val in (("vx", 2.0), ("vy", -1.0), ("vz", 0.5)):
        c.arrays[v] = np.full(H.shape_of(gb), val)
    c.boxes["vel"] = gb
    c.scalars.update(dx=0.1, dy=0.2, dz=0.25, cfl=0.5)
    _, dt = H.run(O(), "lsm3dcomputestableadvectiondt_", c)
    assert abs(dt - 0.5 / (20 + 5 + 2)) < 1e-15


# ---------------------------------------------------------------------------
# Band construction invariants (independent of how the lists were built)
# ---------------------------------------------------------------------------
def test_band_invariants():
    level = 4
    c = H.band_case(level=level)
    phi, nb = c.arrays["phi"], c.arrays["nb"]
    ix, iy, iz = c.arrays["ix"], c.arrays["iy"], c.arrays["iz"]
    n_lo, n_hi = c.arrays["n_lo"], c.arrays["n_hi"]
    gb = c.gb
    width = c.scalars["width"]
    # level 0: exactly the cells with |phi| < width, in (k,j,i) order
    kk, jj, ii = np.nonzero(np.abs(phi) < width)
    assert n_lo[0] == 0 and n_hi[0] == len(ii) - 1
    assert np.array_equal(ix[: len(ii)], ii + gb[0]) and np.array_equal(iz[: len(ii)], kk + gb[4])
    seen = set(zip(ix[: n_hi[level] + 1], iy[: n_hi[level] + 1], iz[: n_hi[level] + 1]))
    assert len(seen) == n_hi[level] + 1  # no duplicates
    prev = set(zip(ix[: len(ii)], iy[: len(ii)], iz[: len(ii)]))
    claimed = set(prev)
    for L in range(1, level + 1):
        lvl = list(zip(ix[n_lo[L]:n_hi[L] + 1], iy[n_lo[L]:n_hi[L] + 1], iz[n_lo[L]:n_hi[L] + 1]))
        # every new point neighbours a parent, and every unclaimed neighbour of a parent is new
        nbrs = set()
        for (i, j, k) in prev:
            for (a, b, e) in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                q = (i + a, j + b, k + e)
                if gb[0] <= q[0] <= gb[1] and gb[2] <= q[1] <= gb[3] and gb[4] <= q[2] <= gb[5] and q not in claimed:
                    nbrs.add(q)
        assert set(lvl) == nbrs, L
        # first-claim order: sorted by (position of first parent in list, direction)
        order = {}
        plist = list(zip(ix[n_lo[L - 1]:n_hi[L - 1] + 1], iy[n_lo[L - 1]:n_hi[L - 1] + 1],
                         iz[n_lo[L - 1]:n_hi[L - 1] + 1]))
        for m, (i, j, k) in enumerate(plist):
            for dno, (a, b, e) in enumerate(((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))):
                q = (i + a, j + b, k + e)
                if q in nbrs and q not in order:
                    order[q] = 6 * m + dno
        assert [order[q] for q in lvl] == sorted(order[q] for q in lvl), L
        claimed |= set(lvl)
        prev = set(lvl)
    # marks: level L points carry L + 1, unbanded cells 0 (boundary layers re-mark the box edges)
    m = c.arrays["nb"].astype(int)
    for L in range(level + 1):
        sl = slice(n_lo[L], n_hi[L] + 1)
        pts = (iz[sl] - gb[4], iy[sl] - gb[2], ix[sl] - gb[0])
        assert np.all((m[pts] == L + 1) | (m[pts] >= c.scalars["mark_fb"])), L
    # outer layer: level-0 points with |phi| >= width_inner, phi <= 0 listed from the
    # front in list order, phi > 0 from the back in reverse list order
    inner_w = c.scalars["width_inner"]
    outer = c.arrays["outer"]
    lo_m, hi_m = c.scalars["nlo_outer_minus"], c.scalars["nhi_outer_minus"]
    lo_p, hi_p = c.scalars["nlo_outer_plus"], c.scalars["nhi_outer_plus"]
    vals = phi[kk, jj, ii]
    want_minus = np.nonzero((np.abs(vals) >= inner_w) & (vals <= 0))[0]
    want_plus = np.nonzero((np.abs(vals) >= inner_w) & (vals > 0))[0]
    assert lo_m == 0 and np.array_equal(outer[lo_m:hi_m + 1], want_minus)
    assert hi_p == outer.size - 1 and np.array_equal(outer[lo_p:hi_p + 1], want_plus[::-1])


def test_band_ties():
    """Exact ties of the level-0 and outer-layer tests: |phi| == width is out
    of the band, |phi| == width_inner is in the outer layer, phi == 0 counts
    as minus (lsm_localization3d.f:174-297)."""
    gb = (0, 5, 0, 1, 0, 0)
    vals = np.array([[0.0, 0.5, -0.5, 1.0, -1.0, 0.25], [-0.25, 0.75, -0.75, 0.0, 2.0, -0.0]])
    phi = vals.reshape(1, 2, 6).copy()
    npts = phi.size
    nb = np.zeros(phi.shape, np.uint8)
    ix, iy, iz, outer = (np.full(npts, -7, np.int32) for _ in range(4))
    n_lo, n_hi = np.full(1, -5, np.int32), np.full(1, -5, np.int32)
    r = O().lsm3ddeterminenarrowband(phi, gb, nb, gb, ix, iy, iz, 0, npts - 1, n_lo, n_hi, outer, 0, npts - 1,
                                     None, None, None, None, 1.0, 0.5, 0)
    nlo_p, nhi_p, nlo_m, nhi_m = r
    flat = vals.reshape(-1)
    band = np.nonzero(np.abs(flat) < 1.0)[0]
    assert n_hi[0] == band.size - 1 and np.array_equal(ix[:band.size] + 6 * iy[:band.size], band)
    b = flat[band]
    assert np.array_equal(outer[nlo_m:nhi_m + 1], np.nonzero((np.abs(b) >= 0.5) & (b <= 0))[0])
    assert np.array_equal(outer[nlo_p:nhi_p + 1], np.nonzero((np.abs(b) >= 0.5) & (b > 0))[0][::-1])
    assert nhi_m - nlo_m + 1 == 2 and nhi_p - nlo_p + 1 == 2
    # width_inner 0: every band point is outer; the two zeros (+0.0, -0.0) go minus
    nb[:] = 0
    r = O().lsm3ddeterminenarrowband(phi, gb, nb, gb, ix, iy, iz, 0, npts - 1, n_lo, n_hi, outer, 0, npts - 1,
                                     None, None, None, None, 1.0, 0.0, 0)
    nlo_p, nhi_p, nlo_m, nhi_m = r
    assert np.array_equal(outer[nlo_m:nhi_m + 1], np.nonzero(b <= 0)[0])
    assert np.array_equal(outer[nlo_p:nhi_p + 1], np.nonzero(b > 0)[0][::-1])


def _full_band_case():
    """Every ghost-box cell at level 0 (width beyond the domain), boundary
    layers marked as usual, higher levels empty (their ranges end at nhi0)."""
    c = H.band_case(level=1, width=100.0, width_inner=50.0)
    c.arrays["mask"] = c.rng.standard_normal(c.arrays["phi"].shape)
    c.boxes["mask"] = c.gb
    h = c.scalars["dx"]
    c.scalars.update(dy=1.3 * h, dz=0.7 * h)
    n0 = c.scalars["nhi0"]
    assert c.scalars["nlo0"] == 0 and n0 + 1 == c.arrays["phi"].size
    for L in range(1, 4):
        c.scalars["nlo%d" % L], c.scalars["nhi%d" % L] = n0 + 1, n0
    return c


TWINS = [s for s in capi.SPECS
         if s.startswith("lsm3d") and s.endswith("local_") and s.replace("local_", "_") in capi.SPECS]
TWINS_2D = [s for s in capi.SPECS
            if s.startswith("lsm2d") and s.endswith("local_") and s.replace("local_", "_") in capi.SPECS]


WHOLE_TWIN = {"lsm3dcomputemeancurvatureorder4local_": "lsm3dcomputegaussiancurvatureorder4_"}


@pytest.mark.parametrize("sym", TWINS)
def test_local_equals_whole_grid_on_full_band(sym):
    """With the band covering the grid, each narrow-band routine computes the
    whole-grid routine on the fill box, bitwise; scalar results (dt, norms)
    agree bitwise.  Exceptions that the reference's local routines share:
    WENO5 needs D1 three cells out, where the outermost ghost layer holds the
    'big' value, so it is compared three cells inside the fill box; the local
    gradient magnitude cuts squared components below zero_tol (1e-11) where
    the whole-grid one cuts below 1e-8, so it is compared where every squared
    component is above 1e-8; the local fourth-order "mean" curvature
    (lsm_curvature3d_local.f.in:285-426) evaluates the Gaussian curvature, so
    its whole-grid twin is the fourth-order Gaussian routine."""
    c = _full_band_case()
    o = O()
    al, ol = H.run(o, sym, c)
    aw, ow = H.run(o, WHOLE_TWIN.get(sym, sym.replace("local_", "_")), c)
    assert ol == ow
    inner = (slice(3, -3),) * 3
    if sym == "lsm3dhjweno5local_":
        inner = (slice(6, -6),) * 3
    keep = np.ones(c.arrays["phi"][inner].shape, bool)
    if sym == "lsm3dgradientmagnitudelocal_":
        keep = np.all([al[n][inner] ** 2 >= 1e-8 for n in ("px", "py", "pz")], axis=0)
        assert keep.mean() > 0.9
    compared = 0
    for name, a in aw.items():
        if name in al and a.dtype.kind == "f":
            assert np.array_equal(al[name][inner][keep], a[inner][keep]), name
            compared += 1
    assert compared or ol is not None


def _full_band_case2d():
    """The 2D counterpart of _full_band_case: every ghost-box cell at level 0,
    boundary lines marked as usual, dy != dx."""
    c = H.band_case2d(level=1, width=100.0, width_inner=50.0)
    c.arrays["mask"] = c.rng.standard_normal(c.arrays["phi"].shape)
    c.boxes["mask"] = c.gb
    c.scalars["dy"] = 1.3 * c.scalars["dx"]
    n0 = c.scalars["nhi0"]
    assert c.scalars["nlo0"] == 0 and n0 + 1 == c.arrays["phi"].size
    for L in range(1, 4):
        c.scalars["nlo%d" % L], c.scalars["nhi%d" % L] = n0 + 1, n0
    return c


@pytest.mark.parametrize("sym", TWINS_2D)
def test_local_2d_equals_whole_grid_on_full_band(sym):
    """The 2D narrow-band routines on a band covering the grid compute the 2D
    whole-grid routines on the fill box, bitwise (scalars too).  HJ-WENO5 and
    HJ-ENO3 read differences three cells out, where the outermost ghost line
    holds the 'big' value, so they are compared three cells inside; the
    orders-4 curvature reads phi_x two rows away and takes the ghost box as
    its gradient box."""
    c = _full_band_case2d()
    if sym == "lsm2dcomputemeancurvatureorder4local_":
        c.boxes["g"] = c.gb
    o = O()
    al, ol = H.run(o, sym, c)
    aw, ow = H.run(o, sym.replace("local_", "_"), c)
    assert ol == ow
    inner = (slice(3, -3),) * 2
    if sym in ("lsm2dhjweno5local_", "lsm2dhjeno3local_"):
        inner = (slice(6, -6),) * 2
    compared = 0
    for name, a in aw.items():
        if name in al and a.dtype.kind == "f" and name not in ("D1", "D2", "D3"):
            assert np.array_equal(al[name][inner], a[inner]), name
            compared += 1
    assert compared or ol is not None


@pytest.mark.parametrize("sym", ["lsm3drk1steplocal_", "lsm3dlaplacianorder2local_", "lsm3dhjeno1local_"])
def test_mark_fb_threshold_is_inclusive(sym):
    """A band point takes part iff its mark is <= mark_fb (signed byte
    compare): a point marked exactly mark_fb is computed, one marked
    mark_fb + 1 keeps its old value."""
    c = _full_band_case()
    nb = c.arrays["nb"]
    mf = c.scalars["mark_fb"]
    nb[10, 10, 10] = mf
    nb[10, 10, 12] = mf + 1
    out, _ = H.run(O(), sym, c)
    whole, _ = H.run(O(), sym.replace("local_", "_"), c)
    name = next(n for n in ("un", "lap", "pxp") if n in out)
    before = c.array(name, H.box_of_array(sym)[name], "R")
    assert out[name][10, 10, 10] == whole[name][10, 10, 10] != before[10, 10, 10]
    assert out[name][10, 10, 12] == before[10, 10, 12]


def test_average_grad_phi_local_linear_field():
    """Central differences are exact on a linear field: the band average of
    |grad phi| is |g|, over exactly the points at or below mark_fb."""
    c = H.band_case(level=2)
    gb = c.gb
    k, j, i = np.meshgrid(*(np.arange(gb[2 * a], gb[2 * a + 1] + 1) for a in (2, 1, 0)), indexing="ij")
    g = np.array([0.3, -0.4, 1.2])
    h = c.scalars["dx"]
    c.arrays["phi"] = g[0] * i * h + g[1] * j * h + g[2] * k * h
    _, ave = H.run(O(), "lsm3dcomputeavegradphilocal_", c)
    assert abs(ave - np.linalg.norm(g)) < 1e-12
    # a point outside the fill box with a steep value must not enter the average
    c.arrays["phi"][0, 0, 0] = 1e6
    _, ave2 = H.run(O(), "lsm3dcomputeavegradphilocal_", c)
    assert ave2 == ave


def test_weno5_local_equals_whole_grid_inside_band():
    """The authored 3D WENO5 local routine equals the whole-grid routine at
    level-0 points whose stencil lies in levels 0..3 with marks <= mark_D1."""
    c = H.band_case(level=4)
    o = O()
    res_l, _ = H.run(o, "lsm3dhjweno5local_", c)
    res_g, _ = H.run(o, "lsm3dhjweno5_", c)
    nb = c.arrays["nb"].view(np.int8)
    ix, iy, iz = c.arrays["ix"], c.arrays["iy"], c.arrays["iz"]
    gb = c.gb
    checked = 0
    for l in range(c.scalars["nlo0"], c.scalars["nhi0"] + 1):
        i, j, k = ix[l] - gb[0], iy[l] - gb[2], iz[l] - gb[4]
        if nb[k, j, i] > c.scalars["mark_fb"]:
            continue
        lo = (max(i - 3, 0), max(j - 3, 0), max(k - 3, 0))
        win = [nb[k, j, max(i - 3, 0):i + 4], nb[k, max(j - 3, 0):j + 4, i], nb[max(k - 3, 0):k + 4, j, i]]
        if min(i, j, k) < 3 or any((w <= 0).any() or (w > c.scalars["mark_D1"]).any() for w in win):
            continue
        del lo
        for name in ("pxp", "pyp", "pzp", "pxm", "pym", "pzm"):
            assert res_l[name][k, j, i] == res_g[name][k, j, i], (name, i, j, k)
        checked += 1
    assert checked > 100


def test_periodic_fill_copies_opposite_interior():
    rng = np.random.default_rng(2)
    gb, fb = (-3, 12, -3, 10, -3, 9), (0, 9, 0, 7, 0, 6)
    phi = rng.standard_normal(H.shape_of(gb))
    for idx in range(6):
        O().lsmhip_periodic_face(phi, gb, fb, idx)
    # periodic extension: phi(i) == phi(i mod n) for every ghost-box index
    n = (10, 8, 7)
    k, j, i = np.meshgrid(np.arange(gb[4], gb[5] + 1), np.arange(gb[2], gb[3] + 1), np.arange(gb[0], gb[1] + 1),
                          indexing="ij")
    src = phi[(k % n[2]) + 3, (j % n[1]) + 3, (i % n[0]) + 3]
    assert np.array_equal(phi, src)


# ---------------------------------------------------------------------------
# Normals, curvatures, orthogonalization and field extension
# ---------------------------------------------------------------------------
def _sphere(n, d, r=0.3):
    """phi = |x - c| - r on cells of spacing d (anisotropic), with its exact
    gradient and Hessian; returns (gb, fb, phi, grad, hess, radius)."""
    gb, fb, i, j, k = _grid((n, n, n))
    x, y, z = (i + 0.5) * d[0] - 0.5, (j + 0.5) * d[1] - 0.5, (k + 0.5) * d[2] - 0.5
    rr = np.sqrt(x * x + y * y + z * z)
    X = (x, y, z)
    grad = [a / rr for a in X]
    hess = {}
    for p, q in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)):
        hess[p, q] = ((1.0 if p == q else 0.0) - X[p] * X[q] / rr ** 2) / rr
    return gb, fb, rr - r, grad, hess, rr


def _curvatures(n, d):
    gb, fb, phi, (px, py, pz), _, rr = _sphere(n, d)
    gm = np.ones(phi.shape)
    out = {}
    for name in ("meancurvatureorder4", "gaussiancurvatureorder4", "gaussiancurvatureorder2"):
        kap = np.zeros(phi.shape)
        getattr(O(), "lsm3dcompute" + name)(kap, gb, phi, gb, px, py, pz, gm, gb, fb, *d)
        near = np.abs(_fb(phi, gb, fb)) < 2 * max(d)
        want = (2 / _fb(rr, gb, fb) if name.startswith("mean") else _fb(rr, gb, fb) ** -2.0)[near]
        out[name] = np.abs(_fb(kap, gb, fb)[near] / want - 1).max()
    return out


def test_curvature_orders_on_sphere_anisotropic():
    """Fourth-order mean and Gaussian curvature and second-order Gaussian
    curvature of a sphere with exact first derivatives on a grid with three
    different spacings: the error against the level set through each cell
    (2/r, 1/r^2) falls at rate >= 3.6 (order 4) and >= 1.8 (order 2) from
    n = 32 to 64.  A spacing applied to the wrong axis, a wrong weight or a
    mixed derivative taken along the wrong axis leaves an O(1) error."""
    d = lambda n: (1.0 / n, 1.2 / n, 0.85 / n)  # noqa: E731
    e1, e2 = _curvatures(32, d(32)), _curvatures(64, d(64))
    for name, order in (("meancurvatureorder4", 3.6), ("gaussiancurvatureorder4", 3.6),
                        ("gaussiancurvatureorder2", 1.8)):
        rate = np.log2(e1[name] / e2[name])
        assert e2[name] < 1e-2 and rate >= order, (name, e1[name], e2[name], rate)


def test_mean_curvature_from_exact_hessian():
    """lsm3dComputeMeanCurvature with the exact gradient and Hessian of a
    sphere gives 2/r to rounding; with the gradient scaled by 7 (|grad|^3 in
    the denominator, |grad|^2 in the numerator) it gives 2/(7 r)."""
    gb, fb, phi, g, hs, rr = _sphere(20, (0.05, 0.05, 0.05))
    o = O()
    for scale in (1.0, 7.0):
        kap = np.zeros(phi.shape)
        o.lsm3dcomputemeancurvature(kap, gb, *[a * scale for a in g], gb, hs[0, 0], hs[0, 1], hs[0, 2], hs[1, 1],
                                    hs[1, 2], hs[2, 2], gb, fb)
        assert np.abs(_fb(kap * rr * scale, gb, fb) / 2 - 1).max() < 1e-12


def test_unit_normals_closed_forms():
    """Unit normal = grad/|grad| (and (1,0,0) where grad = 0); the signed unit
    normal is sgn(phi) grad/|grad| with the smoothed sign
    phi/sqrt(phi^2 + |grad|^2 max(d)^2): -> +-1 far from phi = 0, 0 at phi = 0."""
    d = (0.05, 0.07, 0.04)
    gb, fb, phi, g, _, _ = _sphere(20, d)
    g = [a * 3.0 for a in g]  # |grad| = 3
    g[0].reshape(-1)[::7] = 0.0
    g[1].reshape(-1)[::7] = 0.0
    g[2].reshape(-1)[::7] = 0.0
    phi.reshape(-1)[3::11] = 0.0
    o = O()
    n = [np.zeros(phi.shape) for _ in range(3)]
    o.lsm3dcomputeunitnormal(*n, gb, *g, gb, fb)
    zero = _fb(g[0] == 0, gb, fb)
    norm = np.sqrt(sum(_fb(a, gb, fb) ** 2 for a in g))
    for a, ga, e in zip(n, g, (1.0, 0.0, 0.0)):
        assert np.all(_fb(a, gb, fb)[zero] == e)
        assert np.abs(_fb(a, gb, fb)[~zero] - _fb(ga, gb, fb)[~zero] / norm[~zero]).max() < 1e-15
    o.lsm3dcomputesignedunitnormal(*n, gb, *g, gb, phi, gb, fb, *d)
    P = _fb(phi, gb, fb)
    at_zero = P == 0
    A = [_fb(a, gb, fb) for a in n]
    for a in A:
        assert np.all(a[at_zero] == 0)
    ok = ~zero & ~at_zero
    dot = sum(a * _fb(ga, gb, fb) for a, ga in zip(A, g)) / 3.0  # n . grad/|grad|
    length = np.sqrt(sum(a * a for a in A))
    assert np.all(np.sign(dot[ok]) == np.sign(P[ok]))      # points along sgn(phi) grad phi
    assert np.abs(length[ok] - np.abs(dot[ok])).max() < 1e-15  # and along grad phi only
    assert np.all(np.abs(dot[ok]) < 1) and np.all(np.abs(dot[ok & (np.abs(P) > 0.3)]) > 0.8)
    s = P / np.sqrt(P * P + 9.0 * max(d) ** 2)  # smoothed sign at |grad| = 3
    assert np.abs(dot[ok] - s[ok]).max() < 1e-14


def test_orthogonalization_rhs_linear_fields():
    """For linear phi and psi (one-sided derivatives equal), the RHS is
    -sgn(psi) (grad psi/|grad psi|) . grad phi: 0 for orthogonal gradients;
    with different plus/minus derivatives the upwind side follows the sign of
    sgn(psi) grad psi per component."""
    d = (0.1, 0.12, 0.09)
    gb, fb, i, j, k = _grid((8, 7, 6))
    rng = np.random.default_rng(4)
    o = O()
    for trial in range(4):
        a, b = rng.standard_normal(3), rng.standard_normal(3)
        if trial == 0:
            b = b - a * (a @ b) / (a @ a)  # orthogonal
        psi = 0.2 + b[0] * i * d[0] + b[1] * j * d[1] + b[2] * k * d[2]
        P = [np.full(psi.shape, v) for v in a]
        Q = [np.full(psi.shape, v) for v in b]
        rhs = np.zeros(psi.shape)
        o.lsm3dcomputeorthogonalizationeqnrhs(rhs, gb, *P, gb, *P, gb, psi, gb, *Q, gb, *Q, gb, fb, *d)
        S = _fb(psi, gb, fb)
        s = S / np.sqrt(S * S + max(d) ** 2 * (b @ b))
        want = -s * (a @ b) / np.sqrt(b @ b)
        assert np.abs(_fb(rhs, gb, fb) - want).max() < 1e-12 * max(1.0, np.abs(want).max())
    # upwinding: grad psi = (1,0,0); phi_x_minus = 1, phi_x_plus = 100
    psi = np.where(i < 4, -0.5, 0.5)
    one, zero = np.ones(psi.shape), np.zeros(psi.shape)
    rhs = np.zeros(psi.shape)
    o.lsm3dcomputeorthogonalizationeqnrhs(rhs, gb, 100 * one, zero, zero, gb, one, zero, zero, gb, psi, gb,
                                          one, zero, zero, gb, one, zero, zero, gb, fb, *d)
    S = _fb(psi, gb, fb)
    s = S / np.sqrt(S * S + max(d) ** 2)
    assert np.allclose(_fb(rhs, gb, fb), np.where(S > 0, -s * 1.0, -s * 100.0), rtol=1e-14, atol=0)


def test_field_extension_rhs_on_sphere():
    """S = r (distance from the centre), N the exact signed normal of the
    sphere phi = r - R: the RHS is -N . grad S = -sign(phi) wherever phi has
    the sign of its six face neighbours or |phi| > 3 max(d), and exactly 0 in
    the cells the zero level set crosses."""
    d = (0.05, 0.06, 0.045)
    gb, fb, phi, g, _, rr = _sphere(20, d)
    sg = np.sign(phi)
    N = [sg * a for a in g]
    rhs = np.zeros(phi.shape)
    O().lsm3dcomputefieldextensioneqnrhs(rhs, gb, rr, gb, phi, gb, *g, gb, *N, gb, fb, *d)
    P, R_ = _fb(phi, gb, fb), _fb(rhs, gb, fb)
    lo = [f - b for f, b in zip(fb[::2], gb[::2])]
    same = np.ones(P.shape, bool)
    for ax in range(3):
        for s in (-1, 1):
            sl = [slice(lo[2 - a], lo[2 - a] + P.shape[a]) for a in range(3)]
            sl[2 - ax] = slice(lo[ax] + s, lo[ax] + s + P.shape[2 - ax])
            same &= P * phi[tuple(sl)] > 0
    away = (np.abs(P) > 3 * max(d)) | same
    assert np.all(R_[~away] == 0)
    assert np.abs(R_[away] + np.sign(P[away])).max() < 1e-14
    assert (~away).sum() > 100 and away.sum() > 1000
