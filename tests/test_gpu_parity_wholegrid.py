"""HIP drop-in (liblsmhip.so, FP64) against the oracle, bitwise, on the same
seeded inputs -- through the C-ABI with device arrays."""
import numpy as np
import pytest

import harness as H
from test_oracle_vs_reference import BCS, CONTROL_VOLUME, QUIRKY_BOUNDS, UTILS, WHOLE_GRID, _prepare

pytestmark = pytest.mark.gpu

SUMS = {"lsm3dvolumeintegralphilessthanzero_", "lsm3dvolumeintegralphigreaterthanzero_", "lsm3dsurfaceintegral_",
        "lsm3dvolumeregionphilessthanzero_", "lsm3dvolumeregionphigreaterthanzero_", "lsm3dsurfaceareazerolevelset_",
        "lsm3daveabsdiff_", "lsm3dsurfaceareazerolevelsetdelta_",
        "lsm3dvolumeintegralphilessthanzerocontrolvolume_", "lsm3dvolumeintegralphigreaterthanzerocontrolvolume_",
        "lsm3dsurfaceintegralcontrolvolume_", "lsm3dvolumeregionphilessthanzerocontrolvolume_",
        "lsm3dvolumeregionphigreaterthanzerocontrolvolume_", "lsm3dsurfaceareazerolevelsetcontrolvolume_",
        "lsm3dsurfaceareazerolevelsetdeltacontrolvolume_"}


@pytest.fixture(scope="module")
def hip():
    import lsmlib_amd
    lib = lsmlib_amd.load("f64")
    lib.set_stream(0, True)
    return lib


def _cmp(hip, sym, c):
    a = H.run(H.oracle(), sym, c)
    b = H.run(hip.api, sym, c, device="cuda")
    hip.check()
    if sym in SUMS:
        # FP64 compensated parallel sum vs the reference's serial sum: 1e-13 absolute
        # (tests/test_gpu_integrals.py holds the parallel sum to 2 ulp of the exact sum)
        H.assert_same(a, b, sym, out_abs=1e-13)
    else:
        H.assert_same(a, b, sym, skip=("D1", "D2", "D3"))


@pytest.mark.parametrize("sym", WHOLE_GRID)
@pytest.mark.parametrize("seed", [1, 2])
def test_whole_grid_bitwise(hip, sym, seed):
    c = H.grid_case(seed)
    _prepare(c, sym)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", UTILS)
@pytest.mark.parametrize("seed", [1, 2])
def test_utils(hip, sym, seed):
    """Average difference, voxel counts, the area from a delta function (sums
    to 1e-13, counts exactly), the Hessian, the phi-upwind gradient and the
    first-order delta function (bitwise) against the oracle."""
    c = H.grid_case(seed)
    _prepare(c, sym)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", CONTROL_VOLUME)
@pytest.mark.parametrize("sgn", [1, -1])
def test_control_volume(hip, sym, sgn):
    c = H.grid_case(5)
    c.scalars["cvs"] = sgn
    _prepare(c, sym)
    _cmp(hip, sym, c)


def test_control_volume_with_no_cell(hip):
    """No cell in the control volume: the CFL steps come out negative (the
    reference's initial maximum -1), the max-norm difference is the first
    cell's, sums and counts are zero -- as the oracle's."""
    for sym in CONTROL_VOLUME:
        c = H.grid_case(6)
        _prepare(c, sym)
        c.arrays["cv"] = np.full(H.shape_of(c.box("cv")), -1.0)
        c.scalars["cvs"] = 1
        _cmp(hip, sym, c)


REDUCTIONS = sorted(SUMS) + ["lsm3dmaxnormdiff_", "lsm3dcomputestableadvectiondt_",
                             "lsm3dcomputestablenormalveldt_", "lsm3dcomputestableconstnormalveldt_",
                             "lsm3dvoxelcountgreaterthanzero_", "lsm3dvoxelcountlessthanzerocontrolvolume_",
                             "lsm3dmaxnormdiffcontrolvolume_", "lsm3dcomputestablenormalveldtcontrolvolume_"]


@pytest.mark.parametrize("sym", REDUCTIONS)
def test_reductions_on_long_rows(hip, sym):
    """The reductions visit a box one row per wave, 4 x 64 cells per pass:
    rows of 300 cells (one full pass and a partial one) on 5 x 3 rows, fewer
    rows than a block has waves in the last block, against the oracle."""
    c = H.grid_case(7, gb=(-3, 302, -3, 7, -3, 5), fb=(0, 299, 0, 4, 0, 2))
    _prepare(c, sym)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("offset", [0, 1, 2, 3])
@pytest.mark.parametrize("gb", [(-3, 20, -3, 16, -3, 12), (0, 6, 0, 2, 0, 0), (0, 0, 0, 0, 0, 0),
                                (-3, 300, -2, 40, -1, 30)])
def test_zero_out_whole_ghost_box_at_any_alignment(prec, offset, gb):
    """ZeroOut clears exactly the rhs ghost box (its volume of contiguous
    scalars from the pointer), 2D and 3D, whatever the pointer's 16-byte
    alignment; the guard cells either side keep their values."""
    import torch
    import lsmlib_amd
    lib = lsmlib_amd.load(prec)
    lib.set_stream(0, True)
    dt = torch.float64 if prec == "f64" else torch.float32
    vol = (gb[1] - gb[0] + 1) * (gb[3] - gb[2] + 1) * (gb[5] - gb[4] + 1)
    buf = torch.full((vol + 8,), 7.5, dtype=dt, device="cuda")
    ptr = buf.data_ptr() + offset * buf.element_size()
    lib.api.lsm3dzerooutlevelseteqnrhs(ptr, gb)
    lib.check()
    h = buf.cpu().numpy()
    assert np.all(h[offset:offset + vol] == 0.0)
    assert np.all(h[:offset] == 7.5) and np.all(h[offset + vol:] == 7.5)
    if gb[4] == gb[5]:
        buf.fill_(7.5)
        lib.api.lsm2dzerooutlevelseteqnrhs(ptr, gb[:4])
        lib.check()
        h = buf.cpu().numpy()
        assert np.all(h[offset:offset + vol] == 0.0)
        assert np.all(h[:offset] == 7.5) and np.all(h[offset + vol:] == 7.5)


@pytest.mark.parametrize("sym", WHOLE_GRID)
def test_whole_grid_bitwise_odd_plane_count(hip, sym):
    """A fill box of 11 planes along k (and 19 x 13 in-plane): kernels that
    take two cells along k per thread end on a pair holding one cell."""
    c = H.grid_case(4, gb=(-3, 21, -3, 15, -3, 13), fb=(0, 18, 0, 12, 0, 10))
    _prepare(c, sym)
    _cmp(hip, sym, c)


PLANE_MARCH = ["lsm3dcentralgradorder2_", "lsm3dcentralgradorder4_", "lsm3dlaplacianorder2_",
               "lsm3dcomputemeancurvatureorder2_", "lsm3dcomputegaussiancurvatureorder2_", "lsm3dcentralhessian_",
               "lsm3dhjeno1_", "lsm3dhjeno2_", "lsm3dhjeno3_", "lsm3dhjweno5_", "lsm3dupwindhjeno1_",
               "lsm3dupwindhjeno2_", "lsm3dupwindhjeno3_", "lsm3dupwindhjweno5_"]


# (ghost box, fill box): 141 x 21 x 38 cells; then 129 x 9 x 17, whose last
# tile along i (for 64- and 128-wide tiles), along j and along k is one cell
MANY_TILE_BOXES = [((-3, 143, -3, 23, -3, 40), (0, 140, 0, 20, 0, 37)),
                   ((-3, 131, -3, 11, -3, 19), (0, 128, 0, 8, 0, 16))]


@pytest.mark.parametrize("boxes", MANY_TILE_BOXES)
@pytest.mark.parametrize("sym", PLANE_MARCH)
def test_plane_march_over_many_tiles_and_chunks(hip, sym, boxes):
    """The kernels that march 64 x 8 tiles (128 x 8 for the Laplacian)
    through chunks of 16 planes.  On the 141 x 21 x 38 fill box: three tiles
    along i (two of 128), the last one 13 wide, three along j (the last 5),
    three chunks along k (the last 6).  On the 129 x 9 x 17 box every last
    tile and chunk is a single cell.  phi's ghost box is three cells wider
    than the fill box on every side."""
    gb, fb = boxes
    c = H.grid_case(6, gb=gb, fb=fb)
    _prepare(c, sym)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", QUIRKY_BOUNDS)
def test_quirky_loop_bounds(hip, sym):
    c = H.grid_case(3, mixed_boxes=False)
    c.boxes["rhs"] = (-1, 19, -2, 14, -3, 10)
    _cmp(hip, sym, c)


@pytest.mark.parametrize("sym", BCS)
@pytest.mark.parametrize("idx", range(6))
def test_ghost_fill_faces(hip, sym, idx):
    c = H.grid_case(11, mixed_boxes=False)
    c.boxes["gb"] = c.gb
    c.scalars["idx"] = idx
    _cmp(hip, sym, c)


def test_box_validation_reports_error(hip):
    """A fill box that needs phi outside its ghost box is refused, not read."""
    c = H.grid_case(5, mixed_boxes=False)
    c.boxes["fb"] = (-1, 17, 0, 13, 0, 9)  # WENO5 would need phi(-4)
    before = H.run(H.oracle(), "lsm3dcentralgradorder2_", c)[0]["px"]
    res, _ = H.run(hip.api, "lsm3dhjweno5_", c, device="cuda")
    code, msg = hip.last_error()
    hip.clear_error()
    assert code == 1 and "phi" in msg
    assert np.array_equal(res["pxp"], c.arrays["pxp"])  # untouched
    del before


FACE_SYM = {0: "lsm3dlinearextrapolation", 1: "lsm3dsignedlinearextrapolation", 2: "lsm3dcopyextrapolation",
            3: "lsmhip_periodic_face"}


# ghost box, fill box: asymmetric widths; then ten layers below x (more than
# the eight lanes a row of x ghost cells gets), none on y, one on each z side
FILL_BOXES = [((-3, 12, -2, 9, -3, 7), (0, 9, 0, 6, 0, 4)), ((-10, 12, 0, 6, -1, 5), (0, 10, 0, 6, 0, 4))]


@pytest.mark.parametrize("boxes", FILL_BOXES)
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fill_all_faces_in_one_launch(kind, prec, boxes):
    """lsmhip_fill_ghosts(idx 9) fills the whole ghost shell in one launch; it
    equals the six faces applied in order x_lo .. z_hi bit for bit (FP64:
    against the oracle; FP32: against the library's own face kernels), from
    random stale ghost values and asymmetric ghost widths."""
    import ctypes
    import torch
    import lsmlib_amd
    lib = lsmlib_amd.load(prec)
    gb, fb = boxes
    rng = np.random.default_rng(kind)
    phi = rng.standard_normal(H.shape_of(gb))
    box = [ctypes.POINTER(ctypes.c_int)] * 6
    fill = lib.fn("lsmhip_fill_ghosts", [ctypes.c_void_p] + box + box + [ctypes.c_int, ctypes.c_int])
    bx = lambda b: [ctypes.byref(ctypes.c_int(v)) for v in b]  # noqa: E731
    dt = torch.float64 if prec == "f64" else torch.float32
    d = torch.tensor(phi, dtype=dt, device="cuda")
    assert fill(d.data_ptr(), *bx(gb), *bx(fb), 9, kind) == 0
    lib.check()
    got = d.cpu().numpy()
    if prec == "f64":
        want = phi.copy()
        for idx in range(6):
            getattr(H.oracle(), FACE_SYM[kind])(want, gb, fb, idx)
    else:
        e = torch.tensor(phi, dtype=dt, device="cuda")
        for idx in range(6):
            assert fill(e.data_ptr(), *bx(gb), *bx(fb), idx, kind) == 0
        want = e.cpu().numpy()
    assert np.array_equal(got, want)
