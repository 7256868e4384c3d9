"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Integer work (compaction, face tags, neighbour table) must match bit for bit.
Floating point: rho and v on fluid cells within relative L2 1e-5 of the FP64
oracle (FP32 storage), the tolerance stated for FP32 fields.
"""
import ctypes
import zlib
import os

import numpy as np
import pytest

from oracle import COracle
from oracle import compaction as oc

pytestmark = pytest.mark.gpu

FACES = ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")


def rel_l2(a, b, floor=1e-9):
    """Relative L2 error of a against b; fields whose RMS is below `floor`
    (e.g. v right after start-up, exactly 0 up to the oracle's round-off) are
    measured against that absolute floor instead."""
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), floor * np.sqrt(b.size)))


def make_solver(solid, niu, force, modes, rho_bc, u_bc, conv="reference", dtype="f32", sparse=False,
                band_rows=None):
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    nx, ny, nz = solid.shape
    s = LB3D_Solver_Single_Phase(nx, ny, nz, sparse_storage=sparse, convention=conv, dtype=dtype)
    if band_rows is not None:
        s.band_rows = band_rows
    s.solid.from_numpy(solid)
    s.niu = niu
    s.set_force(force)
    for f, name in enumerate(("x0", "x1", "y0", "y1", "z0", "z1")):
        if modes[f] == 1:
            getattr(s, "set_bc_rho_" + name)(rho_bc[f])
        elif modes[f] == 2:
            getattr(s, "set_bc_vel_" + name)(u_bc[f])
    s.init_simulation()
    return s


def make_oracle(solid, niu, force, modes, rho_bc, u_bc, conv="reference", real="f64", sparse=False):
    nx, ny, nz = solid.shape
    o = COracle(nx, ny, nz, sparse_storage=sparse, real=real, convention=conv, threads=min(16, len(os.sched_getaffinity(0))))
    o.set_solid(solid)
    o.niu = niu
    o.fx, o.fy, o.fz = force
    for f in range(6):
        if modes[f] == 1:
            o.set_bc_rho(FACES[f], rho_bc[f])
        elif modes[f] == 2:
            o.set_bc_vel(FACES[f], u_bc[f])
    o.init_simulation()
    return o


CASES = {
    # name: (shape, solid fraction, niu, force, modes, rho_bc, u_bc)
    "periodic_force": ((10, 9, 8), 0.2, 0.1, (2e-4, -1e-4, 5e-5), (0,) * 6, (1.0,) * 6, [(0, 0, 0)] * 6),
    "pressure_x": ((12, 8, 7), 0.25, 0.1, (0, 0, 0), (1, 1, 0, 0, 0, 0), (1.01, 0.99, 1, 1, 1, 1), [(0, 0, 0)] * 6),
    "mixed_all_faces": ((9, 8, 7), 0.15, 0.05, (1e-4, 0, 2e-4), (1, 2, 2, 1, 2, 2),
                        (1.005, 1, 1, 0.997, 1, 1),
                        [(0, 0, 0), (0.02, 0.01, 0), (0.0, 0.0, 0.03), (0, 0, 0), (0, 0.01, 0), (0.05, 0, 0)]),
    "pressure_force": ((8, 10, 9), 0.2, 0.3, (3e-4, 0, 0), (0, 0, 1, 1, 0, 0), (1, 1, 1.002, 0.998, 1, 1),
                       [(0, 0, 0)] * 6),
}


def case_geometry(name):
    shape, frac = CASES[name][0], CASES[name][1]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    g = (rng.random(shape) < frac).astype(np.int8)
    # a few cells of value -1: not fluid, and not "solid > 0" for the pressure faces (SP:278)
    g[(rng.random(shape) < 0.05) & (g == 1)] = -1
    return g


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("band_rows", [0, 3])
def test_compaction_bit_exact(gpu, name, band_rows):
    """Fluid list, face records and the neighbour ids the step kernel computes
    from the row table equal the brute-force restatement bit for bit, in C
    order and in bands of 3 rows."""
    import torch
    from taichi_lbm3d_amd import _lib
    solid = case_geometry(name)
    modes = CASES[name][4]
    s = make_solver(solid, CASES[name][2], CASES[name][3], modes, CASES[name][5], CASES[name][6],
                    band_rows=band_rows)
    cells, d2c = oc.compact(solid, band_rows)
    assert s.n_fluid == cells.size
    np.testing.assert_array_equal(s._dev["cells"][: cells.size].cpu().numpy(), cells)
    np.testing.assert_array_equal(s.neighbor_table(), oc.neighbors(solid, cells, d2c))
    words = oc.pack_face_bits(oc.face_records(solid, cells, modes))
    np.testing.assert_array_equal(s._dev["face_bits"][: words.size].cpu().numpy().view(np.uint32), words)
    # dense_to_compact through the raw C ABI
    lib = _lib.lib()
    p = s._params
    n_dense = solid.size
    sol = torch.from_numpy(solid.reshape(-1).copy()).to(gpu)
    ws_bytes = lib.size("lbm3d_compact_workspace_bytes", n_dense)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=gpu)
    d2c_g = torch.empty(n_dense, dtype=torch.int32, device=gpu)
    cells_g = torch.empty(cells.size, dtype=torch.int32, device=gpu)
    count = torch.zeros(1, dtype=torch.int32, device=gpu)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    lib.call("lbm3d_compact_count", None, ctypes.byref(p), vp(sol), vp(count), vp(ws), ws_bytes)
    lib.call("lbm3d_compact_fill", None, ctypes.byref(p), vp(sol), vp(d2c_g), vp(cells_g), vp(ws), ws_bytes)
    torch.cuda.synchronize()
    assert int(count.cpu()[0]) == cells.size
    np.testing.assert_array_equal(d2c_g.cpu().numpy(), d2c)
    np.testing.assert_array_equal(cells_g.cpu().numpy(), cells)


def test_compaction_pack32_fixture(gpu):
    """The GPU fluid list, neighbour ids and face records of the seeded 32^3 pack
    equal the plain-loop restatement's, stored in tests/golden."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "pack32_compaction.npz"))
    solid = np.unpackbits(g["bits"])[: 32 ** 3].reshape(32, 32, 32).astype(np.int8)
    modes = tuple(int(m) for m in g["modes"])
    s = make_solver(solid, 0.1, (0, 0, 0), modes, (1.001, 1.0, 1, 1, 1, 1), [(0, 0, 0)] * 6)
    np.testing.assert_array_equal(s._dev["cells"][: s.n_fluid].cpu().numpy(), g["cells"])
    np.testing.assert_array_equal(s.neighbor_table(), g["nbr"])
    words = oc.pack_face_bits(g["face_rec"])
    np.testing.assert_array_equal(s._dev["face_bits"][: words.size].cpu().numpy().view(np.uint32), words)


def test_compaction_cavity_fixture(gpu):
    """The reference's 50^3 cavity mask: 112,896 fluid cells (Single_phase/geo_cavity.dat),
    the list in C order (the checksums are of the C-order list)."""
    import hashlib
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cavity_compaction.npz"))
    from taichi_lbm3d_amd.geometry import cavity
    solid = cavity(50)
    s = make_solver(solid, 0.1, (0, 0, 0), (0, 2, 0, 0, 0, 0), (1,) * 6, [(0, 0, 0), (0, 0, 0.1)] + [(0, 0, 0)] * 4,
                    band_rows=0)
    assert s.n_fluid == 112896
    cells = s._dev["cells"].cpu().numpy()
    assert hashlib.sha256(cells.astype("<i4").tobytes()).hexdigest() == str(g["sha_cells"])
    assert hashlib.sha256(s.neighbor_table().astype("<i4").tobytes()).hexdigest() == str(g["sha_nbr"])
    words = oc.pack_face_bits(g["face_rec"].astype(np.uint32))
    np.testing.assert_array_equal(s._dev["face_bits"][: words.size].cpu().numpy().view(np.uint32), words)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("conv", ["reference", "documented", "textbook"])
@pytest.mark.parametrize("steps", [1, 2, 7, 40])
def test_step_matches_oracle(gpu, name, conv, steps):
    solid = case_geometry(name)
    args = CASES[name][2:]
    s = make_solver(solid, *args, conv=conv)
    o = make_oracle(solid, *args, conv=conv)
    s.run(steps)
    o.step(steps)
    fl = solid == 0
    rho, v = s.rho.to_numpy(), s.v.to_numpy()
    er, ev = rel_l2(rho[fl], o.rho()[fl]), rel_l2(v[fl], o.v()[fl])
    print(f"{name} {conv} {steps}: rel-L2 rho {er:.2e} v {ev:.2e}")
    assert er < 1e-6 and ev < 1e-5
    # non-fluid cells: rho = 1, v = 0 in dense storage (SP:390-392)
    assert np.all(rho[~fl] == 1.0) and np.all(v[~fl] == 0.0)


@pytest.mark.parametrize("name", sorted(CASES))
def test_step_band_order_matches_oracle(gpu, name):
    """The fluid list in bands of 3 rows (every block of the step kernel spans
    several planes and takes the row-table path): the same fields as the
    oracle."""
    solid = case_geometry(name)
    args = CASES[name][2:]
    s = make_solver(solid, *args, band_rows=3)
    o = make_oracle(solid, *args)
    s.run(40)
    o.step(40)
    fl = solid == 0
    rho, v = s.rho.to_numpy(), s.v.to_numpy()
    er, ev = rel_l2(rho[fl], o.rho()[fl]), rel_l2(v[fl], o.v()[fl])
    print(f"{name} bands of 3: rel-L2 rho {er:.2e} v {ev:.2e}")
    assert er < 1e-6 and ev < 1e-5


@pytest.mark.parametrize("band_rows", [0, 16, 20])
def test_porous_band_order_matches_oracle(gpu, band_rows):
    """A 12 x 48 x 128 random porous box, pressure-driven along x, with a force
    along z: in bands of 16 rows 167 of the 202 blocks of 256 cells lie in one
    plane's band and take the LDS-window path, the other 35 span two planes;
    bands of 20 rows leave a last band of 8.  The fields after 60 steps equal
    the FP64 oracle's, and every list order gives the same fields bit for bit
    as C order."""
    rng = np.random.default_rng(11)
    solid = (rng.random((12, 48, 128)) < 0.3).astype(np.int8)
    args = (0.1, (0, 0, 2e-5), (1, 1, 0, 0, 0, 0), (1.002, 1.0, 1, 1, 1, 1), [(0, 0, 0)] * 6)
    s = make_solver(solid, *args, band_rows=band_rows)
    o = make_oracle(solid, *args)
    s.run(60)
    o.step(60)
    fl = solid == 0
    rho, v = s.rho.to_numpy(), s.v.to_numpy()
    er, ev = rel_l2(rho[fl], o.rho()[fl]), rel_l2(v[fl], o.v()[fl])
    print(f"porous bands of {band_rows}: rel-L2 rho {er:.2e} v {ev:.2e}")
    assert er < 1e-6 and ev < 1e-5
    if band_rows:
        c = make_solver(solid, *args, band_rows=0)
        c.run(60)
        np.testing.assert_array_equal(c.rho.to_numpy(), rho)
        np.testing.assert_array_equal(c.v.to_numpy(), v)


def test_initial_state_macros(gpu):
    """Before the first step rho = 1 and v = 0 exactly, even with a body force (SP:165-166)."""
    solid = case_geometry("periodic_force")
    s = make_solver(solid, *CASES["periodic_force"][2:])
    assert np.all(s.rho.to_numpy() == 1.0) and np.all(s.v.to_numpy() == 0.0)


def test_sparse_storage_fields(gpu):
    """sparse_storage=True: padded extent 3*(n//3+1) and zeros outside active 3^3 blocks."""
    rng = np.random.default_rng(3)
    solid = np.ones((11, 10, 9), np.int8)
    solid[2:9, 1:8, 3:8] = (rng.random((7, 7, 5)) < 0.3)
    args = (0.1, (1e-4, 0, 0), (0,) * 6, (1,) * 6, [(0, 0, 0)] * 6)
    s = make_solver(solid, *args, sparse=True)
    o = make_oracle(solid, *args, sparse=True)
    s.run(9)
    o.step(9)
    rho, v = s.rho.to_numpy(), s.v.to_numpy()
    assert rho.shape == o.shape() == (12, 12, 12) and v.shape == (12, 12, 12, 3)
    pad_fluid = np.zeros(rho.shape, bool)
    pad_fluid[:11, :10, :9] = solid == 0
    np.testing.assert_array_equal(rho[~pad_fluid], o.rho()[~pad_fluid])
    np.testing.assert_array_equal(v[~pad_fluid], o.v()[~pad_fluid])
    assert rel_l2(rho[pad_fluid], o.rho()[pad_fluid]) < 1e-6
    assert rel_l2(v[pad_fluid], o.v()[pad_fluid]) < 1e-5


def test_get_max_v(gpu):
    solid = case_geometry("mixed_all_faces")
    args = CASES["mixed_all_faces"][2:]
    s = make_solver(solid, *args)
    o = make_oracle(solid, *args)
    s.run(12)
    o.step(12)
    assert abs(s.get_max_v() - o.get_max_v()) <= 1e-6 * o.get_max_v()
    assert s.max_v[None] == s.get_max_v()


@pytest.mark.parametrize("band_rows", [3, 16])
def test_f64_band_order_matches_fp64_oracle(gpu, band_rows):
    """FP64 storage and arithmetic on the 12 x 48 x 128 porous box with the list
    in bands of rows (in bands of 16, 35 of the 202 blocks span two planes'
    segments and use two-segment windows; in bands of 3 most blocks span
    several segments): the same algorithm as the FP64 oracle to round-off."""
    rng = np.random.default_rng(11)
    solid = (rng.random((12, 48, 128)) < 0.3).astype(np.int8)
    args = (0.1, (0, 0, 2e-5), (1, 1, 0, 0, 0, 0), (1.002, 1.0, 1, 1, 1, 1), [(0, 0, 0)] * 6)
    s = make_solver(solid, *args, band_rows=band_rows, dtype="f64")
    o = make_oracle(solid, *args)
    s.run(41)
    o.step(41)
    rho, v = s.compact_macros(native=True)
    cells = s._dev["cells"].cpu().numpy()
    n = s.n_fluid
    rho = rho.cpu().numpy()[:n]
    v = v.cpu().numpy()[: 3 * n].reshape(3, n).T
    orho = o.rho().reshape(-1)[cells]
    ov = o.v().reshape(-1, 3)[cells]
    assert np.abs(rho - orho).max() < 1e-13
    assert rel_l2(v, ov, floor=1e-12) < 1e-10
