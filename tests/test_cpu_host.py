"""CPU tests: fixtures, compaction restatements, the C ABI surface and host logic."""
import ctypes
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import compaction as oc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_mask(name, shape):
    d = np.load(os.path.join(GOLDEN, name))
    bits = np.unpackbits(d["bits"])[: int(d["n"])]
    return bits.reshape(shape, order="F"), int(d["n_solid"])


# ---- fixtures ---------------------------------------------------------------

def test_cavity_fixture_matches_generator():
    """geo_cavity.dat (50^3) is the mask flow_domain_geo_generation_2D.py:15-27 writes."""
    from taichi_lbm3d_amd.geometry import cavity
    solid, n_solid = load_mask("geo_cavity.npz", (50, 50, 50))
    assert n_solid == 12104 and int((solid == 0).sum()) == 112896
    np.testing.assert_array_equal(solid, cavity(50))


def test_read_geo_layout(tmp_path):
    """init_geo reads one value per line, x fastest (SP:173-177, README.md:17-23)."""
    from taichi_lbm3d_amd.geometry import read_geo, write_geo
    rng = np.random.default_rng(0)
    a = (rng.random((5, 4, 3)) < 0.4).astype(np.uint8)
    p = tmp_path / "g.dat"
    write_geo(str(p), a)
    lines = p.read_text().split()
    assert lines[1] == str(a[1, 0, 0]) and lines[5] == str(a[0, 1, 0]) and lines[20] == str(a[0, 0, 1])
    np.testing.assert_array_equal(read_geo(str(p), 5, 4, 3), a)
    p.write_text("\n".join(["0", "2", "0.5", "-1"] * 15))  # values > 0 are solid
    g = read_geo(str(p), 5, 4, 3)
    assert g.dtype == np.int8 and g.reshape(-1, order="F")[:4].tolist() == [0, 1, 1, -1]


# ---- compaction -------------------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_compaction_loops_equal_numpy(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(x) for x in rng.integers(2, 7, 3))
    # values of the i8 field: 0 fluid, > 0 solid, < 0 neither fluid nor "solid > 0"
    solid = rng.choice(np.array([0, 1, -1, 2], np.int8), size=shape, p=[0.7, 0.15, 0.1, 0.05])
    modes = tuple(int(x) for x in rng.integers(0, 3, 6))
    cells, d2c = oc.compact(solid)
    nbr = oc.neighbors(solid, cells, d2c)
    rec = oc.face_records(solid, cells, modes)
    c2, rec2, nb2 = oc.compact_loops(solid, modes)
    np.testing.assert_array_equal(cells, c2)
    np.testing.assert_array_equal(rec, rec2)
    np.testing.assert_array_equal(nbr, nb2)
    words = oc.pack_face_bits(rec)
    unpacked = (words[np.arange(rec.size) // 16] >> (2 * (np.arange(rec.size) % 16)).astype(np.uint32)) & 3
    np.testing.assert_array_equal(unpacked, rec)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("band_rows", [1, 2, 3, 5])
def test_band_order_loops_equal_numpy(seed, band_rows):
    """The list in bands of R rows: numpy and plain loops agree on the numbering,
    neighbour table and face records, also when R does not divide ny."""
    rng = np.random.default_rng(100 + seed)
    shape = (int(rng.integers(2, 6)), int(rng.integers(4, 12)), int(rng.integers(2, 6)))
    solid = rng.choice(np.array([0, 1, -1], np.int8), size=shape, p=[0.7, 0.2, 0.1])
    modes = tuple(int(x) for x in rng.integers(0, 3, 6))
    cells, d2c = oc.compact(solid, band_rows)
    c2, rec2, nb2 = oc.compact_loops(solid, modes, band_rows)
    np.testing.assert_array_equal(cells, c2)
    np.testing.assert_array_equal(oc.neighbors(solid, cells, d2c), nb2)
    np.testing.assert_array_equal(oc.face_records(solid, cells, modes), rec2)


@pytest.mark.parametrize("shape,band_rows", [((3, 10, 4), 4), ((5, 7, 3), 2), ((2, 9, 6), 3), ((4, 6, 5), 6)])
def test_band_order_closed_form(shape, band_rows):
    """All fluid: the id of (i, j, k) is jb nx R nz + i Rb nz + (j - jb R) nz + k,
    jb = j // R and Rb = min(R, ny - jb R) the rows of its band; R >= ny and
    R = 0 give C order.  With solids, every row's fluid cells keep consecutive
    ids in k order (what the row table encodes)."""
    nx, ny, nz = shape
    R = band_rows
    _, d2c = oc.compact(np.zeros(shape, np.int8), R)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    jb = j // R
    Rb = np.minimum(R, ny - jb * R)
    expect = jb * nx * R * nz + i * Rb * nz + (j - jb * R) * nz + k
    np.testing.assert_array_equal(d2c.reshape(shape), expect)
    solid = (np.random.default_rng(3).random(shape) < 0.3).astype(np.int8)
    for r in (0, ny, ny + 3):
        np.testing.assert_array_equal(oc.compact(solid, r)[0], np.flatnonzero(solid.reshape(-1) == 0))
    _, d2c = oc.compact(solid, R)
    ids = d2c.reshape(shape)
    for a in range(nx):
        for b in range(ny):
            row = ids[a, b][ids[a, b] >= 0]
            assert np.all(np.diff(row) == 1)


def test_compaction_cavity_checksum():
    solid, _ = load_mask("geo_cavity.npz", (50, 50, 50))
    g = np.load(os.path.join(GOLDEN, "cavity_compaction.npz"))
    cells, d2c = oc.compact(solid)
    nbr = oc.neighbors(solid, cells, d2c)
    assert cells.size == int(g["n_cells"]) == 112896
    assert hashlib.sha256(cells.astype("<i4").tobytes()).hexdigest() == str(g["sha_cells"])
    assert hashlib.sha256(nbr.astype("<i4").tobytes()).hexdigest() == str(g["sha_nbr"])
    np.testing.assert_array_equal(oc.face_records(solid, cells, tuple(g["modes"])), g["face_rec"])


def test_compaction_all_fluid_docs_fixture():
    """docs/geo.dat: 48^3 all fluid; every neighbour exists (periodic wrap, no bounce-back)."""
    solid, n_solid = load_mask("geo_docs48.npz", (48, 48, 48))
    assert n_solid == 0
    cells, d2c = oc.compact(solid)
    nbr = oc.neighbors(solid, cells, d2c)
    assert cells.size == 48 ** 3 and nbr.min() >= 0
    # each direction is a permutation of the cells
    for q in range(18):
        assert np.unique(nbr[q]).size == cells.size


def test_compaction_pack32_fixture():
    """Seeded 32^3 sphere pack: the generator is deterministic and the vectorised
    compaction and neighbour table equal the plain-loop ones stored in the fixture."""
    from taichi_lbm3d_amd.geometry import sphere_pack
    g = np.load(os.path.join(GOLDEN, "pack32_compaction.npz"))
    solid = np.unpackbits(g["bits"])[: 32 ** 3].reshape(32, 32, 32)
    np.testing.assert_array_equal(solid, sphere_pack((32, 32, 32), porosity=0.35, rmin=3.0, rmax=6.0, seed=7,
                                                     periodic=(False, True, True)))
    cells, d2c = oc.compact(solid)
    np.testing.assert_array_equal(cells, g["cells"])
    np.testing.assert_array_equal(oc.neighbors(solid, cells, d2c), g["nbr"])
    np.testing.assert_array_equal(oc.face_records(solid, cells, tuple(g["modes"])), g["face_rec"])
    assert (g["nbr"] < 0).any() and (g["nbr"] >= 0).any()


# ---- the C ABI --------------------------------------------------------------

def _header_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(lbm3d_[a-z0-9_]+)\s*\(", txt))
    return sorted(names - {"lbm3d_params"})


def test_abi_exports_every_declared_symbol():
    from taichi_lbm3d_amd import _lib
    lib = _lib.lib()
    names = _header_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib.cdll, n), n
        assert n in _lib.SIGNATURES, n
    assert lib.cdll.lbm3d_abi_version() == _lib.ABI_VERSION
    out = subprocess.run(["nm", "-D", "--defined-only", lib.path], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(lbm3d_\w+)\b", out))
    assert set(names) <= exported


def test_params_struct_layout(tmp_path):
    """ctypes mirror of lbm3d_params has the C layout (size and every offset)."""
    from taichi_lbm3d_amd._lib import Params
    fields = [f[0] for f in Params._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'printf("{f} %zu\\n", offsetof(lbm3d_params, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lbm3d.h"\nint main(void){'
                   f'printf("size %zu\\n", sizeof(lbm3d_params));{body}return 0;}}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n")
               if line)
    assert int(got["size"]) == ctypes.sizeof(Params)
    for f in fields:
        assert int(got[f]) == getattr(Params, f).offset, f


def test_abi_rejects_bad_arguments_without_gpu():
    """Argument checks run before any HIP call and report through lbm3d_last_error."""
    from taichi_lbm3d_amd import _lib
    lib = _lib.lib()
    p = _lib.Params()
    p.nx = p.ny = p.nz = 4
    vp = ctypes.c_void_p(8)
    p.n_cells = 100
    rc = lib.cdll.lbm3d_step(None, ctypes.byref(p), 0, vp, vp, vp, vp, vp, 0)
    assert rc == -1 and "n_cells" in lib.last_error()
    p.n_cells = 10
    rc = lib.cdll.lbm3d_step(None, ctypes.byref(p), 7, vp, vp, vp, vp, vp, 0)
    assert rc == -1 and "dtype" in lib.last_error()
    rc = lib.cdll.lbm3d_step(None, ctypes.byref(p), 0, vp, None, vp, vp, vp, 0)
    assert rc == -1 and "NULL" in lib.last_error()
    p.bc_mode[0] = 1
    p.face_u_off[0] = -1
    rc = lib.cdll.lbm3d_step(None, ctypes.byref(p), 0, vp, vp, vp, vp, vp, 0)
    assert rc == -1 and "face_u_off" in lib.last_error()
    p.nx = p.ny = p.nz = 1000
    p.bc_mode[0] = 0
    p.n_cells = 600_000_000  # two FP32 slots beyond one 32-bit buffer range
    rc = lib.cdll.lbm3d_step(None, ctypes.byref(p), 0, vp, vp, vp, vp, vp, 0)
    assert rc == -3
    p.nx = p.ny = p.nz = 4
    p.n_cells = 10
    p.band_rows = -1
    rc = lib.cdll.lbm3d_compact_count(None, ctypes.byref(p), vp, vp, vp, 1 << 20)
    assert rc == -1 and "band_rows" in lib.last_error()
    p.band_rows = 0
    rc = lib.cdll.lbm3d_round16(None, 0, 0, vp, vp, vp, 8)
    assert rc == -1 and "dtype" in lib.last_error()
    rc = lib.cdll.lbm3d_round16(None, 1, 1, vp, None, vp, 8)
    assert rc == -1 and "NULL" in lib.last_error()
    assert lib.cdll.lbm3d_round16(None, 2, 0, None, None, None, 0) == 0
    p.nx = 1
    p.bc_mode[0] = 1
    rc = lib.cdll.lbm3d_compact_count(None, ctypes.byref(p), vp, vp, vp, 1 << 20)
    assert rc == -1 and "extent" in lib.last_error()
    assert lib.cdll.lbm3d_compact_workspace_bytes(2 * 1024) == 8
    p.nx, p.ny, p.nz = 5, 6, 70
    assert lib.cdll.lbm3d_row_words(ctypes.byref(p)) == 5 * 6 * 3
    assert lib.cdll.lbm3d_face_bit_words(33) == 3
    assert [lib.cdll.lbm3d_face_plane(ctypes.byref(p), f) for f in range(6)] == [420, 420, 350, 350, 30, 30]
    assert lib.cdll.lbm3d_slot_stride(65) == 128


def test_batch_struct_layouts(tmp_path):
    """ctypes mirrors of lbm3d_sample, lbm3d_batch_sizes and lbm3d_launch have
    the C layout."""
    from taichi_lbm3d_amd._lib import BatchSizes, Launch, Sample
    structs = ((Sample, "lbm3d_sample"), (BatchSizes, "lbm3d_batch_sizes"), (Launch, "lbm3d_launch"))
    checks = []
    for cls, cname in structs:
        checks.append(f'printf("{cname}.size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            checks.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    src = tmp_path / "batch_layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "lbm3d.h"\nint main(void){'
                   + "".join(checks) + 'printf("align %d\\n", LBM3D_BATCH_ALIGN);return 0;}\n')
    exe = tmp_path / "batch_layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n")
               if line)
    from taichi_lbm3d_amd import _lib
    assert int(got["align"]) == _lib.BATCH_ALIGN
    for cls, cname in structs:
        assert int(got[cname + ".size"]) == ctypes.sizeof(cls)
        for f in cls._fields_:
            assert int(got[f"{cname}.{f[0]}"]) == getattr(cls, f[0]).offset, (cname, f[0])


def test_batch_layout_host_only():
    """Samples start on LBM3D_BATCH_ALIGN boundaries in order, each owning at least
    one chunk; row and face-state blocks are packed in order (no GPU involved)."""
    from taichi_lbm3d_amd import _lib
    lib = _lib.lib()
    p = _lib.Params()
    p.bc_mode[0] = p.bc_mode[1] = 1          # pressure on both x faces
    p.bc_mode[4] = 2                         # velocity on z-left: no face state
    shapes = [(10, 9, 8), (7, 12, 33), (16, 5, 6), (4, 4, 4)]
    n_cells = [300, 1000, 0, 64]
    smp = (_lib.Sample * 4)()
    for i, (sh, n) in enumerate(zip(shapes, n_cells)):
        smp[i].nx, smp[i].ny, smp[i].nz = sh
        smp[i].n_cells = n
    sz = _lib.BatchSizes()
    lib.call("lbm3d_batch_layout", ctypes.byref(p), smp, 4, ctypes.byref(sz))
    assert [smp[i].cell_begin for i in range(4)] == [0, 512, 1536, 2048]
    assert sz.cells == 2560
    words = [nx * ny * ((nz + 31) // 32) for nx, ny, nz in shapes]
    assert [smp[i].rows_begin for i in range(4)] == list(np.cumsum([0] + words[:-1]))
    assert sz.row_words == sum(words)
    fu = [3 * 2 * ny * nz for _, ny, nz in shapes]
    assert [smp[i].face_u_begin for i in range(4)] == list(np.cumsum([0] + fu[:-1]))
    assert sz.face_u == sum(fu)
    buf = np.zeros(int(sz.desc_bytes), np.uint8)
    lib.call("lbm3d_batch_desc", ctypes.byref(p), smp, 4, ctypes.c_void_p(buf.ctypes.data), int(sz.desc_bytes))
    # errors: a sample with a pressure face of extent 1, descriptors of a foreign layout, a short buffer
    bad = (_lib.Sample * 1)()
    bad[0].nx, bad[0].ny, bad[0].nz, bad[0].n_cells = 1, 4, 4, 16
    assert lib.cdll.lbm3d_batch_layout(ctypes.byref(p), bad, 1, ctypes.byref(sz)) == -1
    assert "extent" in lib.last_error()
    smp[1].cell_begin += 512
    assert lib.cdll.lbm3d_batch_desc(ctypes.byref(p), smp, 4, ctypes.c_void_p(buf.ctypes.data), buf.size) == -1
    smp[1].cell_begin -= 512
    assert lib.cdll.lbm3d_batch_desc(ctypes.byref(p), smp, 4, ctypes.c_void_p(buf.ctypes.data), 64) == -1
    # batch calls check the batch cell count before touching the device
    vp = ctypes.c_void_p(256)
    p.n_cells = 1000                          # not a multiple of the alignment
    rc = lib.cdll.lbm3d_batch_step(None, ctypes.byref(p), 0, vp, vp, 4, vp, vp, vp, vp, None, 0, None)
    assert rc == -1 and "n_cells" in lib.last_error()
    p.n_cells = 2560
    shape = (_lib.Launch * 2)()
    shape[0].block, shape[1].block = 96, 256
    rc = lib.cdll.lbm3d_batch_step(None, ctypes.byref(p), 0, vp, vp, 4, vp, vp, vp, vp, None, 0, shape)
    assert rc == -1 and "block" in lib.last_error()


# ---- host logic ---------------------------------------------------------------

def test_batch_group_ranges():
    """The cache-blocked schedule's groups cover the batch once, in order, on
    sample boundaries, each within the byte budget unless one sample alone
    exceeds it."""
    from taichi_lbm3d_amd.batch import group_ranges
    rng = np.random.default_rng(0)
    for _ in range(50):
        sizes = rng.integers(1, 40, size=int(rng.integers(1, 30))) * 512
        begins = np.concatenate([[0], np.cumsum(sizes)[:-1]]).tolist()
        total = int(sizes.sum())
        budget = int(rng.integers(1, 200)) * 512 * 76
        g = group_ranges(begins, total, 76, budget)
        assert g[0][0] == 0 and g[-1][1] == total
        assert all(a[1] == b[0] for a, b in zip(g[:-1], g[1:]))
        assert all(lo < hi and lo in begins for lo, hi in g)
        for lo, hi in g:
            inside = [b for b in begins if lo <= b < hi]
            assert (hi - lo) * 76 <= budget or len(inside) == 1
    assert group_ranges([0, 512, 1024], 1536, 76, 10 ** 9) == [(0, 1536)]
    assert group_ranges([0, 512, 1024], 1536, 76, 1) == [(0, 512), (512, 1024), (1024, 1536)]


def test_relaxation_rates():
    from taichi_lbm3d_amd import relaxation_rates
    tau, sv, so = relaxation_rates(0.1)
    assert abs(tau - (0.1 / 3 + 0.5)) < 1e-15 and abs(sv - 1 / tau) < 1e-15
    assert abs(so - 8 * (2 - sv) / (8 - sv)) < 1e-15
    tau, _, _ = relaxation_rates(0.1, "documented")
    assert abs(tau - 0.8) < 1e-15
    with pytest.raises(ValueError):
        relaxation_rates(0.1, "nope")


def test_solver_surface_without_gpu():
    """The reference's attribute surface and setters (SP:13-28, SP:405-458)."""
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    s = LB3D_Solver_Single_Phase(5, 4, 3)
    assert (s.nx, s.ny, s.nz) == (5, 4, 3) and s.niu == 0.16667 and s.sparse_storage is False
    assert s.bc_x_left == 0 and s.rho_bcxl == 1.0 and s.vx_bcxl == 0.0e-5
    s.set_bc_vel_x1([0.0, 0.0, 0.1])
    assert s.bc_x_right == 2 and (s.vx_bcxr, s.vy_bcxr, s.vz_bcxr) == (0.0, 0.0, 0.1)
    s.set_bc_rho_y0(1.01)
    assert s.bc_y_left == 1 and s.rho_bcyl == 1.01
    s.set_force([1e-5, 0, 0])
    s.set_viscosity(0.2)
    assert s.fx == 1e-5 and s.niu == 0.2
    s.solid.from_numpy(np.ones((5, 4, 3)) * 0.0)
    assert s.solid.to_numpy().dtype == np.int8 and s.solid.shape == (5, 4, 3)
    with pytest.raises(ValueError):
        s.solid.from_numpy(np.zeros((4, 4, 4)))
    assert s.rho.to_numpy().shape == (5, 4, 3) and s.v.to_numpy().shape == (5, 4, 3, 3)
    sp = LB3D_Solver_Single_Phase(5, 4, 3, sparse_storage=True)
    assert sp.rho.shape == (6, 6, 6)
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            s.init_simulation()
        with pytest.raises(RuntimeError):
            s.step()


# ---- geometry generators and VTK -----------------------------------------------

def test_sphere_pack_properties():
    from taichi_lbm3d_amd.geometry import sphere_pack
    a = sphere_pack((48, 40, 40), porosity=0.4, rmin=3, rmax=6, seed=1)
    b = sphere_pack((48, 40, 40), porosity=0.4, rmin=3, rmax=6, seed=1)
    np.testing.assert_array_equal(a, b)
    phi = (a == 0).mean()
    assert 0.38 < phi <= 0.4
    c = sphere_pack((48, 40, 40), porosity=0.4, rmin=3, rmax=6, seed=2)
    assert not np.array_equal(a, c)
    # periodic in y and z: solid touching y = 0 continues at y = ny-1 somewhere
    assert (a[:, 0, :] & a[:, -1, :]).any() and (a[:, :, 0] & a[:, :, -1]).any()


def sphere_pack_one_at_a_time(shape, porosity, rmin, rmax, seed, periodic):
    """The pack's definition, a sphere at a time: draw the centre (3 uniforms
    times the extents) and the radius, mark the cells within r (wrapped along
    periodic axes, clipped along the others), stop once the solid count
    reaches (1 - porosity) of the box."""
    n = np.array(shape)
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, np.uint8)
    target = (1.0 - porosity) * solid.size
    while solid.sum(dtype=np.int64) < target:
        c = rng.uniform(0.0, 1.0, 3) * n
        r = rng.uniform(rmin, rmax)
        idx, dist2 = [], 0.0
        for a in range(3):
            cells = np.arange(int(np.floor(c[a] - r)), int(np.ceil(c[a] + r)) + 1)
            d = cells - c[a]
            if periodic[a]:
                cells = np.mod(cells, n[a])
            else:
                keep = (cells >= 0) & (cells < n[a])
                cells, d = cells[keep], d[keep]
            shp = [1, 1, 1]
            shp[a] = -1
            idx.append(cells.reshape(shp))
            dist2 = dist2 + (d * d).reshape(shp)
        ii, jj, kk = np.broadcast_arrays(*idx)
        inside = dist2 <= r * r
        solid[ii[inside], jj[inside], kk[inside]] = 1
    return solid


@pytest.mark.parametrize("case", [((32, 32, 32), 0.35, 3.0, 6.0, 7, (False, True, True)),
                                  ((48, 40, 40), 0.4, 3, 6, 1, (False, False, False)),
                                  ((10, 12, 9), 0.3, 3, 7, 4, (True, True, True)),
                                  ((5, 6, 40), 0.5, 2, 9, 3, (True, False, True)),
                                  ((71, 71, 71), 0.37, 4.0, 12.0, 1000, (True, True, True))])
def test_sphere_pack_equals_one_sphere_at_a_time(case):
    """The chunked generator gives the cells of the sequential definition."""
    from taichi_lbm3d_amd.geometry import sphere_pack
    np.testing.assert_array_equal(sphere_pack(*case), sphere_pack_one_at_a_time(*case))


def test_cube_channel_properties():
    from taichi_lbm3d_amd.geometry import cube_channel
    g = cube_channel((40, 32, 32), fraction=0.05, seed=0)
    assert g[:, 0].all() and g[:, -1].all() and g[:, :, 0].all() and g[:, :, -1].all()
    inner = g[:, 1:-1, 1:-1]
    assert 0.04 < inner.mean() < 0.12


def test_vtk_roundtrip(tmp_path):
    from taichi_lbm3d_amd.vtk import grid_to_vtk, read_vtr
    rng = np.random.default_rng(0)
    n = (5, 4, 3)
    solid = (rng.random(n) < 0.5).astype(np.int8)
    rho = rng.random(n).astype(np.float32)
    v = [rng.random(n).astype(np.float32) for _ in range(3)]
    x, y, z = (np.linspace(0, k, k) for k in n)
    fn = grid_to_vtk(str(tmp_path / "LB_SingelPhase_7"), x, y, z,
                     pointData={"Solid": solid, "rho": rho, "velocity": tuple(v)})
    assert fn.endswith("LB_SingelPhase_7.vtr")
    back = read_vtr(fn)
    np.testing.assert_array_equal(back["Solid"], solid)
    np.testing.assert_array_equal(back["rho"], rho)
    np.testing.assert_array_equal(back["velocity"], np.stack(v, -1))
    np.testing.assert_array_equal(back["x"], x)
    head = open(fn, "rb").read(400).decode("ascii", "replace")
    assert 'type="RectilinearGrid"' in head and 'WholeExtent="0 4 0 3 0 2"' in head


# ---- geometry input -----------------------------------------------------------------

def test_native_geo_reader_matches_loadtxt(tmp_path):
    """lbm3d_read_geo_ascii == the reference's np.loadtxt path (oracle) on the
    one-value-per-line format: integers, fractions, exponents, negatives, CRLF."""
    from oracle.geo_ref import read_geo_loadtxt
    from taichi_lbm3d_amd.geometry import read_geo
    rng = np.random.default_rng(3)
    shape = (7, 5, 6)
    vals = rng.choice(["0", "1", "2", "-1", "0.5", "-0.5", "-1.7", "1e-3", "0.0", "+1", "3.0E0"], size=np.prod(shape))
    p = tmp_path / "g.dat"
    p.write_text("\n".join(vals) + "\n")
    np.testing.assert_array_equal(read_geo(str(p), *shape), read_geo_loadtxt(str(p), *shape))
    p.write_text("\r\n".join(" " + v + "\t" for v in vals) + "\r\n\n")  # CRLF, padding, blank last line
    np.testing.assert_array_equal(read_geo(str(p), *shape), read_geo_loadtxt(str(p), *shape))
    with pytest.raises(ValueError):
        read_geo(str(p), 7, 5, 5)
    p.write_text("0\n1\nx\n")
    with pytest.raises(ValueError):
        read_geo(str(p), 3, 1, 1)


def test_stl_read_write_roundtrip(tmp_path):
    from taichi_lbm3d_amd.stl import read_stl, sphere_mesh, write_stl
    tris = sphere_mesh((1.0, 2.0, 3.0), 2.5, subdivisions=2)
    write_stl(str(tmp_path / "s.stl"), tris)
    np.testing.assert_array_equal(read_stl(str(tmp_path / "s.stl")), tris)
    lines = ["solid s"]
    for t in tris:
        lines += ["facet normal 0 0 0", " outer loop"] + [f"  vertex {x!r} {y!r} {z!r}" for x, y, z in t.tolist()]
        lines += [" endloop", "endfacet"]
    (tmp_path / "a.stl").write_text("\n".join(lines + ["endsolid s"]))
    np.testing.assert_allclose(read_stl(str(tmp_path / "a.stl")), tris, rtol=1e-6)


def test_winding_number_oracle_closed_meshes():
    """w = 1 inside and 0 outside a closed outward mesh (box, icosphere); this pins
    the solid-angle formula and the meshes' orientation used by the GPU tests."""
    from oracle.geo_ref import winding_number
    from taichi_lbm3d_amd.stl import box_mesh, sphere_mesh
    rng = np.random.default_rng(0)
    box = box_mesh((0.0, 0.0, 0.0), (2.0, 3.0, 1.5))
    pts = rng.uniform(-1, 4, (400, 3))
    inside = np.all((pts > 0) & (pts < np.array([2.0, 3.0, 1.5])), axis=1)
    w = winding_number(box, pts)
    np.testing.assert_allclose(w, inside.astype(float), atol=1e-9)
    sph = sphere_mesh((0.5, -0.5, 1.0), 2.0, subdivisions=3)
    r = np.linalg.norm(pts - np.array([0.5, -0.5, 1.0]), axis=1)
    w = winding_number(sph, pts)
    far = np.abs(r - 2.0) > 0.05
    np.testing.assert_allclose(w[far], (r[far] < 2.0).astype(float), atol=1e-9)


def test_distance_oracle_matches_brute_force():
    """The periodic distance oracle (scipy's exact EDT on a wrap-padded copy,
    capped at the margin) equals a brute-force minimum over every solid cell and
    every periodic image, on small grids, including one smaller than the margin
    and one without solid cells."""
    from oracle.geo_ref import periodic_distance_edt
    rng = np.random.default_rng(7)
    for shape, frac, margin in [((6, 5, 7), 0.1, 24), ((9, 4, 8), 0.05, 3), ((5, 5, 5), 0.0, 4)]:
        s = (rng.random(shape) < frac).astype(np.int8)
        want = np.full(shape, float(margin))
        sol = np.argwhere(s != 0)
        if len(sol):
            idx = np.indices(shape).reshape(3, -1).T
            d2 = np.full(len(idx), np.inf)
            for p in sol:
                for img in np.ndindex(3, 3, 3):
                    q = p + (np.array(img) - 1) * np.array(shape)
                    d2 = np.minimum(d2, ((idx - q) ** 2).sum(1))
            want = np.minimum(np.sqrt(d2), margin).reshape(shape)
        np.testing.assert_array_equal(periodic_distance_edt(s, margin), want.astype(np.float32))

def test_inference_padding_reaches_multiples_of_eight():
    """Non-wrapping inference pads every axis by at least BORDER cells on both
    sides to a multiple of 8, for odd and even extents alike, and the periodic
    pad keeps the volume at its place."""
    import torch
    from taichi_lbm3d_amd.surrogate import BORDER, inference_pads, periodic_pad
    for shape in [(50, 50, 50), (51, 51, 51), (5, 6, 7), (8, 16, 24), (161, 3, 2)]:
        pads = inference_pads(shape, wrap=False)
        for n, (lo, hi) in zip(shape, pads):
            assert lo == BORDER and hi >= BORDER and (n + lo + hi) % 8 == 0
        t = torch.arange(int(np.prod(shape))).reshape(shape)
        u = periodic_pad(t, pads)
        assert torch.equal(u[BORDER:BORDER + shape[0], BORDER:BORDER + shape[1], BORDER:BORDER + shape[2]], t)
        assert torch.equal(u[0, BORDER:BORDER + shape[1], BORDER:BORDER + shape[2]], t[(-BORDER) % shape[0]])
    assert inference_pads((16, 16, 16), wrap=True) == [(0, 0)] * 3


def test_prediction_fields_keep_the_cold_start_mass():
    """fields_from_prediction inverts make_targets on the fluid; the default
    density is 1, and the predicted density (density=True) is shifted so the
    fluid's total mass is the cold start's, one per fluid cell."""
    import torch
    from taichi_lbm3d_amd.surrogate import fields_from_prediction, make_targets
    rng = np.random.default_rng(3)
    solid = (rng.random((6, 5, 4)) < 0.3).astype(np.int8)
    fx, nu = 1e-5, 1.0 / 6.0
    v = rng.normal(0.0, 1e-4, (6, 5, 4, 3)).astype(np.float32)
    rho = (1.0 + rng.normal(0.0, 3e-4, (6, 5, 4))).astype(np.float32)
    pred = torch.from_numpy(make_targets(rho, v, solid, fx, nu))
    fluid = torch.from_numpy(solid == 0)
    r1, v1, rp = fields_from_prediction(pred, fluid, fx, nu)
    assert torch.all(r1 == 1.0)
    fl = solid == 0
    np.testing.assert_allclose(v1.numpy()[fl], v[fl], rtol=1e-5, atol=1e-12)
    assert np.all(v1.numpy()[~fl] == 0.0)
    np.testing.assert_allclose(rp.numpy()[fl], rho[fl], rtol=0, atol=1e-6)
    r2, _, _ = fields_from_prediction(pred, fluid, fx, nu, density=True)
    r2 = r2.double().numpy()
    assert np.all(r2[~fl] == 1.0)
    assert abs(r2[fl].sum() - fl.sum()) < 1e-4
    # the deviation from the fluid mean is the predicted one
    np.testing.assert_allclose(r2[fl] - 1.0, rho[fl] - rho[fl].mean(), rtol=0, atol=1e-6)
