"""GPU tests of the batched solver: every sample of a batch evolves bit for
bit as the single-grid solver evolves it alone (whose parity with the oracle
the other GPU tests pin), whatever the launch shape, with or without a HIP
graph, and with retired samples left untouched."""
import numpy as np
import pytest

from test_gpu_parity import CASES, case_geometry, make_solver

pytestmark = pytest.mark.gpu

FACE_NAMES = ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")


def batch_case(name, shapes, seed=0, dtype="f32"):
    """Samples of uneven shapes sharing the physics and face modes of CASES[name]."""
    from taichi_lbm3d_amd.batch import BatchSolver
    _, frac, niu, force, modes, rho_bc, u_bc = CASES[name]
    rng = np.random.default_rng(seed)
    solids = []
    for shp in shapes:
        g = (rng.random(shp) < frac).astype(np.int8)
        g[(rng.random(shp) < 0.05) & (g == 1)] = -1
        solids.append(g)
    faces = {}
    for f in range(6):
        if modes[f] == 1:
            faces[FACE_NAMES[f]] = ("pressure", rho_bc[f])
        elif modes[f] == 2:
            faces[FACE_NAMES[f]] = ("velocity", u_bc[f])
    b = BatchSolver(solids, niu=niu, force=force, faces=faces, dtype=dtype)
    singles = [make_solver(s, niu, force, modes, rho_bc, u_bc, dtype=dtype) for s in solids]
    return b, singles


def populations(b, s, single):
    """(batch slice, single-solver array) of sample s's 19 population slots."""
    import torch
    from taichi_lbm3d_amd import _lib
    lib = _lib.lib()
    n = int(b.samples[s].n_cells)
    cb = int(b.samples[s].cell_begin)
    sb = lib.size("lbm3d_slot_stride", int(b.sizes.cells))
    ss = lib.size("lbm3d_slot_stride", n)
    gb = b._dev["g"].view(19, sb)[:, cb:cb + n]
    gs = single._dev["g"].view(19, ss)[:, :n]
    # raw bytes (numpy has no bfloat16)
    return gb.contiguous().view(torch.uint8).cpu().numpy(), gs.contiguous().view(torch.uint8).cpu().numpy()


def assert_same(b, singles):
    for s, single in enumerate(singles):
        gb, gs = populations(b, s, single)
        np.testing.assert_array_equal(gb, gs, err_msg=f"sample {s}")
        fu_b = b._dev["face_u"]
        if int(b.sizes.face_u):
            begin = int(b.samples[s].face_u_begin)
            n = single._face_u_floats
            np.testing.assert_array_equal(fu_b[begin:begin + n].cpu().numpy(),
                                          single._dev["face_u"][:n].cpu().numpy())


SHAPES = [(10, 9, 8), (7, 12, 33), (16, 5, 6), (9, 9, 40)]


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("steps", [1, 2, 7, 20])
def test_batch_bit_exact_vs_single(gpu, name, steps):
    b, singles = batch_case(name, SHAPES, seed=steps)
    b.run(steps)
    for s in singles:
        s.run(steps)
    assert_same(b, singles)
    # macros through the batch path equal the single solver's compact macros
    rho, v = b.macros()
    total = int(b.sizes.cells)
    for s, single in enumerate(singles):
        cb, n = int(b.samples[s].cell_begin), single.n_fluid
        r1, v1 = single.compact_macros()
        np.testing.assert_array_equal(rho[cb:cb + n].cpu().numpy(), r1[:n].cpu().numpy())
        vb = v.view(3, total)[:, cb:cb + n].cpu().numpy()
        np.testing.assert_array_equal(vb, v1[: 3 * n].view(3, n).cpu().numpy())


@pytest.mark.parametrize("shape", [(128, 0, 0), (256, 0, 1), (512, 0, 0), (128, 2, 1), (256, 4, 0), (512, 8, 1)])
def test_batch_launch_shapes_bit_exact(gpu, shape):
    """Every launch shape (block, persistent grid, sweep order) gives the same bits."""
    b, singles = batch_case("mixed_all_faces", SHAPES + [(20, 20, 20)], seed=3)
    b.set_launch(shape, (shape[0], shape[1], 1 - shape[2]))
    b.run(9)
    for s in singles:
        s.run(9)
    assert_same(b, singles)


def test_batch_graph_replay(gpu):
    """Steps replayed from a captured HIP graph equal eager steps, replay after
    replay, with eager work and synchronisations in between."""
    import torch
    a, _ = batch_case("periodic_force", SHAPES, seed=5)
    g, _ = batch_case("periodic_force", SHAPES, seed=5)
    for rnd in range(4):
        a.run(13)
        g.run(13, graph_steps=4)
        assert a.steps_done == g.steps_done
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a._dev["g"].view(torch.uint8).cpu().numpy(),
                                      g._dev["g"].view(torch.uint8).cpu().numpy(), err_msg=f"round {rnd}")
        ra, rg = a.reduce(), g.reduce()  # eager kernels between replays
        np.testing.assert_array_equal(ra, rg)


def test_batch_retire_freezes_sample(gpu):
    b, singles = batch_case("pressure_force", SHAPES, seed=9)
    b.run(6)
    for s in singles:
        s.run(6)
    frozen = populations(b, 1, singles[1])[0].copy()
    b.retire([False, True, False, False])
    b.run(10)
    for i, s in enumerate(singles):
        if i != 1:
            s.run(10)
    np.testing.assert_array_equal(populations(b, 1, singles[1])[0], frozen)
    for i in (0, 2, 3):
        gb, gs = populations(b, i, singles[i])
        np.testing.assert_array_equal(gb, gs)


def test_batch_reduce_matches_single(gpu):
    b, singles = batch_case("periodic_force", SHAPES, seed=2)
    b.run(15)
    r = b.reduce()
    for s, single in enumerate(singles):
        single.run(15)
        rs = single.reduce()
        assert int(r[s, 6]) == single.n_fluid
        assert r[s, 5] == 0
        np.testing.assert_allclose(r[s, 0], rs["sum_rho"], rtol=1e-13)
        np.testing.assert_allclose(r[s, 1:4], rs["sum_v"], rtol=1e-11, atol=1e-15)
        np.testing.assert_allclose(r[s, 4], rs["max_v"], rtol=1e-15)


@pytest.mark.parametrize("dtype", ["f16", "bf16", "f64", "f16sr", "bf16sr"])
def test_batch_storage_types_bit_exact(gpu, dtype):
    """Every storage type, stochastic rounding included: a sample's random bits
    depend on its own cell ids, not on where the batch put them."""
    b, singles = batch_case("pressure_x", SHAPES, seed=4, dtype=dtype)
    b.run(8)
    for s in singles:
        s.run(8)
    assert_same(b, singles)


@pytest.mark.parametrize("dtype", ["f16sr", "bf16sr"])
def test_batch_stochastic_rounding_on_both_parities_and_shapes(gpu, dtype):
    """Stochastic rounding under every launch shape and an odd step count:
    the batch still equals each sample run alone, bit for bit; a graph of such
    steps (whose captured seeds a replay would repeat) is refused."""
    b, singles = batch_case("periodic_force", SHAPES + [(20, 20, 20)], seed=6, dtype=dtype)
    b.set_launch((128, 2, 1), (512, 0, 0))
    b.run(11)
    for s in singles:
        s.run(11)
    assert_same(b, singles)
    with pytest.raises(ValueError, match="round-to-nearest"):
        b.run(20, graph_steps=4)


def test_batch_sample_fields_and_warm_start(gpu):
    b, singles = batch_case("periodic_force", SHAPES, seed=7)
    rng = np.random.default_rng(0)
    fields = []
    for s in b.solids:
        fields.append(((1 + 1e-3 * rng.standard_normal(s.shape)).astype(np.float32),
                       (1e-3 * rng.standard_normal(s.shape + (3,))).astype(np.float32)))
    b.init_from_macros(fields)
    for single, (r, v) in zip(singles, fields):
        single.init_from_macros(r, v)
    b.run(5)
    for i, single in enumerate(singles):
        single.run(5)
        rho, v = b.sample_fields(i)
        np.testing.assert_array_equal(rho, single.rho.to_numpy())
        np.testing.assert_array_equal(v, single.v.to_numpy())


def test_autotune_picks_valid_shapes(gpu, tmp_path, monkeypatch):
    from taichi_lbm3d_amd import autotune
    monkeypatch.setenv("TAICHI_LBM3D_TUNE_CACHE", "0")
    b, singles = batch_case("periodic_force", [(24, 24, 24), (30, 20, 16)], seed=1)
    res = autotune.tune(b, candidates=[(128, 0, 0), (256, 0, 1), (512, 2, 0)], pairs=2)
    for shape in (res["even"], res["odd"]):
        assert shape[0] in autotune.BLOCKS and shape[1] in autotune.GRID_PER_CU and shape[2] in (0, 1)
    assert len(res["table"]) == 3
    # the chosen pair is the fastest of the pairs timed together
    assert (res["even"], res["odd"]) == min(res["pairs"], key=lambda r: r[2])[:2]
    t = b.steps_done
    b.run(4)
    for s in singles:
        s.run(t + 4)
    assert_same(b, singles)


def test_batch_tile_descriptors_change_nothing(gpu):
    """Batched odd steps that take their windows from the per-sample tile
    descriptors give the same populations, bit for bit, as the cell-list path:
    packs with ragged rows, one with rows too long for the window (table
    path), samples of uneven size."""
    import torch
    from taichi_lbm3d_amd.batch import BatchSolver
    from taichi_lbm3d_amd.geometry import sphere_pack
    shapes = [(64, 48, 40), (40, 33, 70), (24, 20, 520), (30, 30, 30)]
    solids = [sphere_pack(sh, porosity=0.4, rmin=2, rmax=5, seed=7 + i, periodic=(True, True, True))
              for i, sh in enumerate(shapes)]
    out = []
    for tiles in (False, True):
        b = BatchSolver(solids, niu=0.1, force=(1e-5, 0.0, 0.0), use_tiles=tiles)
        assert (b.p.flags & 2) == (2 if tiles else 0)
        b.run(13)
        torch.cuda.synchronize()
        out.append(b._dev["g"].cpu())
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("dtype,visit", [("f32", 2), ("f32", 6), ("f16sr", 4), ("f64", 3)])
def test_batch_cache_blocked_schedule_bit_exact(gpu, dtype, visit):
    """Samples advanced group by group, several steps per visit
    (lbm3d_batch_run_range), end bit for bit where one launch per step over
    the whole batch leaves them: odd and even step counts, a group budget
    that splits the batch into several groups, graph replay for round-to-
    nearest storage, and retired samples (a whole group no longer launched)."""
    import torch
    from taichi_lbm3d_amd.batch import BatchSolver
    from taichi_lbm3d_amd.geometry import sphere_pack
    shapes = [(24, 20, 18), (30, 16, 22), (12, 40, 20), (26, 26, 26), (18, 18, 40)]
    solids = [sphere_pack(sh, porosity=0.4, rmin=2, rmax=5, seed=40 + i, periodic=(True, True, True))
              for i, sh in enumerate(shapes)]
    out = []
    for grouped in (False, True):
        item = {"f32": 4, "f16sr": 2, "f64": 8}[dtype]
        kw = dict(group_bytes=8000 * 19 * item, steps_per_visit=visit) if grouped else {}
        b = BatchSolver(solids, niu=0.1, force=(1e-5, 0.0, 0.0), dtype=dtype, **kw)
        if grouped:
            assert len(b.groups) >= 3, b.groups
        b.run(7)
        b.run(13)
        if dtype != "f16sr":
            b.run(24, graph_steps=8)
        else:
            b.run(24)
        # retire the first two samples (a whole group when grouped): they stay as they are
        b.retire([True, True, False, False, False])
        if grouped:
            assert len(b._live) < len(b.groups)
        b.run(10)
        torch.cuda.synchronize()
        out.append((b._dev["g"].view(torch.uint8).cpu(), b._dev["face_u"].cpu(), b.steps_done))
    assert out[0][2] == out[1][2] == 54
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])

