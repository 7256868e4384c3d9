"""The batched data-generation workload at full size (GPU): the 64 seeded
periodic sphere packs (edges 48-160, porosity 0.25-0.45, radii 4-12, fx 1e-5,
niu 1/6, FP32) stepped in one launch per step with autotuned launch shapes
and a HIP-graph-captured step loop.  Every sample must equal, bit for bit,
its own run on the single-grid solver, and the smallest samples must match the
FP64 CPU oracle (the reference's four-kernel step, SP:477-481, restated) to
the FP32 field tolerance."""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 400
THREADS = min(16, len(os.sched_getaffinity(0)))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def packs():
    from taichi_lbm3d_amd.batch_bench import make_samples
    t0 = time.time()
    solids, specs = make_samples(64, workers=1)
    print(f"64 packs generated in {time.time() - t0:.1f}s")
    return solids, specs


def test_batch_of_64_equals_single_grid_and_oracle(gpu, packs, monkeypatch):
    import torch
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase, autotune, _lib
    from taichi_lbm3d_amd.batch import BatchSolver
    from taichi_lbm3d_amd.batch_bench import FX, NIU
    from oracle import COracle
    monkeypatch.setenv("TAICHI_LBM3D_TUNE_CACHE", "0")
    solids, specs = packs
    assert len(solids) == 64 and min(s[0] for s in specs) >= 48 and max(s[0] for s in specs) <= 160
    t0 = time.time()
    b = BatchSolver(solids, niu=NIU, force=(FX, 0.0, 0.0))
    tuned = autotune.tune(b, pairs=2, use_cache=False)  # advances the batch by real steps
    assert b.steps_done % 2 == 0 and b.steps_done < STEPS - 100
    b.run(STEPS - 100 - b.steps_done)
    b.run(100, graph_steps=50)  # the last 100 steps replayed from a captured graph
    torch.cuda.synchronize()
    assert b.steps_done == STEPS
    print(f"batch: {int(b.n_fluid.sum())} fluid cells, tuned even {tuned['even']} odd {tuned['odd']}, "
          f"{time.time() - t0:.1f}s")

    lib = _lib.lib()
    sb = lib.size("lbm3d_slot_stride", int(b.sizes.cells))
    gb = b._dev["g"].view(19, sb)
    t0 = time.time()
    for s, solid in enumerate(solids):
        one = LB3D_Solver_Single_Phase(*solid.shape, sparse_storage=True)
        one.solid.from_numpy(solid)
        one.niu = NIU
        one.set_force([FX, 0.0, 0.0])
        one.init_simulation()
        one.run(STEPS)
        n, cb = int(b.samples[s].n_cells), int(b.samples[s].cell_begin)
        assert n == one.n_fluid
        ss = lib.size("lbm3d_slot_stride", n)
        assert torch.equal(gb[:, cb:cb + n], one._dev["g"].view(19, ss)[:, :n]), f"sample {s}"
        del one
    print(f"64 single-grid runs compared in {time.time() - t0:.1f}s")

    t0 = time.time()
    smallest = sorted(range(64), key=lambda s: solids[s].size)[:4]
    for s in smallest:
        solid = solids[s]
        o = COracle(*solid.shape, real="f64", threads=THREADS)
        o.set_solid(solid)
        o.niu = NIU
        o.fx = FX
        o.init_simulation()
        o.step(STEPS)
        rho, v = b.sample_fields(s)
        fl = solid == 0
        er, ev = rel_l2(rho[fl], o.rho()[fl]), rel_l2(v[fl], o.v()[fl])
        print(f"sample {s} ({solid.shape[0]}^3): rel-L2 rho {er:.2e} v {ev:.2e}")
        assert er < 1e-6 and ev < 1e-5
    print(f"oracle comparisons {time.time() - t0:.1f}s")
