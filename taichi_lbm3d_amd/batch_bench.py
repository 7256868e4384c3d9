"""Batched data-generation workload: 64 seeded sphere packs of uneven size.

    python taichi_lbm3d_amd/batch_bench.py [--samples 64] [--steps 1000] [--steady]
                                            [--save-fields PATH]

Sample i (seed base + i): cube edge L uniform in {48, ..., 160}, porosity
uniform in [0.25, 0.45], sphere radii uniform in [4, 12], every axis periodic.
Body force fx = 1e-5, all faces periodic, niu = 1/6, FP32 storage.

Reports the batch's aggregate MLUPS (one launch per step for all samples, the
launch shapes picked by the autotuner, the loop replayed from a HIP graph)
against the same samples stepped one by one with the single-grid solver, and
with --steady runs every sample to |d<vx>|/|<vx>| < 1e-6 over 500 steps (each
retired as it converges) and reports steps and permeabilities; --save-fields
writes the converged fields (the surrogate's training data) to an .npz.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

if __package__ in (None, ""):  # run as a file
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

BYTES_PER_FLUP = 152
NIU = 1.0 / 6.0
FX = 1e-5


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sample_spec(i: int, base_seed: int = 1000):
    rng = np.random.default_rng(base_seed + i)
    edge = int(rng.integers(48, 161))
    porosity = float(rng.uniform(0.25, 0.45))
    return edge, porosity, base_seed + i


def _make(spec):
    from taichi_lbm3d_amd.geometry import sphere_pack
    edge, porosity, seed = spec
    return sphere_pack((edge,) * 3, porosity=porosity, rmin=4.0, rmax=12.0, seed=seed, periodic=(True, True, True))


def make_samples(n: int, base_seed: int = 1000, workers: int | None = None):
    specs = [sample_spec(i, base_seed) for i in range(n)]
    workers = workers or min(16, len(os.sched_getaffinity(0)))
    if workers > 1:
        with ProcessPoolExecutor(workers) as ex:
            solids = list(ex.map(_make, specs))
    else:
        solids = [_make(s) for s in specs]
    return solids, specs


def time_batch(b, steps: int, graph_steps: int):
    import torch
    stream = torch.cuda.current_stream(b.device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    b.run(steps, graph_steps=graph_steps)
    e1.record(stream)
    torch.cuda.synchronize(b.device)
    return e0.elapsed_time(e1)


def time_one_by_one(solids, steps: int, warmup: int = 20):
    """Each sample alone on the single-grid solver: summed event time of `steps` steps."""
    import torch
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    total_ms, flups = 0.0, 0
    for solid in solids:
        s = LB3D_Solver_Single_Phase(*solid.shape, sparse_storage=True)
        s.solid.from_numpy(solid)
        s.niu = NIU
        s.set_force([FX, 0.0, 0.0])
        s.init_simulation()
        s.run(warmup)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        s.run(steps)
        e1.record()
        torch.cuda.synchronize()
        total_ms += e0.elapsed_time(e1)
        flups += s.n_fluid * steps
        del s
    return total_ms, flups


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--graph-steps", type=int, default=100)
    ap.add_argument("--single-steps", type=int, default=200)
    ap.add_argument("--steady", action="store_true")
    ap.add_argument("--max-steady-steps", type=int, default=50000)
    ap.add_argument("--save-fields", default="")
    args = ap.parse_args(argv)

    import torch
    from taichi_lbm3d_amd import autotune
    from taichi_lbm3d_amd.batch import BatchSolver
    assert torch.cuda.is_available(), "the batch workload runs on the GPU"
    t0 = time.time()
    solids, specs = make_samples(args.samples, args.seed)
    log(f"{len(solids)} samples generated ({time.time() - t0:.1f}s); edges {min(s[0] for s in specs)}-"
        f"{max(s[0] for s in specs)}, cells {sum(s[0] ** 3 for s in specs)}")

    b = BatchSolver(solids, niu=NIU, force=(FX, 0.0, 0.0))
    n_fluid = int(b.n_fluid.sum())
    log(f"batch: {n_fluid} fluid cells, {int(b.sizes.cells)} ids")
    b.run(args.warmup)
    tuned = autotune.tune(b, pairs=3, use_cache=False)
    log(f"autotune: even {tuned['even']} odd {tuned['odd']}")
    b.run(2)
    ms_eager = time_batch(b, args.steps, 0)
    ms_graph = time_batch(b, args.steps, args.graph_steps)
    b.set_launch((256, 0, 1), (256, 0, 0))
    ms_default = time_batch(b, args.steps, args.graph_steps)
    b.set_launch(tuned["even"], tuned["odd"])
    mlups = n_fluid * args.steps / (ms_graph / 1e3) / 1e6
    log(f"batched: graph {ms_graph / args.steps * 1e3:.1f} us/step ({mlups:.0f} MLUPS), eager "
        f"{ms_eager / args.steps * 1e3:.1f} us/step, default shapes {ms_default / args.steps * 1e3:.1f} us/step")

    ms_single, flups_single = time_one_by_one(solids, args.single_steps)
    mlups_single = flups_single / (ms_single / 1e3) / 1e6
    log(f"one by one: {mlups_single:.0f} MLUPS")

    out = {
        "metric": "aggregate MLUPS, batched D3Q19 MRT over 64 uneven sphere packs",
        "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "dtype": "f32",
        "ms_per_step": round(ms_graph / args.steps, 5),
        "config": {"workload": f"{args.samples} sphere packs, edges 48-160, porosity 0.25-0.45, radii 4-12, "
                               f"all periodic, fx {FX}, niu 1/6 (seeds {args.seed}..{args.seed + args.samples - 1})",
                   "fluid_cells": n_fluid, "graph_steps": args.graph_steps},
        "roofline": {"bound": "hbm", "achieved": round(n_fluid * BYTES_PER_FLUP / (ms_graph / args.steps / 1e3) / 1e9, 1),
                     "peak": 8000.0, "unit": "GB/s"},
        "launch": {"even": list(tuned["even"]), "odd": list(tuned["odd"]),
                   "table_us": [[list(s), round(e, 1), round(o, 1)] for s, e, o in tuned["table"]]},
        "eager_ms_per_step": round(ms_eager / args.steps, 5),
        "default_shape_ms_per_step": round(ms_default / args.steps, 5),
        "one_by_one": {"value": round(mlups_single, 1), "unit": "MLUPS", "steps_per_sample": args.single_steps},
        "speedup_vs_one_by_one": round(mlups / mlups_single, 3),
    }
    out["roofline"]["frac"] = round(out["roofline"]["achieved"] / 8000.0, 4)

    if args.steady:
        b2 = BatchSolver(solids, niu=NIU, force=(FX, 0.0, 0.0))
        b2.set_launch(tuned["even"], tuned["odd"])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        steps, conv = b2.run_to_steady(tol=1e-6, interval=500, max_steps=args.max_steady_steps,
                                       graph_steps=args.graph_steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        k = b2.permeability()
        out["steady"] = {"wall_s": round(wall, 2), "converged": int(conv.sum()), "steps_min": int(steps.min()),
                         "steps_median": int(np.median(steps)), "steps_max": int(steps.max()),
                         "permeability_median": float(np.median(k))}
        log(f"steady: {int(conv.sum())}/{len(conv)} converged in {wall:.1f}s, steps {steps.min()}-{steps.max()}")
        if args.save_fields:
            data = {}
            for s in range(b2.n_samples):
                rho, v = b2.sample_fields(s)
                data[f"solid_{s}"] = solids[s]
                data[f"rho_{s}"] = rho.astype(np.float32)
                data[f"v_{s}"] = v.astype(np.float32)
            data["steps"] = steps
            data["converged"] = conv
            data["permeability"] = k
            data["specs"] = np.array(specs, dtype=np.float64)
            np.savez(args.save_fields, **data)
            log(f"fields saved to {args.save_fields}")
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
