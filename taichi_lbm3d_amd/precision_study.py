"""Storage-precision study: FP32 / FP16 / BF16 population storage against an
FP64 run of the same GPU kernels.

    python taichi_lbm3d_amd/precision_study.py --domain pack256 [--dtypes f32,f16,bf16]
    python taichi_lbm3d_amd/precision_study.py --domain channel512 [--steps 4000]

Every run uses the same solver and kernels; only the storage type of the
shifted populations g = f - w changes (FP32 arithmetic for f32/f16/bf16, FP64
storage and arithmetic for the reference run "f64").

Domains:
  pack256     256^3 overlapping-sphere pack (seed 0, radii 6-14, porosity 0.35,
              periodic in y and z), pressure x faces rho 1.0005 / 1.0, niu 0.1.
              Each format runs to its own steady state (|d<vx>|/|<vx>| < 1e-6
              over 500 steps, at most 50k steps); the steady fields and Darcy
              permeabilities are compared.
  channel512  512^3 channel periodic in x with bounce-back walls at y and z and
              seeded random solid cubes (edges 4-16) covering 5 % of the volume,
              body force fx = 1e-6, niu 0.5 (so that the steady centre-line
              velocity F H^2 / (8 niu) stays near 0.065).  Steady state is
              millions of steps away here, so every format runs the same fixed
              number of steps from rest and the fields at that step are compared;
              "permeability" is then the apparent one, niu <u_x> / fx, at that step.

For each format the study prints one JSON line: velocity relative L2 error on
fluid cells, density max error, permeability relative error, steps, wall time
and MLUPS of a separately timed run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

if __package__ in (None, ""):  # run as a file
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

BYTES_PER_FLUP = {"f64": 304, "f32": 152, "f16": 76, "bf16": 76, "f16sr": 76, "bf16sr": 76}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pack256(n=256, seed=0):
    from taichi_lbm3d_amd.geometry import sphere_pack
    solid = sphere_pack((n, n, n), porosity=0.35, rmin=6.0, rmax=14.0, seed=seed, periodic=(False, True, True))

    def setup(s):
        s.niu = 0.1
        s.set_bc_rho_x0(1.0005)
        s.set_bc_rho_x1(1.0)
    return solid, setup, True


def channel512(n=512, seed=0):
    from taichi_lbm3d_amd.geometry import cube_channel
    solid = cube_channel((n, n, n), fraction=0.05, emin=4, emax=16, seed=seed)

    def setup(s):
        s.niu = 0.5
        s.fx = 1e-6
    return solid, setup, False


DOMAINS = {"pack256": pack256, "channel512": channel512}


def run_format(solid, setup, sparse, dtype, steady, steps, timed_steps):
    import torch
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    s = LB3D_Solver_Single_Phase(*solid.shape, sparse_storage=sparse, dtype=dtype)
    s.solid.from_numpy(solid)
    setup(s)
    s.init_simulation()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if steady:
        n, ok, _ = s.run_to_steady(tol=1e-6, interval=500, max_steps=steps)
    else:
        s.run(steps)
        n, ok = steps, None
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    r = s.reduce()
    if r["nonfinite"]:
        log(f"{dtype}: {r['nonfinite']} non-finite cells after {n} steps")
    k = s.permeability() if not r["nonfinite"] else float("nan")
    rho, v = s.compact_macros(native=True)
    rho = rho.double()
    v = v.double()
    # throughput on the evolved state, continuing from it (an even number of steps
    # keeps the AA parity of the compared state irrelevant: fields are taken above)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    s.run(timed_steps)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    mlups = s.n_fluid * timed_steps / (ms / 1e3) / 1e6
    res = {"dtype": dtype, "steps": n, "converged": ok, "wall_s": round(wall, 3), "permeability": k,
           "mlups": round(mlups, 1), "hbm_gbs": round(mlups * BYTES_PER_FLUP[dtype] / 1e3, 1),
           "nonfinite": r["nonfinite"], "fluid_cells": s.n_fluid}
    del s
    torch.cuda.empty_cache()
    return res, rho, v


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--domain", choices=sorted(DOMAINS), required=True)
    ap.add_argument("--n", type=int, default=None, help="cube edge (default 256 / 512)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--dtypes", default="f32,f16,bf16,f16sr,bf16sr")
    ap.add_argument("--steps", type=int, default=None,
                    help="max steps to steady (pack256, default 50000) or fixed steps (channel512, default 4000)")
    ap.add_argument("--timed-steps", type=int, default=400)
    args = ap.parse_args()

    import torch
    assert torch.cuda.is_available(), "the precision study runs on the GPU"
    steady = args.domain == "pack256"
    steps = args.steps or (50_000 if steady else 4000)
    t0 = time.time()
    kw = {"seed": args.seed}
    if args.n:
        kw["n"] = args.n
    solid, setup, sparse = DOMAINS[args.domain](**kw)
    log(f"{args.domain}: geometry {solid.shape}, porosity {float((solid == 0).mean()):.4f} "
        f"({time.time() - t0:.1f}s)")

    ref, rho64, v64 = run_format(solid, setup, sparse, "f64", steady, steps, args.timed_steps)
    ref["domain"] = args.domain
    print(json.dumps(ref), flush=True)
    n = ref["fluid_cells"]
    vnorm = torch.linalg.vector_norm(v64).item()
    for dtype in [d for d in args.dtypes.split(",") if d and d != "f64"]:
        res, rho, v = run_format(solid, setup, sparse, dtype, steady, steps, args.timed_steps)
        res["domain"] = args.domain
        res["v_rel_l2"] = torch.linalg.vector_norm(v[: 3 * n] - v64[: 3 * n]).item() / max(vnorm, 1e-300)
        res["rho_max_abs_err"] = (rho[:n] - rho64[:n]).abs().max().item()
        res["permeability_rel_err"] = (res["permeability"] - ref["permeability"]) / ref["permeability"]
        print(json.dumps(res), flush=True)
        del rho, v
        torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
