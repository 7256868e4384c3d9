"""Short run of the step kernels for hardware-counter profiles.

    python taichi_lbm3d_amd/profile_step.py [--n 256] [--steps 20] [--dtype f32]
                                            [--calib-copy-mib 2048] [--band-rows R]

Runs the 256^3 sphere-pack workload of bench.py (seed 0, pressure x faces,
niu 0.1) for a few steps and prints the per-kernel time of each parity.
With --calib-copy-mib it first copies a buffer of that size three times with
torch's device copy: a plain copy of known size, against which counters such
as FETCH_SIZE / WRITE_SIZE are scaled.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

if __package__ in (None, ""):  # run as a file (e.g. under rocprofv3 from another directory)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "f16sr", "bf16sr"])
    ap.add_argument("--calib-copy-mib", type=int, default=0)
    ap.add_argument("--band-rows", type=int, default=None, help="fluid-list bands (default: the solver's)")
    args = ap.parse_args()

    import torch
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    from taichi_lbm3d_amd.geometry import sphere_pack

    if args.calib_copy_mib:
        n = args.calib_copy_mib * 1024 * 1024 // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_mib": args.calib_copy_mib, "dispatches": 3}), flush=True)
        del a, b
        torch.cuda.empty_cache()

    solid = sphere_pack((args.n,) * 3, porosity=0.35, rmin=6.0, rmax=14.0, seed=args.seed,
                        periodic=(False, True, True))
    s = LB3D_Solver_Single_Phase(*solid.shape, sparse_storage=True, dtype=args.dtype)
    if args.band_rows is not None:
        s.band_rows = args.band_rows
    s.solid.from_numpy(solid)
    s.niu = 0.1
    s.set_bc_rho_x0(1.0005)
    s.set_bc_rho_x1(1.0)
    s.init_simulation()
    s.run(args.warmup)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    ev[0].record()
    for i in range(args.steps):
        s.step()
        ev[i + 1].record()
    torch.cuda.synchronize()
    dts = np.array([ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]) * 1e3
    par = np.array([(args.warmup + i) % 2 for i in range(args.steps)])
    print(json.dumps({"fluid_cells": s.n_fluid, "dtype": args.dtype, "band_rows": s.band_rows, "steps": args.steps,
                      "even_us": float(np.median(dts[par == 0])), "odd_us": float(np.median(dts[par == 1]))}))


if __name__ == "__main__":
    main()
