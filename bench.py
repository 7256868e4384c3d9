#!/usr/bin/env python3
"""Benchmark: sparse D3Q19 MRT on a 256^3 porous sphere pack, one MI355X.

Workload: 256^3 uint8 geometry from a seeded overlapping-sphere pack (radii
uniform in [6,14], spheres added until porosity <= 0.35, periodic in y and z),
fixed pressure on the x faces (rho 1.0005 left, 1.0 right), y/z periodic,
niu = 0.1, FP32 populations.  A step is one fused collide+stream launch over
the fluid list.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--dtype f32|f16|bf16|f64|f16sr|bf16sr]
                  [--no-steady] [--no-precision] [--no-cpu-baseline]

With FP32 storage the line also carries `precision`: every 16-bit storage
format on the same pack (MLUPS, steps and wall time to the steady criterion,
permeability error against an FP64 run to the same criterion).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, GB/s
BYTES_PER_FLUP = {"f64": 304, "f32": 152, "f16": 76, "bf16": 76, "f16sr": 76, "bf16sr": 76}  # 19 reads + 19 writes per fluid-cell update
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_traffic_pack256.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_geometry(n, seed):
    from taichi_lbm3d_amd.geometry import sphere_pack
    t0 = time.time()
    solid = sphere_pack((n, n, n), porosity=0.35, rmin=6.0, rmax=14.0, seed=seed, periodic=(False, True, True))
    log(f"geometry {n}^3 seed {seed}: porosity {(solid == 0).mean():.4f} ({time.time() - t0:.1f}s)")
    return solid


def make_solver(solid, dtype):
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    n = solid.shape
    s = LB3D_Solver_Single_Phase(*n, sparse_storage=True, dtype=dtype)
    s.solid.from_numpy(solid)
    s.niu = 0.1
    s.set_bc_rho_x0(1.0005)
    s.set_bc_rho_x1(1.0)
    s.init_simulation()
    return s


def cpu_model():
    """The host CPU as /proc/cpuinfo names it, with family/model/stepping and
    the logical CPUs this process may use."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                info[k] = v.strip()
    except OSError:
        pass
    return (f"{info.get('model name', 'unknown CPU')} (family {info.get('cpu family', '?')}, model "
            f"{info.get('model', '?')}, stepping {info.get('stepping', '?')}); {len(os.sched_getaffinity(0))} of "
            f"{os.cpu_count()} logical CPUs in this process's affinity mask")


def cpu_baseline(solid, warmup=50, planned=500, budget_s=55.0):
    """The C restatement of the reference's four-kernel step (sparse storage,
    FP32, OpenMP over this host's cores), built -O3 -march=native on this host
    and timed on a bounded step count."""
    import oracle
    from oracle import COracle
    flags = oracle.build_native()
    threads = len(os.sched_getaffinity(0))
    n = solid.shape
    o = COracle(*n, sparse_storage=True, real="f32native", threads=threads)
    o.set_solid(solid)
    o.niu = 0.1
    o.set_bc_rho("x_left", 1.0005)
    o.set_bc_rho("x_right", 1.0)
    o.init_simulation()
    n_fluid = int((solid == 0).sum())
    t0 = time.perf_counter()
    o.step(warmup)
    per_step = (time.perf_counter() - t0) / warmup
    steps = int(min(planned, max(20, budget_s / per_step)))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    why = "" if steps == planned else (f" ({planned} planned; cut to fit {budget_s:.0f}s of CPU time so the default "
                                       f"bench finishes within minutes)")
    return {"value": n_fluid * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{steps} timed steps after {warmup} warm-up steps of the {n[0]}^3 pack, {dt:.1f}s wall{why}; "
                      f"C/OpenMP restatement of the reference's colission/streaming1/Boundary_condition/streaming3 "
                      f"over active 3^3 blocks (not Taichi), gcc {flags}; CPU {cpu_model()}"}


def time_steps(s, steps, warmup):
    """Median-free HIP-event time of `steps` steps after `warmup` steps, in ms."""
    import torch
    s.run(warmup)
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    s.run(steps)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end)


def run_steady(solid, dtype, max_steps):
    import torch
    s = make_solver(solid, dtype)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps, ok, hist = s.run_to_steady(tol=1e-6, interval=500, max_steps=max_steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    k, kg = s.permeability(), s.permeability("gradient")
    del s
    torch.cuda.empty_cache()
    return {"steps": steps, "converged": ok, "wall_s": round(wall, 3), "permeability_lu2": k,
            "permeability_gradient_lu2": kg, "mean_vx": hist[-1]}


def precision_study(solid, f32_mlups, f32_steady, max_steps):
    """Each 16-bit storage format on the same pack: MLUPS (400 timed steps), its
    own run to the 1e-6 criterion, and the permeability against an FP64 run to
    the same criterion."""
    import torch
    ref = run_steady(solid, "f64", max_steps)
    k64 = ref["permeability_lu2"]
    out = {"reference": {"dtype": "f64", **ref}}
    rows = {"f32": (f32_mlups, f32_steady)}
    for dt in ("f16", "bf16", "f16sr", "bf16sr"):
        s = make_solver(solid, dt)
        ms = time_steps(s, 400, 100)
        del s
        torch.cuda.empty_cache()
        rows[dt] = (s_mlups(solid, ms, 400), run_steady(solid, dt, max_steps))
        log(f"precision {dt}: {rows[dt][0]:.0f} MLUPS, {rows[dt][1]['steps']} steps, {rows[dt][1]['wall_s']}s")
    for dt, (mlups, st) in rows.items():
        out[dt] = {"mlups": round(mlups, 1), "vs_f32_mlups": round(mlups / f32_mlups, 3),
                   "steady_steps": st["steps"], "converged": st["converged"], "steady_wall_s": st["wall_s"],
                   "permeability_lu2": st["permeability_lu2"],
                   "k_rel_err_vs_f64": st["permeability_lu2"] / k64 - 1.0}
    return out


def s_mlups(solid, ms, steps):
    return int((solid == 0).sum()) * steps / (ms / 1e3) / 1e6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "f64", "f16sr", "bf16sr"])
    ap.add_argument("--no-steady", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--max-steady-steps", type=int, default=50000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the solver is single-GPU; "
              "see the sharding design for the multi-GPU layout)", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    solid = make_geometry(args.n, args.seed)
    s = make_solver(solid, args.dtype)
    n_fluid = s.n_fluid
    band_rows = s.band_rows
    log(f"fluid cells {n_fluid}")

    ms = time_steps(s, args.steps, args.warmup)
    ms_per_step = ms / args.steps
    mlups = s_mlups(solid, ms, args.steps)
    r = s.reduce()
    if r["nonfinite"]:
        raise FloatingPointError("non-finite fields after the timed run")
    del s
    torch.cuda.empty_cache()
    bpf = BYTES_PER_FLUP[args.dtype]
    achieved = n_fluid * bpf / (ms_per_step / 1e3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and args.n == 256 and args.dtype == "f32":
        with open(TRAFFIC_FILE) as fh:
            tf = json.load(fh)
        traffic = tf.get("bytes_per_launch")
    log(f"timed {args.steps} steps: {ms:.1f} ms, {mlups:.0f} MLUPS, {achieved:.0f} GB/s algorithmic")

    steady = None
    if not args.no_steady:
        steady = run_steady(solid, args.dtype, args.max_steady_steps)
        steady["criterion"] = f"|d<vx>|/|<vx>| < 1e-6 over 500 steps, max {args.max_steady_steps}"
        log(f"steady: {steady['steps']} steps converged={steady['converged']} {steady['wall_s']:.2f}s "
            f"k={steady['permeability_lu2']:.5g}")

    precision = None
    if not args.no_precision and not args.no_steady and args.dtype == "f32":
        precision = precision_study(solid, mlups, steady, args.max_steady_steps)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(solid)
        if steady is not None:
            # CPU time to the same steady state, extrapolated from its measured step rate
            cpu["steady_s_extrapolated"] = round(steady["steps"] * n_fluid / (cpu["value"] * 1e6), 1)
        log(f"cpu baseline: {cpu['value']:.1f} MLUPS on {cpu['cores']} cores")

    schedule = "one launch per step"
    out = {
        "metric": "MLUPS + % HBM roofline, D3Q19 MRT 256^3 porous; time-to-steady permeability",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": f"sparse D3Q19 MRT, {args.n}^3 overlapping-sphere pack (seed {args.seed}, radii "
                               f"6-14, porosity {float((solid == 0).mean()):.4f}), pressure x faces 1.0005/1.0, "
                               f"y/z periodic, niu 0.1, reference convention",
                   "fluid_cells": n_fluid, "storage": args.dtype, "schedule": schedule,
                   "fluid_list": f"bands of {band_rows} grid rows" if 0 < band_rows < args.n else "C order",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_flup": bpf, "kernel": "k_step (even and odd updates alternate)"},
        "cpu_baseline": cpu,
        "steady": steady,
        "precision": precision,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
