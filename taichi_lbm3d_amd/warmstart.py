"""Learned warm start, end to end on one GPU.

    python taichi_lbm3d_amd/warmstart.py [--train-samples 64] [--iters 32000] [--batch 2]
                                          [--crop 128] [--eval-samples 32] [--eval-size 192]
                                          [--save-model PATH] [--out PATH]

1. Training data: the batched workload's 64 seeded periodic sphere packs
   (edges 48-160, porosity 0.25-0.45, radii 4-12, fx = 1e-5, niu = 1/6) run to
   steady state by the batched solver; their converged fields are the targets.
2. Training: the 3-D U-Net of surrogate.py with Adam (lr 1e-3, cosine decay)
   on random periodic 128^3 crops of those samples (a periodic sample tiles
   space, so a crop may wrap; y/z flips and the y-z swap are applied at
   random), masked MSE on the fluid cells away from the crop faces, `iters`
   steps of `batch` crops.
3. Evaluation: held-out seeded 192^3 packs (porosity 0.3, radii 4-12).  Cold
   start (rest state) and warm start (f at the equilibrium of the predicted
   fields) each run to the 1e-6 steady criterion in one batched launch;
   reported: steps, wall-clock (the warm side includes inference), the
   permeability error of the warm result against the cold one, the steps and
   time until each sample's permeability is within 1 % of its cold-start value,
   and the samples that fell back to a cold start (a non-finite prediction or
   a run that does not converge).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):  # run as a file
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from taichi_lbm3d_amd import surrogate as sg  # noqa: E402
from taichi_lbm3d_amd.batch_bench import FX, NIU, make_samples  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def converge(solids, graph_steps=50, max_steps=50_000, tol=1e-6, init_fields=None, check_every=50):
    """Run a batch to steady state (criterion over 500 steps, checked every
    `check_every`); returns (batch, steps, converged, wall_s, history of
    (t, wall_s, permeability per sample)), the history starting at t = 0."""
    import torch
    from taichi_lbm3d_amd.batch import BatchSolver
    b = BatchSolver(solids, niu=NIU, force=(FX, 0.0, 0.0))
    if init_fields is not None:
        b.init_from_macros(init_fields)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = [(0, 0.0, b.permeability())]

    def cb(t, r):
        hist.append((t, time.perf_counter() - t0, b.permeability(r)))
        if t % 5000 == 0:
            log(f"  step {t}: {time.perf_counter() - t0:.1f}s")

    steps, conv = b.run_to_steady(tol=tol, interval=500, max_steps=max_steps, graph_steps=graph_steps, callback=cb,
                                  check_every=check_every)
    torch.cuda.synchronize()
    return b, steps, conv, time.perf_counter() - t0, hist


def initial_permeability(net, solids, nu):
    """Per sample: the permeability of the surrogate's warm start at t = 0
    (samples of any extents, predicted one at a time)."""
    from taichi_lbm3d_amd.batch import BatchSolver
    preds = []
    for s in solids:
        preds += sg.predict(net, [s], FX, nu, batch=1)
    b = BatchSolver(solids, niu=NIU, force=(FX, 0.0, 0.0))
    b.init_from_macros(preds)
    k = b.permeability()
    del b, preds
    return k


def within(hist, k_ref, rel=0.01):
    """Per sample: first (step, wall_s) from which |k - k_ref| / k_ref stays below rel."""
    n = len(k_ref)
    steps = np.full(n, -1, np.int64)
    walls = np.full(n, np.nan)
    for s in range(n):
        for idx in range(len(hist)):
            if all(abs(h[2][s] - k_ref[s]) <= rel * abs(k_ref[s]) for h in hist[idx:]):
                steps[s], walls[s] = hist[idx][0], hist[idx][1]
                break
    return steps, walls


class CropSampler:
    """Random periodic crops of converged samples, on the GPU."""

    def __init__(self, inputs, targets, solids, device, seed=0):
        import torch
        self.torch = torch
        self.inp = [i.to(device) if torch.is_tensor(i) else torch.from_numpy(i).to(device) for i in inputs]
        self.tgt = [torch.from_numpy(t).to(device) for t in targets]
        self.fluid = [torch.from_numpy((s == 0).astype(np.float32)).to(device) for s in solids]
        self.device = device
        self.rng = np.random.default_rng(seed)

    def crop(self, size):
        torch = self.torch
        s = int(self.rng.integers(len(self.inp)))
        L = self.inp[s].shape[:3]
        idx = [((torch.arange(size, device=self.device) + int(self.rng.integers(L[a]))) % L[a]) for a in range(3)]
        x = self.inp[s][idx[0]][:, idx[1]][:, :, idx[2]]
        y = self.tgt[s][idx[0]][:, idx[1]][:, :, idx[2]]
        m = self.fluid[s][idx[0]][:, idx[1]][:, :, idx[2]]
        if self.rng.random() < 0.5:  # y flip: vy changes sign
            x, y, m = x.flip(1), y.flip(1), m.flip(1)
            y = y * torch.tensor([1.0, -1.0, 1.0, 1.0], device=self.device)
        if self.rng.random() < 0.5:  # z flip: vz changes sign
            x, y, m = x.flip(2), y.flip(2), m.flip(2)
            y = y * torch.tensor([1.0, 1.0, -1.0, 1.0], device=self.device)
        if self.rng.random() < 0.5:  # y-z swap: vy and vz swap
            x, y, m = x.transpose(1, 2), y.transpose(1, 2), m.transpose(1, 2)
            y = y[..., [0, 2, 1, 3]]
        return x.contiguous(), y.contiguous(), m.contiguous()

    def batch(self, n, size):
        torch = self.torch
        xs, ys, ms = zip(*(self.crop(size) for _ in range(n)))
        x = torch.zeros((n, size, size, size, sg.CIN_PAD), dtype=torch.bfloat16, device=self.device)
        x[..., :2] = torch.stack(xs).to(torch.bfloat16)
        return x, torch.stack(ys), torch.stack(ms)


def train(net, sampler, iters, batch, crop, lr=1e-3, log_every=50, border=sg.BORDER, rho_weight=1.0):
    """Adam from lr with a cosine decay to lr/100; masked MSE over the fluid cells
    at least `border` cells inside the crop (the network's zero padding at a
    crop face stands where the true field goes on; inference pads periodically).
    The density channel's squared error is weighted by rho_weight (the warm
    start itself uses the velocity only)."""
    import math
    import torch
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    losses = []
    chw = torch.tensor([1.0, 1.0, 1.0, rho_weight], device=net.device)
    inner = torch.zeros((crop,) * 3, device=net.device)
    inner[border:crop - border, border:crop - border, border:crop - border] = 1.0
    t0 = time.perf_counter()
    for it in range(iters):
        for group in opt.param_groups:
            group["lr"] = lr * (0.01 + 0.99 * 0.5 * (1.0 + math.cos(math.pi * it / max(iters, 1))))
        x, y, m = sampler.batch(batch, crop)
        m = m * inner
        out = net(x)
        err = (out - y) ** 2 * m[..., None] * chw
        loss = err.sum() / (m.sum() * chw.sum() + 1e-9)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        net.zero_head_padding()
        losses.append(float(loss.detach()) if (it % log_every == 0 or it == iters - 1) else None)
        if it % log_every == 0 or it == iters - 1:
            torch.cuda.synchronize()
            log(f"iter {it}: loss {losses[-1]:.4g} ({time.perf_counter() - t0:.1f}s)")
    torch.cuda.synchronize()
    return [(i, l) for i, l in enumerate(losses) if l is not None], time.perf_counter() - t0


def training_data(solids, max_steps=50_000):
    """The batched solver's converged fields of `solids` as network inputs and
    targets; returns (inputs, targets, k_train, nu, info)."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    b, steps, conv, wall, _ = converge(solids, max_steps=max_steps)
    log(f"training data: {int(conv.sum())}/{len(conv)} converged, steps {steps.min()}-{steps.max()}, {wall:.1f}s")
    nu = b.effective_viscosity()
    k_train = b.permeability()
    fields = [b.sample_fields(s) for s in range(b.n_samples)]
    del b
    torch.cuda.empty_cache()
    inputs = [sg.make_inputs(s, device=dev) for s in solids]
    targets = [sg.make_targets(r, v, s, FX, nu) for (r, v), s in zip(fields, solids)]
    info = {"samples": len(solids), "converged": int(conv.sum()), "steps_max": int(steps.max()),
            "solver_wall_s": round(wall, 2)}
    return inputs, targets, k_train, nu, info


def velocity_scale(net, solids, k_train, nu):
    """Median ratio of the training samples' converged permeability to the
    permeability of the network's own warm start at t = 0."""
    k_tr_pred = initial_permeability(net, solids, nu)
    scale = float(np.median(k_train / k_tr_pred))
    log(f"calibration: velocity scale {scale:.4f} (training k ratio {np.min(k_train / k_tr_pred):.3f}"
        f"-{np.max(k_train / k_tr_pred):.3f})")
    return scale


def evaluate(net, ev, nu, scale=1.0, density=False, max_steps=50_000):
    """Cold and warm starts of the packs `ev`, each batch run to the 1e-6
    steady criterion; returns the report dict (steps, wall-clock with the warm
    side including inference, permeability errors, steps and time to within
    1 % of the cold permeability, fallbacks) with the per-sample arrays under
    "per_sample"."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    bc, steps_c, conv_c, wall_c, hist_c = converge(ev, max_steps=max_steps)
    k_cold = bc.permeability()
    ref = [bc.sample_fields(s) for s in range(bc.n_samples)]
    del bc
    torch.cuda.empty_cache()
    log(f"cold: {int(conv_c.sum())} converged, steps median {np.median(steps_c):.0f}, {wall_c:.1f}s")

    # the geometry goes to the device once, as the solver's own copy does, before the clock
    ev_dev = [torch.from_numpy(np.ascontiguousarray(s)).to(dev) for s in ev]
    torch.cuda.synchronize()
    t_inf = time.perf_counter()
    preds = sg.predict(net, ev_dev, FX, nu, density=density)
    if scale != 1.0:
        preds = [(r, v * scale) for r, v in preds]
    torch.cuda.synchronize()
    inf_s = time.perf_counter() - t_inf
    fallback = np.array([not (torch.isfinite(r).all() and torch.isfinite(v).all()) for r, v in preds])
    preds = [(torch.ones_like(r), torch.zeros_like(v)) if fb else (r, v) for (r, v), fb in zip(preds, fallback)]
    # the surrogate's error, judged by the converged cold-start fields
    v_err, rho_err = [], []
    for (r, v), (rr, vr), s in zip(preds, ref, ev):
        fl = s == 0
        vp = v.cpu().numpy()
        v_err.append(float(np.linalg.norm(vp[fl] - vr[fl]) / np.linalg.norm(vr[fl])))
        # density deviation from the sample's fluid mean, against the cold start's
        dr, dp = rr[fl] - rr[fl].mean(), r.cpu().numpy()[fl] - 1.0
        rho_err.append(float(np.linalg.norm(dp - dr) / np.linalg.norm(dr)))
    try:
        bw, steps_w, conv_w, wall_w, hist_w = converge(ev, max_steps=max_steps, init_fields=preds)
        k_warm = bw.permeability()
        del bw
    except FloatingPointError as e:  # the warm run diverged: every sample falls back to the cold start
        log(f"warm start diverged ({e}); falling back to cold start")
        fallback[:] = True
        steps_w, conv_w, wall_w, hist_w, k_warm = steps_c, conv_c, wall_c, hist_c, k_cold
    fallback |= ~conv_w
    k_warm = np.where(fallback, k_cold, k_warm)
    k_pred = hist_w[0][2]
    signed = (k_pred - k_cold) / np.abs(k_cold)
    s1_c, w1_c = within(hist_c, k_cold)
    s1_w, w1_w = within(hist_w, k_cold)
    err = np.abs(k_warm - k_cold) / np.abs(k_cold)
    perr = np.abs(k_pred - k_cold) / np.abs(k_cold)
    return {
        "samples": len(ev), "size": int(ev[0].shape[0]), "fallbacks": int(fallback.sum()),
        "surrogate": {"v_rel_l2_median": float(np.median(v_err)), "v_rel_l2_max": float(np.max(v_err)),
                      "rho_dev_rel_l2_median": float(np.median(rho_err)),
                      "k_rel_err_median": float(np.median(perr)), "k_rel_err_max": float(np.max(perr)),
                      "k_signed_err_median": float(np.median(signed)), "velocity_scale": scale,
                      "inference_s": round(inf_s, 3)},
        "cold": {"steps_median": float(np.median(steps_c)), "steps_max": int(steps_c.max()),
                 "wall_s": round(wall_c, 2), "converged": int(conv_c.sum()),
                 "steps_to_1pct_median": float(np.median(s1_c)), "steps_to_1pct_max": int(s1_c.max()),
                 "wall_to_1pct_s": round(float(np.nanmax(w1_c)), 3)},
        "warm": {"steps_median": float(np.median(steps_w)), "steps_max": int(steps_w.max()),
                 "wall_s": round(wall_w + inf_s, 2), "converged": int(conv_w.sum()),
                 "steps_to_1pct_median": float(np.median(s1_w)), "steps_to_1pct_max": int(s1_w.max()),
                 "wall_to_1pct_s": round(float(np.nanmax(w1_w)) + inf_s, 3)},
        "permeability_rel_err_warm_vs_cold": {"median": float(np.median(err)), "max": float(err.max())},
        "speedup_to_steady": round(wall_c / (wall_w + inf_s), 3),
        "speedup_to_1pct": round(float(np.nanmax(w1_c)) / (float(np.nanmax(w1_w)) + inf_s), 3),
        # the same for the median sample: the batch's wall-clock when that
        # sample came within 1 % (cold), against inference plus the same (warm)
        "speedup_to_1pct_median_sample": round(float(np.nanmedian(w1_c)) / (float(np.nanmedian(w1_w)) + inf_s), 3),
        "per_sample": {"k_cold": k_cold.tolist(), "k_warm": k_warm.tolist(), "steps_cold": steps_c.tolist(),
                       "steps_warm": steps_w.tolist(), "steps_to_1pct_cold": s1_c.tolist(),
                       "steps_to_1pct_warm": s1_w.tolist(), "fallback": fallback.tolist()},
    }


def eval_packs(n, size, seed=5000):
    """The held-out packs: periodic, porosity 0.3, radii 4-12, seeds seed .. seed+n-1."""
    from taichi_lbm3d_amd.geometry import sphere_pack
    return [sphere_pack((size,) * 3, porosity=0.3, rmin=4.0, rmax=12.0, seed=seed + i, periodic=(True, True, True))
            for i in range(n)]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-samples", type=int, default=64)
    ap.add_argument("--train-seed", type=int, default=1000)
    ap.add_argument("--iters", type=int, default=32000)
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--eval-samples", type=int, default=32)
    ap.add_argument("--eval-size", type=int, default=192)
    ap.add_argument("--eval-seed", type=int, default=5000)
    ap.add_argument("--max-steps", type=int, default=50_000)
    ap.add_argument("--save-model", default="")
    ap.add_argument("--load-model", default="", help="skip training: load the network from this file")
    ap.add_argument("--rho-weight", type=float, default=0.1, help="weight of the density channel in the loss")
    ap.add_argument("--no-calibrate", dest="calibrate", action="store_false",
                    help="skip the velocity scale fitted on the training samples")
    ap.add_argument("--density", action="store_true",
                    help="start from the predicted density (at mean 1 per sample) instead of rho = 1")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)

    result = {"config": vars(args)}
    # sample geometries come from a worker pool: nothing may fork after the GPU is initialised
    t0 = time.time()
    solids, specs = make_samples(args.train_samples, args.train_seed)
    log(f"{len(solids)} training geometries ({time.time() - t0:.1f}s)")

    import torch
    assert torch.cuda.is_available(), "the warm-start pipeline runs on the GPU"
    dev = torch.device("cuda", torch.cuda.current_device())

    # 1. training data from the batched solver
    inputs, targets, k_train, nu, info = training_data(solids, args.max_steps)
    info["prep_wall_s"] = round(time.time() - t0, 1)
    result["data"] = info

    # 2. training
    net = sg.UNet3D(dev, seed=0)
    if args.load_model:
        net.load_state_dict(torch.load(args.load_model, map_location=dev, weights_only=True))
        result["train"] = {"loaded": args.load_model}
    else:
        sampler = CropSampler(inputs, targets, solids, dev, seed=1)
        curve, train_s = train(net, sampler, args.iters, args.batch, args.crop, rho_weight=args.rho_weight)
        result["train"] = {"iters": args.iters, "batch": args.batch, "crop": args.crop, "wall_s": round(train_s, 1),
                           "loss_first": curve[0][1], "loss_last": curve[-1][1], "curve": curve}
        del sampler
    del inputs, targets
    torch.cuda.empty_cache()
    # 3. velocity scale fitted on the training samples
    scale = velocity_scale(net, solids, k_train, nu) if args.calibrate else 1.0
    result["train"]["velocity_scale"] = scale
    if args.save_model:
        torch.save(net.state_dict(), args.save_model)

    # 4. evaluation on held-out packs
    ev = eval_packs(args.eval_samples, args.eval_size, args.eval_seed)
    log(f"evaluation: {len(ev)} packs of {args.eval_size}^3")
    result["eval"] = evaluate(net, ev, nu, scale, density=args.density, max_steps=args.max_steps)
    log(json.dumps({k: v for k, v in result["eval"].items() if k != "per_sample"}))
    line = json.dumps(result)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(line + "\n")
    print(line, flush=True)
    return result


if __name__ == "__main__":
    main()
