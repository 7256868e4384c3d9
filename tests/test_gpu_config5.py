"""The learned warm start at full size (GPU): the 3-D U-Net (4 levels,
16-32-64-128 channels, BF16 MFMA convolutions, Adam) trained in the test on
periodic 128^3 crops of the 64 batched packs' converged fields for a bounded
number of iterations, then the evaluation of the full workload: 32 held-out
seeded 192^3 packs (porosity 0.3) started cold and warm (f at the equilibrium
of the prediction, periodic inference on each whole volume), each batch run to
the 1e-6 steady criterion.  Both starts must reach the same permeability, with
no fallback to a cold start; the speedups are printed (the full training run
and its figures are warmstart.py's)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITERS = 3000


def test_warm_start_full_size(gpu):
    import torch
    from taichi_lbm3d_amd import surrogate as sg
    from taichi_lbm3d_amd.batch_bench import make_samples
    from taichi_lbm3d_amd.warmstart import CropSampler, eval_packs, evaluate, train, training_data, velocity_scale
    t0 = time.time()
    solids, _ = make_samples(64, workers=1)
    inputs, targets, k_train, nu, info = training_data(solids)
    assert info["converged"] == 64
    print(f"training data {time.time() - t0:.1f}s: {info}")
    net = sg.UNet3D(gpu, seed=0)
    sampler = CropSampler(inputs, targets, solids, gpu, seed=1)
    curve, train_s = train(net, sampler, ITERS, 2, 128, rho_weight=0.1, log_every=500)
    print(f"training: {ITERS} iterations of 2 x 128^3 in {train_s:.1f}s, loss {curve[0][1]:.3g} -> {curve[-1][1]:.3g}")
    assert curve[-1][1] < 0.5 * curve[0][1]
    del sampler, inputs, targets
    torch.cuda.empty_cache()
    scale = velocity_scale(net, solids, k_train, nu)
    assert 0.5 < scale < 2.0

    t0 = time.time()
    ev = eval_packs(32, 192)
    print(f"32 held-out packs generated in {time.time() - t0:.1f}s")
    r = evaluate(net, ev, nu, scale)
    per = r.pop("per_sample")
    print(r)
    assert r["samples"] == 32 and r["size"] == 192
    assert r["fallbacks"] == 0
    assert r["cold"]["converged"] == 32 and r["warm"]["converged"] == 32
    k_cold, k_warm = np.array(per["k_cold"]), np.array(per["k_warm"])
    assert np.all(np.abs(k_warm / k_cold - 1.0) < 1e-3), np.abs(k_warm / k_cold - 1.0).max()
    # the prediction starts far closer to the answer than the rest state (error 1)
    assert r["surrogate"]["k_rel_err_max"] < 0.2
    # the warm start gets within 1 % of the cold permeability in fewer steps,
    # for every sample, and reaches the steady criterion no later
    s1c, s1w = np.array(per["steps_to_1pct_cold"]), np.array(per["steps_to_1pct_warm"])
    assert np.all(s1w < s1c), (s1w, s1c)
    assert r["warm"]["steps_median"] <= r["cold"]["steps_median"]
    assert r["warm"]["steps_max"] <= r["cold"]["steps_max"]
    assert r["speedup_to_1pct"] > 1.5
