"""Storage-precision study at full size (GPU): the 256^3 sphere pack run to
steady state and the 512^3 cube channel at a fixed step count, every storage
format against an FP64 run of the same kernels (FP64 storage and arithmetic,
itself matched to the FP64 CPU oracle in test_gpu_features).

The bounds are the figures the precision study reports, with margin; BF16 and
FP16 with round-to-nearest are asserted to fail where they do (the weak-force
channel), so a regression that fixed or broke them would show."""
import time

import pytest

pytestmark = pytest.mark.gpu


def study(domain, dtypes, steps):
    import torch
    from taichi_lbm3d_amd import precision_study as ps
    t0 = time.time()
    solid, setup, sparse = ps.DOMAINS[domain]()
    print(f"{domain}: geometry {time.time() - t0:.1f}s")
    steady = domain == "pack256"
    out = {}
    ref, rho64, v64 = ps.run_format(solid, setup, sparse, "f64", steady, steps, 200)
    n = ref["fluid_cells"]
    vnorm = torch.linalg.vector_norm(v64).item()
    out["f64"] = ref
    print(ref)
    for dt in dtypes:
        res, rho, v = ps.run_format(solid, setup, sparse, dt, steady, steps, 200)
        res["v_rel_l2"] = torch.linalg.vector_norm(v[: 3 * n] - v64[: 3 * n]).item() / vnorm
        res["rho_max_abs_err"] = (rho[:n] - rho64[:n]).abs().max().item()
        res["k_rel_err"] = res["permeability"] / ref["permeability"] - 1.0
        print(res)
        out[dt] = res
        del rho, v
    del rho64, v64
    torch.cuda.empty_cache()
    print(f"{domain}: {time.time() - t0:.1f}s in all")
    return out


def test_pack256_storage_formats_to_steady_state(gpu):
    """Each format runs to its own 1e-6 steady state (at most 50k steps)."""
    r = study("pack256", ["f32", "f16", "bf16", "f16sr", "bf16sr"], 50_000)
    assert r["f64"]["converged"] and r["f64"]["nonfinite"] == 0
    k64 = r["f64"]["permeability"]
    assert 0.30 < k64 < 0.38
    f32 = r["f32"]
    assert f32["converged"] and f32["v_rel_l2"] < 1e-4 and abs(f32["k_rel_err"]) < 2e-4
    assert abs(f32["steps"] - r["f64"]["steps"]) <= 1000
    # 16-bit storage within 0.5 % of the FP64 permeability: FP16 (round to
    # nearest, slow to settle) and both stochastic-rounding formats
    for dt in ("f16", "f16sr", "bf16sr"):
        assert r[dt]["converged"] and r[dt]["nonfinite"] == 0, dt
        assert abs(r[dt]["k_rel_err"]) < 5e-3, dt
        assert r[dt]["v_rel_l2"] < 2e-2, dt
    # FP16 with round to nearest needs about 3x the FP32 steps to settle (its
    # rounding error slows the last digits of <vx>); with stochastic rounding
    # the rounding noise of <vx> is of the order of the 1e-6 criterion, so the
    # step that first meets it is random (0.8x-3.1x the FP32 count over the
    # code versions seen); both converge well inside the 50k-step limit
    for dt in ("f16sr", "bf16sr"):
        assert r[dt]["steps"] <= 40_000, dt
    assert r["f16"]["steps"] > 2.5 * f32["steps"]
    # BF16 with round to nearest misses the target by several percent
    assert r["bf16"]["converged"] and 5e-3 < abs(r["bf16"]["k_rel_err"]) < 0.1
    # 16-bit storage moves half the bytes
    for dt in ("f16", "bf16", "f16sr", "bf16sr"):
        assert r[dt]["mlups"] > 1.3 * f32["mlups"], dt


def test_channel512_weak_force_storage_formats(gpu):
    """fx = 1e-6, 4,000 steps from rest: FP16 with round to nearest never
    starts the flow (the force increment is below FP16's spacing at 0), BF16
    stalls below the FP64 velocity, stochastic rounding tracks FP64."""
    r = study("channel512", ["f32", "f16", "bf16", "f16sr", "bf16sr"], 4000)
    assert r["f64"]["nonfinite"] == 0 and r["f64"]["permeability"] > 0
    assert r["f32"]["v_rel_l2"] < 1e-4 and abs(r["f32"]["k_rel_err"]) < 1e-4
    assert r["f16"]["v_rel_l2"] > 0.9 and r["f16"]["k_rel_err"] < -0.9
    assert r["bf16"]["v_rel_l2"] > 0.1 and r["bf16"]["k_rel_err"] < -0.1
    for dt in ("f16sr", "bf16sr"):
        assert r[dt]["v_rel_l2"] < 5e-3, dt
        assert abs(r[dt]["k_rel_err"]) < 5e-4, dt
