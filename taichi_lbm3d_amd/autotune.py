"""Launch-shape autotuner for the batched step kernel.

The batched step (lbm3d_batch_step) takes, per AA parity, a launch shape
(block threads, persistent workgroups per compute unit, sweep direction).
Results do not depend on the shape (the kernel is bit-identical for every
shape), only the time does, so the tuner times real steps of the caller's own
batch: every candidate runs a few (even, odd) step pairs, each step timed with
HIP events on the solver's stream, and the fastest shape per parity is kept.
A step's speed also depends on the step before it: an even step that sweeps
against the preceding odd step's direction starts on the slots that step wrote
last, still in the last-level cache.  So the per-parity winners, their two
alternating-direction variants and the default pair are then timed as pairs,
and the fastest pair wins.
The winners are cached per device, storage type, force switch and batch size
(bucketed by powers of two) in ~/.cache/taichi_lbm3d_amd/launch.json;
TAICHI_LBM3D_TUNE_CACHE=0 disables the cache.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

BLOCKS = (128, 256, 512)
GRID_PER_CU = (0, 2, 4, 8)
CANDIDATES = tuple((b, g, r) for b in BLOCKS for g in GRID_PER_CU for r in (0, 1))
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "taichi_lbm3d_amd", "launch.json")


def _key(batch) -> str:
    import torch
    name = torch.cuda.get_device_name(batch.device)
    bucket = int(math.log2(max(int(batch.sizes.cells), 1)))
    return f"{name}|{batch.dtype}|force{int(batch.p.force_on)}|cells2^{bucket}"


def _load_cache() -> dict:
    if os.environ.get("TAICHI_LBM3D_TUNE_CACHE", "1") == "0":
        return {}
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, entry: dict) -> None:
    if os.environ.get("TAICHI_LBM3D_TUNE_CACHE", "1") == "0":
        return
    data = _load_cache()
    data[key] = entry
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        tmp = CACHE + f".{os.getpid()}"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def time_shape(batch, shape, pairs: int = 3, odd_shape=None):
    """Median (even_us, odd_us) of `pairs` step pairs with launch shape `shape`
    for both parities (or `odd_shape` for odd steps); the batch advances by
    2 * (pairs + 1) steps."""
    import torch
    batch.set_launch(shape, odd_shape or shape)
    if batch.steps_done % 2:
        batch.run(1)
    batch.run(2)  # first launches of this shape
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * pairs + 1)]
    stream = torch.cuda.current_stream(batch.device)
    ev[0].record(stream)
    for i in range(2 * pairs):
        batch.run(1)
        ev[i + 1].record(stream)
    torch.cuda.synchronize(batch.device)
    dt = np.array([ev[i].elapsed_time(ev[i + 1]) * 1e3 for i in range(2 * pairs)])
    return float(np.median(dt[0::2])), float(np.median(dt[1::2]))


def tune(batch, candidates=CANDIDATES, pairs: int = 3, use_cache: bool = True):
    """Pick and set the fastest (even, odd) launch shapes for `batch`.

    Returns {"even": shape, "odd": shape, "table": [(shape, even_us, odd_us)],
    "cached": bool}.  The batch's state advances by real steps while tuning.
    """
    key = _key(batch)
    # shapes are timed step by step over the whole batch, outside the
    # cache-blocked group schedule (restored before returning)
    groups, live = batch.groups, batch._live
    batch.groups, batch._live = [(0, int(batch.sizes.cells))], [0]
    try:
        return _tune(batch, key, candidates, pairs, use_cache)
    finally:
        batch.groups, batch._live = groups, live
        batch._graph = None


def _tune(batch, key, candidates, pairs, use_cache):
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            batch.set_launch(tuple(hit["even"]), tuple(hit["odd"]))
            return {"even": tuple(hit["even"]), "odd": tuple(hit["odd"]), "table": hit.get("table", []),
                    "cached": True}
    table = []
    for shape in candidates:
        e, o = time_shape(batch, shape, pairs)
        table.append((tuple(shape), e, o))
    best_even = min(table, key=lambda r: r[1])[0]
    best_odd = min(table, key=lambda r: r[2])[0]
    pairs_tried = []
    for ev, od in {(best_even, best_odd),
                   (best_even[:2] + (1,), best_odd[:2] + (0,)),
                   (best_even[:2] + (0,), best_odd[:2] + (1,)),
                   ((256, 0, 1), (256, 0, 0))}:
        e, o = time_shape(batch, ev, pairs, odd_shape=od)
        pairs_tried.append((ev, od, e + o))
    best_even, best_odd, _ = min(pairs_tried, key=lambda r: r[2])
    batch.set_launch(best_even, best_odd)
    entry = {"even": list(best_even), "odd": list(best_odd),
             "table": [[list(s), round(e, 2), round(o, 2)] for s, e, o in table],
             "pairs": [[list(e), list(o), round(t, 2)] for e, o, t in pairs_tried]}
    if use_cache:
        _store_cache(key, entry)
    return {"even": best_even, "odd": best_odd, "table": table, "pairs": pairs_tried, "cached": False}
