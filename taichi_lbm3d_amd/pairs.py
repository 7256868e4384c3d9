"""Schedule of step pairs in one pass (lbm3d_run_pairs, include/lbm3d.h).

The fluid list is cut into tiles of 256 consecutive cells, the tiles into 8
slabs of consecutive tiles, one per XCD.  The even update of a tile touches
only its own cells' slots; the odd update of a tile touches slots of cells in
the planes of its cells and the planes next to them (the AA pattern's odd step
reads and writes each neighbour's slot, SP:259-269 streaming with the
unconditional periodic wrap of SP:247-257).  So the odd update of tile t may
run once the even updates of every tile holding cells of planes
first(t) - 1 .. last(t) + 1 are done.  When all those tiles lie in t's slab,
t is interior and its odd update joins the slab's item list right after the
even update of the last of them; the odd updates of the other tiles run in a
second launch.

No reference counterpart: a launch schedule of the same per-cell arithmetic.
"""
from __future__ import annotations

import numpy as np

TILE = 256
SLABS = 8


def pair_plan(cells: np.ndarray, nx: int, ny: int, nz: int, tile: int = TILE, slabs: int = SLABS,
              lag: int = 0) -> dict:
    """Item lists of lbm3d_pair_plan for the C-order fluid list `cells` (dense
    indices i*ny*nz + j*nz + k, increasing).

    Returns item_off (slabs + 1 offsets), items (tile id for an even update,
    ~tile for an odd update), list (tiles whose odd update runs in the second
    launch), need (per plane: tiles holding its cells) and tile_planes (first
    plane | last plane << 16), all int32 (tile_planes uint32).  lag > 0 moves
    each odd update behind `lag` more even updates of its slab (at most to the
    slab's end), so that the even updates it waits for have likely finished
    when a workgroup takes it."""
    cells = np.asarray(cells, np.int64)
    n = len(cells)
    if nx >= 1 << 16:
        raise ValueError("pair schedules index planes with 16 bits")
    n_tiles = (n + tile - 1) // tile
    plane = cells // (ny * nz)
    first = np.arange(n_tiles, dtype=np.int64) * tile
    last = np.minimum(first + tile - 1, n - 1)
    pf, pl = plane[first], plane[last]
    # tiles holding cells of each plane: a difference array over [pf, pl]
    d = np.zeros(nx + 1, np.int64)
    np.add.at(d, pf, 1)
    np.add.at(d, pl + 1, -1)
    need = np.cumsum(d[:nx]).astype(np.int32)
    # first and last tile of each plane (-1 where the plane has no fluid)
    start = np.searchsorted(plane, np.arange(nx), side="left")
    stop = np.searchsorted(plane, np.arange(nx), side="right")
    has = stop > start
    t_lo = np.where(has, start // tile, -1)
    t_hi = np.where(has, (stop - 1) // tile, -1)
    bounds = (np.arange(slabs + 1, dtype=np.int64) * n_tiles) // slabs
    slab_of = np.searchsorted(bounds[1:], np.arange(n_tiles), side="right")

    items = [[] for _ in range(slabs)]
    boundary = []
    after = {}  # tile -> interior tiles whose odd update follows its even update
    for t in range(n_tiles):
        s = slab_of[t]
        # the kernel's wait gives one lane to each plane: at most 64 of them
        lo_ok, dep_last = int(pl[t] - pf[t]) + 3 <= 64, -1
        for q in range(int(pf[t]) - 1, int(pl[t]) + 2) if lo_ok else ():
            q %= nx
            if not has[q]:
                continue
            if t_lo[q] < bounds[s] or t_hi[q] >= bounds[s + 1]:
                lo_ok = False
                break
            dep_last = max(dep_last, int(t_hi[q]))
        if lo_ok and nx >= 3:
            after.setdefault(min(dep_last + lag, int(bounds[s + 1]) - 1), []).append(t)
        else:
            boundary.append(t)
    for t in range(n_tiles):
        s = slab_of[t]
        items[s].append(t)
        for u in after.get(t, ()):
            items[s].append(~u)
    item_off = np.zeros(slabs + 1, np.int32)
    for s in range(slabs):
        item_off[s + 1] = item_off[s] + len(items[s])
    flat = np.array([v for run in items for v in run], np.int32) if n_tiles else np.zeros(0, np.int32)
    return {
        "item_off": item_off,
        "items": flat,
        "list": np.array(boundary, np.int32),
        "need": need,
        "tile_planes": (pf.astype(np.uint32) | (pl.astype(np.uint32) << 16)).astype(np.uint32),
        "slab_bounds": bounds,
    }
