"""The in-place AA access pattern of the step kernel against streaming1 (CPU).

The reference streams by pushing each post-collision population into a second
array: F[x + e_q][q] = f[x][q] when x + e_q is fluid, else F[x][LR[q]] =
f[x][q] (SP:259-269).  Every slot of F is written exactly once per step, which
is what makes the two-array push race-free.  The HIP kernel keeps one array
and alternates two address sets per cell x and direction q:

  even steps: read slot q of x, write f_q to slot LR[q] of x;
  odd steps:  A_q(x) = slot q of x + e_q when that cell is fluid, else slot
              LR[q] of x (the kernel's odd_offsets); read population LR[q]
              from A_q(x) and write f_q back to A_q(x).

These tests check, from neighbour tables built independently of the kernel
(oracle.compaction, a brute force over periodic_index, SP:247-257), that each
address set hits every slot exactly once (the write-once property of
SP:259-269 that makes the in-place update race-free), and that an even step
followed by an odd step moves populations exactly as two streaming1 pushes do.
A planted mistake in the address rule must fail the checker.
"""
import os

import numpy as np
import pytest

from oracle import compaction as oc
from oracle.oracle_np import LR, NumpySolver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def odd_addresses(nbr, lr=LR):
    """(slot, cell) of A_q(x) for q = 0..18 and every fluid cell x, shape (19, n)."""
    n = nbr.shape[1]
    own = np.arange(n, dtype=np.int64)
    slot = np.empty((19, n), np.int64)
    cell = np.empty((19, n), np.int64)
    slot[0], cell[0] = 0, own
    for q in range(1, 19):
        nb = nbr[q - 1].astype(np.int64)
        fluid = nb >= 0
        slot[q] = np.where(fluid, q, lr[q])
        cell[q] = np.where(fluid, nb, own)
    return slot, cell


def even_addresses(n):
    """(slot, cell) written by the even step for q = 0..18: slot LR[q] of x."""
    own = np.broadcast_to(np.arange(n, dtype=np.int64), (19, n))
    return np.broadcast_to(LR[:, None].astype(np.int64), (19, n)), own


def hits(slot, cell, n):
    """Number of times each of the 19 n slots is addressed."""
    return np.bincount((slot * n + cell).reshape(-1), minlength=19 * n)


def geometries():
    rng = np.random.default_rng(3)
    out = {
        "random_11x9x7": (rng.random((11, 9, 7)) < 0.3).astype(np.int8),
        "all_fluid_6x5x4": np.zeros((6, 5, 4), np.int8),
        "thin_1x6x5": (rng.random((1, 6, 5)) < 0.25).astype(np.int8),
        "thin_2x3x9": (rng.random((2, 3, 9)) < 0.25).astype(np.int8),
        "isolated_cells": np.ones((5, 5, 5), np.int8),
    }
    out["isolated_cells"][1, 1, 1] = out["isolated_cells"][3, 2, 4] = 0
    g = np.load(os.path.join(GOLDEN, "pack32_compaction.npz"))
    out["pack32_fixture"] = np.unpackbits(g["bits"])[: 32 ** 3].reshape(32, 32, 32).astype(np.int8)
    d = np.load(os.path.join(GOLDEN, "geo_cavity.npz"))
    out["cavity50_fixture"] = np.unpackbits(d["bits"])[: int(d["n"])].reshape((50, 50, 50), order="F").astype(np.int8)
    return out


GEOMS = geometries()


@pytest.mark.parametrize("name", sorted(GEOMS))
def test_each_step_addresses_every_slot_once(name):
    solid = GEOMS[name]
    cells, d2c = oc.compact(solid)
    nbr = oc.neighbors(solid, cells, d2c)
    n = cells.size
    assert n > 0
    odd = hits(*odd_addresses(nbr), n)
    even = hits(*even_addresses(n), n)
    assert odd.min() == 1 and odd.max() == 1, name
    assert even.min() == 1 and even.max() == 1, name


def test_checker_catches_a_wrong_bounce_back_slot():
    """Bouncing back into slot q of x instead of LR[q] leaves slots unwritten
    and others written twice on any grid with a solid neighbour."""
    solid = GEOMS["random_11x9x7"]
    cells, d2c = oc.compact(solid)
    nbr = oc.neighbors(solid, cells, d2c)
    wrong = np.arange(19)
    h = hits(*odd_addresses(nbr, lr=wrong), cells.size)
    assert h.min() == 0 and h.max() == 2


def dense_f(solid, cells, post):
    """(nx, ny, nz, 19) array holding post[q, x] at the fluid cells."""
    f = np.zeros(solid.shape + (19,))
    f.reshape(-1, 19)[cells] = post.T
    return f


def push(solid, cells, post):
    """streaming1 (SP:259-269) of post-collision populations post[q, x], read
    back on the fluid list: F[q, x]."""
    ns = NumpySolver(solid)
    ns.f = dense_f(solid, cells, post)
    ns.F = np.full(solid.shape + (19,), np.nan)
    ns.streaming1()
    return ns.F.reshape(-1, 19)[cells].T


@pytest.mark.parametrize("name", ["random_11x9x7", "thin_1x6x5", "thin_2x3x9", "pack32_fixture", "isolated_cells"])
def test_even_then_odd_step_equals_two_pushes(name):
    """Post-collision populations written by an even step and read by the
    next odd step arrive as streaming1 delivers them; those the odd step
    writes and the following even step reads arrive likewise."""
    solid = GEOMS[name]
    cells, d2c = oc.compact(solid)
    nbr = oc.neighbors(solid, cells, d2c)
    n = cells.size
    rng = np.random.default_rng(5)
    slot, cell = odd_addresses(nbr)
    store = np.full((19, n), np.nan)

    # even step at t: f_q(x) -> slot LR[q] of x
    post1 = rng.random((19, n))
    store[LR[:, None], np.arange(n)[None, :]] = post1
    # odd step at t+1 reads population LR[q] at A_q(x)
    got1 = np.empty((19, n))
    got1[LR] = store[slot, cell]
    np.testing.assert_array_equal(got1, push(solid, cells, post1))

    # odd step at t+1 writes f_q(x) to A_q(x); the even step at t+2 reads slot q of x
    post2 = rng.random((19, n))
    store[:] = np.nan
    store[slot, cell] = post2
    assert not np.isnan(store).any()
    np.testing.assert_array_equal(store, push(solid, cells, post2))
