"""Brute-force fluid-list compaction and neighbour table (test infrastructure).

The reference has no fluid list: its kernels test ``solid == 0`` per cell
(SP:225, SP:262, SP:376) and stream with ``periodic_index`` (SP:247-257),
pushing ``F[ip][s] = f[i][s]`` when ``ip`` is fluid and bouncing back into
``F[i][LR[s]]`` otherwise (SP:264-268).  The HIP path stores the same
information as a numbered fluid list and a table; these functions restate it
two ways (plain loops, and numpy) so the GPU build can be checked bit for bit.

Numbering: interior fluid cells first, then cells owned by a face with a
pressure or velocity condition, each group in C order.  A node on several such
faces belongs to the last of x-left, x-right, y-left, y-right, z-left, z-right,
the order Boundary_condition writes them (SP:274-370).
"""
from __future__ import annotations

import numpy as np

from .oracle_np import E


def face_owner(solid: np.ndarray, bc_mode) -> tuple[np.ndarray, np.ndarray]:
    """owner face per cell (-1 none) and the face tag (owner | mode<<3 | inward_solid<<5)."""
    raw = np.asarray(solid)
    positive = raw > 0          # "solid" for the inward test, SP:278
    solid = raw != 0            # not fluid, SP:225
    nx, ny, nz = solid.shape
    n = (nx, ny, nz)
    owner = np.full(solid.shape, -1, np.int32)
    tag = np.zeros(solid.shape, np.int32)
    for f in range(6):
        m = int(bc_mode[f])
        if m not in (1, 2):
            continue
        axis, right = divmod(f, 2)
        pos = n[axis] - 1 if right else 0
        inw = n[axis] - 2 if right else 1
        sl = [slice(None)] * 3
        sl[axis] = pos
        sq = [slice(None)] * 3
        sq[axis] = inw
        owner[tuple(sl)] = f
        tag[tuple(sl)] = f | (m << 3) | (positive[tuple(sq)].astype(np.int32) << 5)
    owner[solid] = -1
    tag[solid] = 0
    return owner, tag


def compact(solid: np.ndarray, bc_mode=(0,) * 6):
    """(cells, n_interior, face_tag, dense_to_compact) with numpy."""
    owner, tag = face_owner(solid, bc_mode)
    solid = np.asarray(solid) != 0
    fluid = ~solid
    interior = np.flatnonzero((fluid & (owner < 0)).reshape(-1))
    face = np.flatnonzero((fluid & (owner >= 0)).reshape(-1))
    cells = np.concatenate([interior, face]).astype(np.int32)
    d2c = np.full(solid.size, -1, np.int32)
    d2c[cells] = np.arange(cells.size, dtype=np.int32)
    return cells, int(interior.size), tag.reshape(-1)[face].astype(np.int32), d2c


def neighbors(solid: np.ndarray, cells: np.ndarray, d2c: np.ndarray) -> np.ndarray:
    """nbr[q-1, x] for q = 1..18: compact id of wrap((i,j,k)+e_q), -1 if solid (numpy)."""
    nx, ny, nz = np.asarray(solid).shape
    c = cells.astype(np.int64)
    i, j, k = c // (ny * nz), (c // nz) % ny, c % nz
    out = np.empty((18, cells.size), np.int32)
    for q in range(1, 19):
        a = np.mod(i + E[q, 0], nx)
        b = np.mod(j + E[q, 1], ny)
        d = np.mod(k + E[q, 2], nz)
        out[q - 1] = d2c[(a * ny + b) * nz + d]
    return out


def compact_loops(solid, bc_mode=(0,) * 6):
    """The same numbering and table with plain Python loops (small grids only),
    following periodic_index (SP:247-257) and streaming1 (SP:259-269) literally."""
    solid = np.asarray(solid)
    nx, ny, nz = solid.shape
    order = []
    for face_pass in (False, True):
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    if solid[i, j, k] != 0:
                        continue
                    owner = -1
                    for f in range(6):
                        if int(bc_mode[f]) not in (1, 2):
                            continue
                        axis, right = divmod(f, 2)
                        pos = (i, j, k)[axis]
                        if pos == ((nx, ny, nz)[axis] - 1 if right else 0):
                            owner = f
                    if (owner >= 0) == face_pass:
                        order.append((i, j, k, owner))
    ids = {(i, j, k): x for x, (i, j, k, _) in enumerate(order)}
    n_int = sum(1 for o in order if o[3] < 0)
    tags = []
    for (i, j, k, owner) in order[n_int:]:
        axis, right = divmod(owner, 2)
        n = (nx, ny, nz)[axis]
        q = [i, j, k]
        q[axis] = n - 2 if right else 1
        tags.append(owner | (int(bc_mode[owner]) << 3) | ((1 if solid[q[0], q[1], q[2]] > 0 else 0) << 5))
    nbr = np.empty((18, len(order)), np.int32)
    for x, (i, j, k, _) in enumerate(order):
        for q in range(1, 19):
            ip = [i + int(E[q, 0]), j + int(E[q, 1]), k + int(E[q, 2])]
            if ip[0] < 0: ip[0] = nx - 1
            if ip[0] > nx - 1: ip[0] = 0
            if ip[1] < 0: ip[1] = ny - 1
            if ip[1] > ny - 1: ip[1] = 0
            if ip[2] < 0: ip[2] = nz - 1
            if ip[2] > nz - 1: ip[2] = 0
            nbr[q - 1, x] = ids.get(tuple(ip), -1) if solid[ip[0], ip[1], ip[2]] == 0 else -1
    cells = np.array([(i * ny + j) * nz + k for (i, j, k, _) in order], np.int32)
    return cells, n_int, np.array(tags, np.int32), nbr
