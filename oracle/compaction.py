"""Brute-force fluid-list compaction and neighbour table (test infrastructure).

The reference has no fluid list: its kernels test ``solid == 0`` per cell
(SP:225, SP:262, SP:376) and stream with ``periodic_index`` (SP:247-257),
pushing ``F[ip][s] = f[i][s]`` when ``ip`` is fluid and bouncing back into
``F[i][LR[s]]`` otherwise (SP:264-268).  The HIP path stores the same
information as a numbered fluid list and a table; these functions restate it
two ways (plain loops, and numpy) so the GPU build can be checked bit for bit.

Numbering: fluid cells in list order.  With band_rows = R > 0 and R < ny the
grid rows (i, j) are cut into bands of R consecutive j (the last band may be
narrower) and the cells are numbered band after band, inside a band plane
after plane (i), row after row (j), k fastest; R = 0 or R >= ny is C order.
Face records: a fluid node on a face with
a pressure or velocity condition belongs to the last such face of x-left,
x-right, y-left, y-right, z-left, z-right, the order Boundary_condition
writes them (SP:274-370); its record is 1, plus 2 when that face's inward
neighbour has solid value > 0 (SP:278).
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


def list_order(shape, band_rows: int = 0) -> np.ndarray:
    """C-order dense indices of every cell of the grid in list order (numpy)."""
    nx, ny, nz = shape
    R = ny if band_rows <= 0 or band_rows >= ny else int(band_rows)
    idx = np.arange(nx * ny * nz, dtype=np.int64).reshape(nx, ny, nz)
    return np.concatenate([idx[:, b:b + R, :].reshape(-1) for b in range(0, ny, R)])


def compact(solid: np.ndarray, band_rows: int = 0):
    """(cells, dense_to_compact) with numpy: fluid cells (value 0) in list order."""
    solid = np.asarray(solid)
    fluid = (solid == 0).reshape(-1)
    order = list_order(solid.shape, band_rows)
    cells = order[fluid[order]].astype(np.int32)
    d2c = np.full(fluid.size, -1, np.int32)
    d2c[cells] = np.arange(cells.size, dtype=np.int32)
    return cells, d2c


def face_records(solid: np.ndarray, cells: np.ndarray, bc_mode) -> np.ndarray:
    """Per fluid cell: 0, or 1 for a face-owned node, +2 if the owner's inward neighbour is solid."""
    owner, tag = face_owner(solid, bc_mode)
    own = owner.reshape(-1)[cells]
    inward = (tag.reshape(-1)[cells] >> 5) & 1
    return np.where(own >= 0, 1 + 2 * inward, 0).astype(np.uint32)


def pack_face_bits(rec: np.ndarray) -> np.ndarray:
    """2-bit records packed 16 per uint32 word, cell x at bits 2*(x%16)."""
    n = rec.size
    words = np.zeros((n + 15) // 16, np.uint32)
    for b in range(16):
        part = rec[b::16].astype(np.uint32)
        words[: part.size] |= part << np.uint32(2 * b)
    return words


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


def compact_loops(solid, bc_mode=(0,) * 6, band_rows: int = 0):
    """The same numbering, face records and table with plain Python loops
    (small grids only), following periodic_index (SP:247-257), streaming1
    (SP:259-269) and the face order of Boundary_condition (SP:274-370) literally."""
    solid = np.asarray(solid)
    nx, ny, nz = solid.shape
    R = ny if band_rows <= 0 or band_rows >= ny else int(band_rows)
    order = []
    for j0 in range(0, ny, R):
        for i in range(nx):
            for j in range(j0, min(j0 + R, ny)):
                for k in range(nz):
                    if solid[i, j, k] == 0:
                        order.append((i, j, k))
    ids = {c: x for x, c in enumerate(order)}
    rec = []
    for (i, j, k) in order:
        owner = -1
        for f in range(6):
            if int(bc_mode[f]) not in (1, 2):
                continue
            axis, right = divmod(f, 2)
            if (i, j, k)[axis] == ((nx, ny, nz)[axis] - 1 if right else 0):
                owner = f
        if owner < 0:
            rec.append(0)
            continue
        axis, right = divmod(owner, 2)
        q = [i, j, k]
        q[axis] = (nx, ny, nz)[axis] - 2 if right else 1
        rec.append(1 + 2 * (1 if solid[q[0], q[1], q[2]] > 0 else 0))
    nbr = np.empty((18, len(order)), np.int32)
    for x, (i, j, k) in enumerate(order):
        for q in range(1, 19):
            ip = [i + int(E[q, 0]), j + int(E[q, 1]), k + int(E[q, 2])]
            if ip[0] < 0: ip[0] = nx - 1
            if ip[0] > nx - 1: ip[0] = 0
            if ip[1] < 0: ip[1] = ny - 1
            if ip[1] > ny - 1: ip[1] = 0
            if ip[2] < 0: ip[2] = nz - 1
            if ip[2] > nz - 1: ip[2] = 0
            nbr[q - 1, x] = ids[tuple(ip)] if solid[ip[0], ip[1], ip[2]] == 0 else -1
    cells = np.array([(i * ny + j) * nz + k for (i, j, k) in order], np.int32)
    return cells, np.array(rec, np.uint32), nbr
