"""Voxel geometries: the reference's ASCII format and seeded synthetic media.

All arrays are uint8 C-order ``(nx, ny, nz)``, 0 = fluid, 1 = solid, like the
``solid`` field of the reference solver (SP:57).
"""
from __future__ import annotations

import numpy as np


def read_geo(path: str, nx: int, ny: int, nz: int) -> np.ndarray:
    """ASCII voxel file -> solid mask, as init_geo does (SP:173-177).

    One number per line, x fastest, then y, then z (README.md:17-23); any value
    > 0 becomes 1 and the field is int8 like the reference's i8 ``solid``
    (np.loadtxt, a Fortran-order reshape to (nx, ny, nz) and a cast), so a
    negative value stays negative: such a cell is not fluid (the kernels test
    ``solid == 0``) yet not solid for the pressure-face test ``solid > 0``.
    Parsed by the library's native reader (lbm3d_read_geo_ascii).
    """
    import ctypes
    from . import _lib
    out = np.empty((nx, ny, nz), np.int8)
    lib = _lib.lib()
    n = lib.cdll.lbm3d_read_geo_ascii(str(path).encode(), nx, ny, nz, out.ctypes.data_as(ctypes.c_void_p))
    if n < 0:
        raise ValueError(lib.cdll.lbm3d_io_last_error().decode())
    return out


def write_geo(path: str, solid: np.ndarray) -> None:
    """Inverse of :func:`read_geo` (the layout flow_domain_geo_generation_2D.py:27-30 writes)."""
    flat = np.asarray(solid).reshape(-1, order="F")
    np.savetxt(path, flat.astype(np.int64), fmt="%d")


def sphere_pack(shape=(256, 256, 256), porosity: float = 0.35, rmin: float = 6.0, rmax: float = 14.0,
                seed: int = 0, periodic=(False, True, True), max_spheres: int = 10_000_000) -> np.ndarray:
    """Seeded random overlapping-sphere pack.

    Centres uniform in the box, radii uniform in [rmin, rmax]; a cell (i,j,k)
    is solid when its integer centre lies within r of a sphere centre, with
    minimum-image wrap along the axes marked periodic and clipping along the
    others.  Spheres are added one at a time until the fluid fraction is at or
    below ``porosity``.

    Sphere s draws four doubles from ``default_rng(seed)``: the centre
    (u0, u1, u2) * shape and the radius rmin + (rmax - rmin) u3.  Cell (i,j,k)
    lies in the sphere when ((0 + dx^2) + dy^2) + dz^2 <= r*r, d the offset of
    the cell (before wrapping) from the centre.  The spheres are evaluated in
    chunks: within a chunk, each newly covered cell is credited to the first
    sphere covering it, and the pack keeps the spheres up to the first one at
    which the solid count reaches the target -- the same cells as adding them
    one at a time.
    """
    nx, ny, nz = (int(a) for a in shape)
    n = np.array([nx, ny, nz], np.float64)
    rng = np.random.default_rng(seed)
    flat = np.zeros(nx * ny * nz, np.uint8)
    target = (1.0 - porosity) * flat.size
    n_solid = 0
    box = 2 * int(np.ceil(rmax)) + 3  # cells per axis of a sphere's bounding range
    # spheres per chunk: about a quarter of the count a pack without overlap would
    # need, at most ~4M candidate cells per chunk
    ball = 4.0 / 3.0 * np.pi * (0.5 * (rmin + rmax)) ** 3
    chunk = int(max(8, min(target / ball / 4 + 1, 4_000_000 // box ** 3)))
    off = np.arange(box, dtype=np.int64)
    used = 0
    while n_solid < target and used < max_spheres:
        k = min(chunk, max_spheres - used)
        u = rng.random((k, 4))
        cen = u[:, :3] * n
        r = rmin + (rmax - rmin) * u[:, 3]
        axes = []
        for a, ext in enumerate((nx, ny, nz)):
            lo = np.floor(cen[:, a] - r).astype(np.int64)
            hi = np.ceil(cen[:, a] + r).astype(np.int64)
            cells = lo[:, None] + off[None, :]
            d = cells - cen[:, a:a + 1]
            ok = cells <= hi[:, None]
            if periodic[a]:
                cells = np.mod(cells, ext)
            else:
                ok &= (cells >= 0) & (cells < ext)
            axes.append((cells, d * d, ok))
        (cx, dx2, okx), (cy, dy2, oky), (cz, dz2, okz) = axes
        dist2 = ((0.0 + dx2)[:, :, None, None] + dy2[:, None, :, None]) + dz2[:, None, None, :]
        inside = (dist2 <= (r * r)[:, None, None, None]) & okx[:, :, None, None] & oky[:, None, :, None] \
            & okz[:, None, None, :]
        sid, a_, b_, c_ = np.nonzero(inside)
        idx = (cx[sid, a_] * ny + cy[sid, b_]) * nz + cz[sid, c_]
        fresh = flat[idx] == 0
        sid, idx = sid[fresh], idx[fresh]
        # first sphere of the chunk covering each cell (nonzero order is by sphere)
        uidx, first = np.unique(idx, return_index=True)
        owner = sid[first]
        gained = np.cumsum(np.bincount(owner, minlength=k))
        reach = np.nonzero(n_solid + gained >= target)[0]
        last = int(reach[0]) if reach.size else k - 1
        flat[uidx[owner <= last]] = 1
        n_solid += int(gained[last])
        used += last + 1
    return flat.reshape(nx, ny, nz)


def cube_channel(shape=(512, 512, 512), fraction: float = 0.05, emin: int = 4, emax: int = 16,
                 seed: int = 0) -> np.ndarray:
    """Channel periodic in x with solid walls at y = 0, ny-1 and z = 0, nz-1,
    filled with seeded random axis-aligned solid cubes (integer edge uniform
    in [emin, emax], corner uniform in the box, clipped) until the cubes cover
    ``fraction`` of the volume."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, np.uint8)
    covered = 0
    target = fraction * nx * ny * nz
    while covered < target:
        e = int(rng.integers(emin, emax + 1))
        x0, y0, z0 = (int(rng.integers(0, n)) for n in shape)
        block = solid[x0:x0 + e, y0:y0 + e, z0:z0 + e]
        covered += int((block == 0).sum())
        block[...] = 1
    solid[:, 0, :] = 1
    solid[:, -1, :] = 1
    solid[:, :, 0] = 1
    solid[:, :, -1] = 1
    return solid


def cavity(n: int = 50) -> np.ndarray:
    """The lid-driven cavity mask of flow_domain_geo_generation_2D.py:15-25:
    solid at x = 0 and on both y and z faces, open at x = n-1."""
    g = np.zeros((n, n, n), np.uint8)
    g[0, :, :] = 1
    g[:, 0, :] = 1
    g[:, -1, :] = 1
    g[:, :, 0] = 1
    g[:, :, -1] = 1
    return g


def porosity(solid: np.ndarray) -> float:
    return float((np.asarray(solid) == 0).mean())
