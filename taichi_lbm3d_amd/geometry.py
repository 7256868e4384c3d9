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


def _ball_offsets(r: float):
    ri = int(np.ceil(r))
    ax = np.arange(-ri, ri + 1)
    return ri, ax


def sphere_pack(shape=(256, 256, 256), porosity: float = 0.35, rmin: float = 6.0, rmax: float = 14.0,
                seed: int = 0, periodic=(False, True, True), max_spheres: int = 10_000_000) -> np.ndarray:
    """Seeded random overlapping-sphere pack.

    Centres uniform in the box, radii uniform in [rmin, rmax]; a cell (i,j,k)
    is solid when its integer centre lies within r of a sphere centre, with
    minimum-image wrap along the axes marked periodic and clipping along the
    others.  Spheres are added one at a time until the fluid fraction is at or
    below ``porosity``.
    """
    nx, ny, nz = shape
    n = np.array(shape)
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, np.uint8)
    total = nx * ny * nz
    n_solid = 0
    target = (1.0 - porosity) * total
    for _ in range(max_spheres):
        if n_solid >= target:
            break
        c = rng.uniform(0.0, 1.0, 3) * n
        r = rng.uniform(rmin, rmax)
        ri = int(np.ceil(r))
        idx = []
        dist2 = 0.0
        for a in range(3):
            lo = int(np.floor(c[a] - r))
            hi = int(np.ceil(c[a] + r))
            cells = np.arange(lo, hi + 1)
            d = cells - c[a]
            if periodic[a]:
                cells = np.mod(cells, n[a])
            else:
                keep = (cells >= 0) & (cells < n[a])
                cells, d = cells[keep], d[keep]
            shp = [1, 1, 1]
            shp[a] = -1
            idx.append(cells.reshape(shp))
            dist2 = dist2 + (d * d).reshape(shp)
        inside = dist2 <= r * r
        del ri
        sub = solid[idx[0], idx[1], idx[2]]
        # cells may repeat along a periodic axis only when 2r+1 > n; count once
        newly = inside & (sub == 0)
        if np.any(newly):
            ii, jj, kk = np.broadcast_arrays(idx[0], idx[1], idx[2])
            sel = np.unique(np.stack([ii[newly], jj[newly], kk[newly]], 0), axis=1)
            solid[sel[0], sel[1], sel[2]] = 1
            n_solid += sel.shape[1]
    return solid


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
