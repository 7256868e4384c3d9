"""CPU restatements of the geometry inputs (test infrastructure only).

* read_geo_loadtxt -- the reference's init_geo (SP:173-177): np.loadtxt, values
  > 0 set to 1, Fortran-order reshape to (nx, ny, nz), cast to the i8 field.
* winding_number -- the generalized winding number of a triangle mesh at
  points (FP64 numpy), the quantity whose sign the reference's STL converter
  (Convert_stl_to_binary.cpp:108-127, libigl fast winding number) thresholds:
  a node is solid when w >= 1/2.  Pinned in the tests by closed meshes with
  known interiors (w = 1 inside, 0 outside) and brute-force agreement.
* periodic_distance_edt -- Euclidean distance to the nearest solid cell under
  periodic wrap with scipy's exact EDT on a wrap-padded copy, capped at
  `margin`: the check of the surrogate's GPU distance input.
"""
from __future__ import annotations

import numpy as np


def read_geo_loadtxt(path: str, nx: int, ny: int, nz: int) -> np.ndarray:
    data = np.loadtxt(path)
    data[data > 0] = 1
    if data.size != nx * ny * nz:
        raise ValueError(f"{path}: {data.size} values, expected {nx}*{ny}*{nz}")
    return np.ascontiguousarray(np.reshape(data, (nx, ny, nz), order="F")).astype(np.int8)


def winding_number(tris: np.ndarray, points: np.ndarray) -> np.ndarray:
    """tris (T, 3, 3), points (P, 3) -> w (P,), Van Oosterom-Strackee solid angles / 4 pi."""
    tris = np.asarray(tris, np.float64)
    pts = np.asarray(points, np.float64)
    w = np.zeros(len(pts))
    for t in tris:
        a = t[0] - pts
        b = t[1] - pts
        c = t[2] - pts
        la, lb, lc = (np.linalg.norm(x, axis=1) for x in (a, b, c))
        det = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = la * lb * lc + np.einsum("ij,ij->i", a, b) * lc + np.einsum("ij,ij->i", b, c) * la \
            + np.einsum("ij,ij->i", c, a) * lb
        w += 2.0 * np.arctan2(det, den)
    return w / (4.0 * np.pi)


def periodic_distance_edt(solid: np.ndarray, margin: int = 24) -> np.ndarray:
    from scipy import ndimage
    s = np.asarray(solid) != 0
    if not s.any():
        return np.full(s.shape, float(margin), np.float32)
    m = min(margin, *s.shape)
    padded = np.pad(~s, m, mode="wrap")
    d = ndimage.distance_transform_edt(padded)[m:-m, m:-m, m:-m]
    return np.minimum(d, margin).astype(np.float32)
