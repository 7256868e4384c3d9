"""STL input: read a triangle mesh and voxelize it on the GPU.

The reference converts an STL surface to the solver's 0/1 voxel grid with a
libigl program (Single_phase/Convert_stl_to_binary.cpp:29-157): grid nodes at
(x0 + i*dx, y0 + j*dx, z0 + k*dx), nx = ceil((x1 - x0)/dx) etc., a node solid
(1) unless its winding-number signed distance is positive (outside), the flags
written as a legacy ASCII VTK file.  Here the mesh is read in Python (binary or
ASCII STL) and the inside test runs as the HIP kernel lbm3d_voxelize (exact
generalized winding number per node, solid when w >= 1/2).
"""
from __future__ import annotations

import ctypes
import math
import struct

import numpy as np

from . import _lib


def read_stl(path: str) -> np.ndarray:
    """Triangles (T, 3, 3) float32 of a binary or ASCII STL file."""
    with open(path, "rb") as fh:
        data = fh.read()
    if len(data) >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            return np.ascontiguousarray(rec["v"], dtype=np.float32)
    text = data.decode("ascii", errors="replace")
    verts = [line.split()[1:4] for line in text.splitlines() if line.strip().startswith("vertex")]
    if len(verts) % 3:
        raise ValueError(f"{path}: vertex count {len(verts)} is not a multiple of 3")
    return np.asarray(verts, np.float32).reshape(-1, 3, 3)


def write_stl(path: str, tris: np.ndarray) -> None:
    """Binary STL (normals from the corner order)."""
    t = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    nrm = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    ln = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(ln > 0, nrm / np.maximum(ln, 1e-30), 0).astype(np.float32)
    rec = np.zeros(len(t), dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    rec["n"], rec["v"] = nrm, t
    with open(path, "wb") as fh:
        fh.write(b"taichi_lbm3d_amd binary STL".ljust(80, b" "))
        fh.write(struct.pack("<I", len(t)))
        fh.write(rec.tobytes())


def grid_extent(lower, upper, dx: float):
    """Node counts of the reference's grid: ceil((upper - lower) / dx) per axis."""
    return tuple(int(math.ceil((u - l) / dx)) for l, u in zip(lower, upper))


def voxelize(tris, lower, upper, dx: float, device=None) -> np.ndarray:
    """int8 (nx, ny, nz): 1 where the node is inside the closed mesh (or on it)."""
    import torch
    tris = np.ascontiguousarray(np.asarray(tris, np.float32).reshape(-1, 3, 3))
    nx, ny, nz = grid_extent(lower, upper, dx)
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    t = torch.from_numpy(tris).to(dev)
    out = torch.empty((nx, ny, nz), dtype=torch.int8, device=dev)
    with torch.cuda.device(dev):
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        lib = _lib.lib()
        rc = lib.cdll.lbm3d_voxelize(stream, ctypes.c_void_p(t.data_ptr()), len(tris), float(lower[0]),
                                     float(lower[1]), float(lower[2]), float(dx), nx, ny, nz,
                                     ctypes.c_void_p(out.data_ptr()))
        if rc != 0:
            raise _lib.LbmError(f"lbm3d_voxelize failed ({rc}): {lib.cdll.lbm3d_io_last_error().decode()}")
    return out.cpu().numpy()


def voxelize_stl(path: str, lower, upper, dx: float, device=None) -> np.ndarray:
    return voxelize(read_stl(path), lower, upper, dx, device)


def write_vtk_flags(path: str, flags: np.ndarray) -> None:
    """The reference converter's check file: legacy ASCII VTK STRUCTURED_POINTS,
    x fastest (Convert_stl_to_binary.cpp:131-153)."""
    f = np.asarray(flags)
    nx, ny, nz = f.shape
    with open(path, "w") as fh:
        fh.write("# vtk DataFile Version 2.0\n")
        fh.write("Lattice Boltzmann 3D single phase solid geometry\n")
        fh.write("ASCII\nDATASET STRUCTURED_POINTS\n")
        fh.write(f"DIMENSIONS {nx} {ny} {nz}\nORIGIN 0 0 0\nSPACING 1 1 1\n")
        fh.write(f"POINT_DATA {nx * ny * nz}\nSCALARS sample_scalars float\nLOOKUP_TABLE default\n")
        fh.write(" ".join(str(int(v)) for v in f.reshape(-1, order="F")))
        fh.write("\n")


def box_mesh(lo, hi) -> np.ndarray:
    """Closed, outward-oriented triangle mesh (12 triangles) of an axis-aligned box."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
         (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (3, 0, 4), (3, 4, 7)]
    return c[np.array(f)]


def sphere_mesh(center, radius: float, subdivisions: int = 3) -> np.ndarray:
    """Closed, outward-oriented icosphere."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdivisions):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    vv = np.array(verts) * radius + np.asarray(center, float)
    return vv[np.array(f)].astype(np.float32)
