"""GPU tests of the STL voxelizer against the FP64 winding-number oracle and
closed-form interiors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def nodes(lower, shape, dx):
    i, j, k = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    return np.stack([lower[0] + dx * i, lower[1] + dx * j, lower[2] + dx * k], -1).reshape(-1, 3)


def test_voxelize_box_exact(gpu):
    from taichi_lbm3d_amd.stl import box_mesh, grid_extent, voxelize
    lo, hi, dx = (-1.03, -0.51, -2.07), (4.0, 3.0, 2.5), 0.25
    flags = voxelize(box_mesh((0.0, 0.0, 0.0), (2.0, 1.5, 1.0)), lo, hi, dx)
    assert flags.shape == grid_extent(lo, hi, dx) and flags.dtype == np.int8
    p = nodes(lo, flags.shape, dx).reshape(flags.shape + (3,))
    inside = np.all((p > 0) & (p < np.array([2.0, 1.5, 1.0])), axis=-1)
    np.testing.assert_array_equal(flags, inside.astype(np.int8))


def test_voxelize_sphere_matches_oracle(gpu, tmp_path):
    from oracle.geo_ref import winding_number
    from taichi_lbm3d_amd.stl import sphere_mesh, voxelize_stl, write_stl, write_vtk_flags
    tris = sphere_mesh((0.3, 0.1, -0.2), 1.7, subdivisions=3)
    path = str(tmp_path / "ball.stl")
    write_stl(path, tris)
    lo, hi, dx = (-2.0, -2.0, -2.5), (2.5, 2.0, 2.0), 0.1
    flags = voxelize_stl(path, lo, hi, dx)
    p = nodes(lo, flags.shape, dx)
    w = winding_number(tris, p)
    sure = np.abs(w - 0.5) > 1e-6
    assert sure.mean() > 0.99
    np.testing.assert_array_equal(flags.reshape(-1)[sure], (w[sure] >= 0.5).astype(np.int8))
    r = np.linalg.norm(p - np.array([0.3, 0.1, -0.2]), axis=1)
    assert np.all(flags.reshape(-1)[r < 1.6] == 1) and np.all(flags.reshape(-1)[r > 1.75] == 0)
    write_vtk_flags(str(tmp_path / "check.vtk"), flags)
    body = (tmp_path / "check.vtk").read_text().split("LOOKUP_TABLE default\n")[1].split()
    np.testing.assert_array_equal(np.array(body, int), flags.reshape(-1, order="F"))
