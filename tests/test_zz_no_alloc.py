"""lbm3d_reduce_noalloc, the reduction entry point that allocates nothing (its
partials live in a device array inside the library), against lbm3d_reduce
with a caller workspace: the same ten FP64 results bit for bit, on two calls
in a row, the second on a smaller input, and both against numpy FP64 sums."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def solver(shape, seed, steps):
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    from taichi_lbm3d_amd.geometry import sphere_pack
    solid = sphere_pack(shape, porosity=0.4, rmin=3.0, rmax=7.0, seed=seed, periodic=(False, True, True))
    s = LB3D_Solver_Single_Phase(*shape, sparse_storage=True)
    s.solid.from_numpy(solid)
    s.niu = 0.1
    s.set_bc_rho_x0(1.001)
    s.set_bc_rho_x1(1.0)
    s.set_force([1e-5, 2e-6, 0.0])
    s.init_simulation()
    s.run(steps)
    return s


def both(s, i_lo, i_hi):
    import torch
    from taichi_lbm3d_amd import _lib
    lib = _lib.lib()
    rho, v = s.compact_macros()
    cells = s._dev["cells"]
    dev = rho.device
    vp = ctypes.c_void_p
    stream = vp(torch.cuda.current_stream(dev).cuda_stream)
    p = ctypes.byref(s._params)
    ws_bytes = lib.size("lbm3d_reduce_workspace_bytes", s.n_fluid)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    a = torch.full((10,), -7.0, dtype=torch.float64, device=dev)
    b = torch.full((10,), -7.0, dtype=torch.float64, device=dev)
    lib.call("lbm3d_reduce", stream, p, vp(cells.data_ptr()), vp(rho.data_ptr()), vp(v.data_ptr()), i_lo, i_hi,
             vp(a.data_ptr()), vp(ws.data_ptr()), ws_bytes)
    lib.call("lbm3d_reduce_noalloc", stream, p, vp(cells.data_ptr()), vp(rho.data_ptr()), vp(v.data_ptr()), i_lo,
             i_hi, vp(b.data_ptr()))
    torch.cuda.synchronize()
    n = s.n_fluid
    i = cells[:n].long().cpu().numpy() // (s.ny * s.nz)
    r64 = rho[:n].double().cpu().numpy()
    v64 = v[: 3 * n].double().cpu().numpy().reshape(3, n)
    return a.cpu().numpy(), b.cpu().numpy(), r64, v64, i


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_reduce_matches_workspace_reduce(gpu, order):
    big, small = solver((120, 96, 80), 3, 30), solver((14, 12, 10), 4, 9)
    runs = [(big, 5, 90), (small, 2, 9)] if order == "large_then_small" else [(small, 0, 13), (big, 0, 119)]
    for s, lo, hi in runs:
        a, b, r64, v64, i = both(s, lo, hi)
        assert s.n_fluid > 0
        np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64))
        sl = (i >= lo) & (i <= hi)
        assert np.isclose(b[0], r64.sum(), rtol=1e-12, atol=0)
        np.testing.assert_allclose(b[1:4], v64.sum(axis=1), rtol=1e-9, atol=1e-15)
        assert b[4] == np.sqrt((v64 ** 2).sum(axis=0)).max() or np.isclose(b[4], np.sqrt((v64 ** 2).sum(axis=0)).max(),
                                                                            rtol=1e-15)
        assert b[5] == 0 and b[7] == sl.sum() and b[9] == 0
        assert np.isclose(b[6], v64[0][sl].sum(), rtol=1e-9, atol=1e-15)
        assert np.isclose(b[8], r64[sl].sum(), rtol=1e-12, atol=0)
