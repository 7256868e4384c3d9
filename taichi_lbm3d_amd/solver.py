"""LB3D_Solver_Single_Phase on MI355X: the reference's Python surface over liblbm3d.so.

Mirrors ``Single_phase/LBM_3D_SinglePhase_Solver.py`` (SP), class
``LB3D_Solver_Single_Phase`` (SP:9-481): same constructor, attribute names
and defaults (SP:13-28), setters (SP:405-458), ``init_geo`` (SP:173-177),
``init_simulation`` (SP:118-149), ``step`` (SP:477-481), ``get_max_v``
(SP:394-402), ``export_VTK`` (SP:462-475) and the ``solid`` / ``rho`` / ``v``
fields with ``from_numpy`` / ``to_numpy``.

Device memory is held as PyTorch tensors; every computation runs in the HIP
kernels of liblbm3d.so (there is no CPU path).
"""
from __future__ import annotations

import math
import os
from typing import Optional

import numpy as np

from . import _lib
from . import geometry as _geometry
from . import vtk as _vtk

FACE_NAMES = ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")
_FACE_ATTRS = (("bc_x_left", "rho_bcxl", "vx_bcxl", "vy_bcxl", "vz_bcxl"),
               ("bc_x_right", "rho_bcxr", "vx_bcxr", "vy_bcxr", "vz_bcxr"),
               ("bc_y_left", "rho_bcyl", "vx_bcyl", "vy_bcyl", "vz_bcyl"),
               ("bc_y_right", "rho_bcyr", "vx_bcyr", "vy_bcyr", "vz_bcyr"),
               ("bc_z_left", "rho_bczl", "vx_bczl", "vy_bczl", "vz_bczl"),
               ("bc_z_right", "rho_bczr", "vx_bczr", "vy_bczr", "vz_bczr"))

# relaxation time and force-term scalings per convention
#   reference : tau = niu/3 + 1/2, Guo term with (e-u).F/3 and (e.u)(e.F)/9   (SP:127, SP:236)
#   documented: tau = 3 niu + 1/2, Guo term unscaled                          (SP:126 commented form)
#   textbook  : tau = 3 niu + 1/2, Guo term 3 (e-u).F + 9 (e.u)(e.F)
CONVENTIONS = {"reference": (1.0 / 3.0, 1.0 / 9.0), "documented": (1.0, 1.0), "textbook": (3.0, 9.0)}


def relaxation_rates(niu: float, convention: str = "reference"):
    """(tau, s_v, s_other) of init_simulation (SP:126-129), in Python doubles."""
    if convention not in CONVENTIONS:
        raise ValueError(f"unknown convention {convention!r}")
    tau = niu / 3.0 + 0.5 if convention == "reference" else 3.0 * niu + 0.5
    s_v = 1.0 / tau
    s_other = 8.0 * (2.0 - s_v) / (8.0 - s_v)
    return tau, s_v, s_other


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("LB3D_Solver_Single_Phase needs a ROCm GPU (torch.cuda is not available)")
    return torch


class _SolidField:
    """The ``solid`` field (ti.i8, shape (nx,ny,nz), SP:57)."""

    def __init__(self, owner):
        self._o = owner
        self.shape = (owner.nx, owner.ny, owner.nz)
        self._data = np.zeros(self.shape, np.int8)

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr)
        if a.shape != self.shape:
            raise ValueError(f"solid.from_numpy: shape {a.shape} != {self.shape}")
        self._data = np.ascontiguousarray(a.astype(np.int8))
        self._o._geometry_changed()

    def to_numpy(self) -> np.ndarray:
        return self._data.copy()

    def __getitem__(self, idx):
        return self._data[idx]

    def __setitem__(self, idx, val):
        self._data[idx] = val
        self._o._geometry_changed()


class _MacroField:
    """``rho`` (SP:34/39) or ``v`` (SP:35/40) as to_numpy() sees them."""

    def __init__(self, owner, kind: str):
        self._o = owner
        self._kind = kind

    @property
    def shape(self):
        return self._o._field_shape()

    def to_numpy(self) -> np.ndarray:
        rho, v = self._o._dense_macros()
        return rho if self._kind == "rho" else v


class _Scalar:
    """0-d field such as ``max_v`` (SP:20), indexed with ``[None]``."""

    def __init__(self, value=0.0):
        self.value = value

    def __getitem__(self, idx):
        return self.value

    def __setitem__(self, idx, val):
        self.value = val


class LB3D_Solver_Single_Phase:
    """Single-phase D3Q19 MRT solver with the reference class's surface (SP:9-481).

    Extra keyword arguments (not in the reference):
      convention -- "reference" (SP's tau = niu/3+1/2 and /3, /9 force term;
                    the default), "documented" or "textbook";
      dtype      -- population storage "f32" (default), "f16" or "bf16",
                    stored shifted as f - w with FP32 arithmetic ("f16sr" /
                    "bf16sr": the same with stochastic rounding of every
                    store); or "f64", FP64 storage and arithmetic (the FP64
                    GPU reference run);
      device     -- torch device index or string (default: current device).
    """

    def __init__(self, nx, ny, nz, sparse_storage=False, *, convention="reference", dtype="f32",
                 device=None):
        if convention not in CONVENTIONS:
            raise ValueError(f"unknown convention {convention!r}")
        if dtype not in _lib.DTYPES:
            raise ValueError(f"dtype must be one of {sorted(_lib.DTYPES)}")
        self.enable_projection = True
        self.sparse_storage = bool(sparse_storage)
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        if min(self.nx, self.ny, self.nz) < 1:
            raise ValueError("grid extents must be >= 1")
        self.fx, self.fy, self.fz = 0.0e-6, 0.0, 0.0
        self.niu = 0.16667
        self.max_v = _Scalar(0.0)
        # Boundary condition mode: 0 periodic, 1 fixed pressure, 2 fixed velocity (SP:22-28)
        self.bc_x_left, self.rho_bcxl, self.vx_bcxl, self.vy_bcxl, self.vz_bcxl = 0, 1.0, 0.0e-5, 0.0, 0.0
        self.bc_x_right, self.rho_bcxr, self.vx_bcxr, self.vy_bcxr, self.vz_bcxr = 0, 1.0, 0.0, 0.0, 0.0
        self.bc_y_left, self.rho_bcyl, self.vx_bcyl, self.vy_bcyl, self.vz_bcyl = 0, 1.0, 0.0, 0.0, 0.0
        self.bc_y_right, self.rho_bcyr, self.vx_bcyr, self.vy_bcyr, self.vz_bcyr = 0, 1.0, 0.0, 0.0, 0.0
        self.bc_z_left, self.rho_bczl, self.vx_bczl, self.vy_bczl, self.vz_bczl = 0, 1.0, 0.0, 0.0, 0.0
        self.bc_z_right, self.rho_bczr, self.vx_bczr, self.vy_bczr, self.vz_bczr = 0, 1.0, 0.0, 0.0, 0.0
        self.convention = convention
        self.dtype = dtype
        self.solid = _SolidField(self)
        self.rho = _MacroField(self, "rho")
        self.v = _MacroField(self, "v")
        # coordinates handed to the VTK writer (SP:112-114)
        self.x = np.linspace(0, self.nx, self.nx)
        self.y = np.linspace(0, self.ny, self.ny)
        self.z = np.linspace(0, self.nz, self.nz)
        self._device = device
        # odd steps locate their cells through tile descriptors (LBM3D_FLAG_TILES);
        # results are identical either way
        self.use_tiles = True
        # order of the fluid list: bands of this many grid rows j, band after
        # band, plane after plane inside a band (0: C order).  Odd steps then
        # find a cell's neighbours in planes i +- 1 32 rows of cells away
        # instead of a whole plane, so more of the slot lines they share are
        # still in the XCD's L2.  Results are identical either way.
        self.band_rows = 32
        self._initialised = False
        self._compiled = False  # parameters frozen by the first step() (SP kernels compile then)
        self._t = 0
        self._macro_cache = None
        self._params = _lib.Params()
        self._dev = {}

    # ---- setters, SP:405-458 ------------------------------------------------
    def set_bc_vel_x1(self, vel):
        self.bc_x_right = 2
        self.vx_bcxr, self.vy_bcxr, self.vz_bcxr = vel[0], vel[1], vel[2]

    def set_bc_vel_x0(self, vel):
        self.bc_x_left = 2
        self.vx_bcxl, self.vy_bcxl, self.vz_bcxl = vel[0], vel[1], vel[2]

    def set_bc_vel_y1(self, vel):
        self.bc_y_right = 2
        self.vx_bcyr, self.vy_bcyr, self.vz_bcyr = vel[0], vel[1], vel[2]

    def set_bc_vel_y0(self, vel):
        self.bc_y_left = 2
        self.vx_bcyl, self.vy_bcyl, self.vz_bcyl = vel[0], vel[1], vel[2]

    def set_bc_vel_z1(self, vel):
        self.bc_z_right = 2
        self.vx_bczr, self.vy_bczr, self.vz_bczr = vel[0], vel[1], vel[2]

    def set_bc_vel_z0(self, vel):
        self.bc_z_left = 2
        self.vx_bczl, self.vy_bczl, self.vz_bczl = vel[0], vel[1], vel[2]

    def set_bc_rho_x0(self, rho):
        self.bc_x_left = 1
        self.rho_bcxl = rho

    def set_bc_rho_x1(self, rho):
        self.bc_x_right = 1
        self.rho_bcxr = rho

    def set_bc_rho_y0(self, rho):
        self.bc_y_left = 1
        self.rho_bcyl = rho

    def set_bc_rho_y1(self, rho):
        self.bc_y_right = 1
        self.rho_bcyr = rho

    def set_bc_rho_z0(self, rho):
        self.bc_z_left = 1
        self.rho_bczl = rho

    def set_bc_rho_z1(self, rho):
        self.bc_z_right = 1
        self.rho_bczr = rho

    def set_viscosity(self, niu):
        self.niu = niu

    def set_force(self, force):
        self.fx, self.fy, self.fz = force[0], force[1], force[2]

    # ---- geometry ----------------------------------------------------------
    def init_geo(self, filename):
        """ASCII 0/1 voxel file, x fastest; values > 0 are solid (SP:173-177)."""
        self.solid.from_numpy(_geometry.read_geo(filename, self.nx, self.ny, self.nz))

    def _geometry_changed(self):
        if self._compiled:
            raise RuntimeError("the geometry is frozen once the simulation has stepped")
        self._initialised = False

    # ---- parameter freezing --------------------------------------------------
    def _bc_modes(self):
        return [int(getattr(self, a[0])) for a in _FACE_ATTRS]

    def _freeze_init(self):
        """What init_simulation() fixes (SP:118-140): S, force_flag, face velocities."""
        p = self._params
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        self.tau_f, self.s_v, self.s_other = relaxation_rates(float(self.niu), self.convention)
        p.s_v, p.s_o = self.s_v, self.s_other
        a, b = CONVENTIONS[self.convention]
        p.guo_a, p.guo_b = a, b
        self.force_flag = 1 if (abs(self.fx) > 0 or abs(self.fy) > 0 or abs(self.fz) > 0) else 0
        p.force_on = self.force_flag
        p.flags = _lib.STORAGE_FLAGS.get(self.dtype, 0)
        self.bc_vel_x_left = [self.vx_bcxl, self.vy_bcxl, self.vz_bcxl]
        self.bc_vel_x_right = [self.vx_bcxr, self.vy_bcxr, self.vz_bcxr]
        self.bc_vel_y_left = [self.vx_bcyl, self.vy_bcyl, self.vz_bcyl]
        self.bc_vel_y_right = [self.vx_bcyr, self.vy_bcyr, self.vz_bcyr]
        self.bc_vel_z_left = [self.vx_bczl, self.vy_bczl, self.vz_bczl]
        self.bc_vel_z_right = [self.vx_bczr, self.vy_bczr, self.vz_bczr]
        vels = (self.bc_vel_x_left, self.bc_vel_x_right, self.bc_vel_y_left, self.bc_vel_y_right,
                self.bc_vel_z_left, self.bc_vel_z_right)
        for f in range(6):
            for c in range(3):
                p.u_bc[f][c] = float(vels[f][c])
        self._freeze_compile()

    def _freeze_compile(self):
        """What the first kernel compilation bakes in (SP:219, SP:230, SP:274-365, SP:385)."""
        p = self._params
        modes = self._bc_modes()
        for f, attrs in enumerate(_FACE_ATTRS):
            m = modes[f]
            n = (self.nx, self.ny, self.nz)[f // 2]
            if m in (1, 2) and n < 2:
                raise ValueError(f"face {FACE_NAMES[f]} has a boundary condition but extent {n} < 2")
            p.bc_mode[f] = m
            p.rho_bc[f] = float(getattr(self, attrs[1]))
        # state of each pressure face: 3 floats per node of its plane
        off = 0
        planes = (self.ny * self.nz, self.ny * self.nz, self.nx * self.nz, self.nx * self.nz,
                  self.nx * self.ny, self.nx * self.ny)
        for f in range(6):
            if modes[f] == 1:
                p.face_u_off[f] = off
                off += 3 * planes[f]
            else:
                p.face_u_off[f] = -1
        self._face_u_floats = off
        p.fx, p.fy, p.fz = float(self.fx), float(self.fy), float(self.fz)
        return modes

    # ---- device state --------------------------------------------------------
    def _torch_device(self):
        torch = _torch()
        if self._device is None:
            self._device = torch.device("cuda", torch.cuda.current_device())
        return torch.device(self._device) if not isinstance(self._device, torch.device) else self._device

    def _stream(self):
        torch = _torch()
        return ctypes_void(torch.cuda.current_stream(self._torch_device()).cuda_stream)

    def _build_lists(self):
        """Fluid list (in the order of band_rows), row table and face bits on the GPU."""
        torch = _torch()
        lib = _lib.lib()
        dev = self._torch_device()
        p = self._params
        n_dense = self.nx * self.ny * self.nz
        solid = torch.from_numpy(np.ascontiguousarray(self.solid._data.reshape(-1))).to(dev)
        ws_bytes = lib.size("lbm3d_compact_workspace_bytes", n_dense)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        count = torch.zeros(1, dtype=torch.int32, device=dev)
        p.band_rows = int(self.band_rows)
        with torch.cuda.device(dev):
            s = self._stream()
            lib.call("lbm3d_compact_count", s, _ref(p), _ptr(solid), _ptr(count), _ptr(ws), ws_bytes)
            p.n_cells = int(count.cpu()[0])
            d2c = torch.empty(n_dense, dtype=torch.int32, device=dev)
            cells = torch.empty(max(p.n_cells, 1), dtype=torch.int32, device=dev)
            lib.call("lbm3d_compact_fill", s, _ref(p), _ptr(solid), _ptr(d2c), _ptr(cells), _ptr(ws), ws_bytes)
            # row table, then the odd step's tile descriptors (LBM3D_FLAG_TILES)
            n_row = lib.size("lbm3d_row_words", _ref(p))
            n_tile = lib.size("lbm3d_tile_words", _ref(p))
            rows = torch.empty(max(4 * (n_row + n_tile), 4), dtype=torch.int32, device=dev)
            lib.call("lbm3d_build_rows", s, _ref(p), _ptr(solid), _ptr(d2c), _ptr(rows))
            if n_tile and self.use_tiles:
                lib.call("lbm3d_build_tiles", s, _ref(p), _ptr(cells), _ptr(rows[4 * n_row:]))
                p.flags |= _lib.FLAG_TILES
            else:
                p.flags &= ~_lib.FLAG_TILES
            face_bits = torch.zeros(max(lib.size("lbm3d_face_bit_words", p.n_cells), 1), dtype=torch.int32,
                                    device=dev)
            lib.call("lbm3d_build_face_bits", s, _ref(p), _ptr(solid), _ptr(cells), _ptr(face_bits))
            torch.cuda.current_stream(dev).synchronize()
        del d2c, ws, solid
        self._dev.update(cells=cells, rows=rows, face_bits=face_bits)
        self._modes_built = self._bc_modes()

    def neighbor_table(self) -> np.ndarray:
        """The explicit neighbour table (18, n_fluid), built on the GPU with the
        lookup the step kernel uses: compact id of the cell at x + e_q, -1 if
        that cell is not fluid (streaming1 bounce-back, SP:259-269)."""
        torch = _torch()
        p = self._params
        dev = self._torch_device()
        nbr = torch.empty(max(18 * p.n_cells, 1), dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            _lib.lib().call("lbm3d_build_neighbors", self._stream(), _ref(p), _ptr(self._dev["cells"]),
                            _ptr(self._dev["rows"]), _ptr(nbr))
        return nbr[: 18 * p.n_cells].cpu().numpy().reshape(18, p.n_cells)

    def _alloc_state(self):
        torch = _torch()
        dev = self._torch_device()
        p = self._params
        tdt = _lib.torch_storage_dtype(self.dtype)
        stride = _lib.lib().size("lbm3d_slot_stride", p.n_cells)
        self._dev["g"] = torch.zeros(max(19 * stride, 1), dtype=tdt, device=dev)
        self._dev["face_u"] = torch.zeros(max(self._face_u_floats, 1), dtype=self._real_dtype(), device=dev)

    def _real_dtype(self):
        """Arithmetic type of the kernels: FP64 for f64 storage, FP32 otherwise."""
        import torch
        return torch.float64 if self.dtype == "f64" else torch.float32

    def _init_state(self, rho0=None, u0=None):
        """init kernel (SP:160-170): equilibrium at (rho0, u0), default rest state."""
        torch = _torch()
        lib = _lib.lib()
        dev = self._torch_device()
        p = self._params
        with torch.cuda.device(dev):
            lib.call("lbm3d_init", self._stream(), _ref(p), _lib.DTYPES[self.dtype], _ptr(self._dev["g"]),
                     _ptr(self._dev["cells"]), _ptr(self._dev["face_u"]),
                     _ptr(rho0) if rho0 is not None else None, _ptr(u0) if u0 is not None else None)
        self._t = 0
        self._macro_cache = None

    def init_simulation(self):
        """Freeze parameters and initialise the populations (SP:118-149)."""
        _torch()
        self._compiled = False
        self._freeze_init()
        self._build_lists()
        self._alloc_state()
        self._init_state()
        self._initialised = True

    def init_from_macros(self, rho, v):
        """Warm start: f = feq(rho, v) on every fluid cell (replaces init, SP:160-170).

        ``rho`` (nx,ny,nz) and ``v`` (nx,ny,nz,3) dense arrays; the first step
        then collides with exactly these macros, as the reference's first step
        collides with init's rho = 1, v = 0.
        """
        torch = _torch()
        if not self._initialised:
            self.init_simulation()
        dev = self._torch_device()
        cells = self._dev["cells"][: self._params.n_cells].long()
        rho_d = torch.as_tensor(np.asarray(rho, np.float32).reshape(-1), device=dev)
        v_d = torch.as_tensor(np.asarray(v, np.float32).reshape(-1, 3), device=dev)
        rho_c = rho_d[cells].contiguous()
        u_c = v_d[cells].t().contiguous()
        self._init_state(rho_c, u_c)

    # ---- time stepping --------------------------------------------------------
    def _prepare_first_step(self):
        if not self._initialised:
            raise RuntimeError("call init_simulation() before step()")
        if not self._compiled:
            modes = self._freeze_compile()
            if modes != self._modes_built:
                # boundary modes changed after init_simulation: number the faces again
                self._build_lists()
                self._alloc_state()
                self._init_state()
            self._compiled = True

    def step(self):
        """One time step: colission, streaming1, Boundary_condition, streaming3 (SP:477-481)."""
        self.run(1)

    def run(self, nsteps: int):
        """``nsteps`` calls of step(), enqueued back to back on the current stream."""
        self._prepare_first_step()
        torch = _torch()
        p = self._params
        d = self._dev
        args = (self._stream(), _ref(p), _lib.DTYPES[self.dtype], _ptr(d["g"]), _ptr(d["cells"]), _ptr(d["rows"]),
                _ptr(d["face_bits"]), _ptr(d["face_u"]), self._t, int(nsteps))
        with torch.cuda.device(self._torch_device()):
            _lib.lib().call("lbm3d_run", *args)
        self._t += int(nsteps)
        self._macro_cache = None

    @property
    def steps_done(self) -> int:
        return self._t

    @property
    def n_fluid(self) -> int:
        return int(self._params.n_cells)

    # ---- macroscopic fields ---------------------------------------------------
    def compact_macros(self, native: bool = False):
        """(rho [n], v [3*n]) FP32 torch tensors over the fluid list, after steps_done
        steps; with native=True in the kernels' arithmetic type (FP64 for f64)."""
        if native and self.dtype == "f64":
            return self._macros_native()
        return self._macros_f32()

    def _macros_native(self):
        torch = _torch()
        p = self._params
        d = self._dev
        dev = self._torch_device()
        n = max(p.n_cells, 1)
        rho = torch.empty(n, dtype=self._real_dtype(), device=dev)
        v = torch.empty(3 * n, dtype=self._real_dtype(), device=dev)
        with torch.cuda.device(dev):
            _lib.lib().call("lbm3d_macros", self._stream(), _ref(p), _lib.DTYPES[self.dtype], _ptr(d["g"]),
                            _ptr(d["cells"]), _ptr(d["rows"]), _ptr(d["face_bits"]), _ptr(d["face_u"]), self._t,
                            _ptr(rho), _ptr(v))
        return rho, v

    def _macros_f32(self):
        if self._macro_cache is not None:
            return self._macro_cache
        if not self._initialised:
            raise RuntimeError("call init_simulation() first")
        if self.dtype == "f64":
            rho, v = self._macros_native()
            self._macro_cache = (rho.float(), v.float())
            return self._macro_cache
        torch = _torch()
        p = self._params
        d = self._dev
        dev = self._torch_device()
        n = max(p.n_cells, 1)
        rho = torch.empty(n, dtype=torch.float32, device=dev)
        v = torch.empty(3 * n, dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _lib.lib().call("lbm3d_macros", self._stream(), _ref(p), _lib.DTYPES[self.dtype], _ptr(d["g"]),
                            _ptr(d["cells"]), _ptr(d["rows"]), _ptr(d["face_bits"]), _ptr(d["face_u"]), self._t,
                            _ptr(rho), _ptr(v))
        self._macro_cache = (rho, v)
        return self._macro_cache

    def _field_shape(self):
        if self.sparse_storage:  # pointer((n//3+1)^3) -> dense(3^3) (SP:41-44)
            return (3 * (self.nx // 3 + 1), 3 * (self.ny // 3 + 1), 3 * (self.nz // 3 + 1))
        return (self.nx, self.ny, self.nz)

    def _dense_macros(self):
        """rho and v as to_numpy() returns them.

        Fluid cells hold streaming3's values.  Other cells: dense storage has
        rho = 1, v = 0 everywhere (SP:390-392); sparse storage has that inside
        active 3^3 blocks (those holding a fluid cell, activated by init,
        SP:164-169) and 0 elsewhere, over the padded extent.
        """
        shape = self._field_shape()
        if not self._initialised:
            return np.zeros(shape, np.float32), np.zeros(shape + (3,), np.float32)
        torch = _torch()
        dev = self._torch_device()
        fluid = self.solid._data == 0
        if self.sparse_storage:
            bx, by, bz = (s // 3 for s in shape)
            pad = np.zeros(shape, bool)
            pad[: self.nx, : self.ny, : self.nz] = fluid
            active = pad.reshape(bx, 3, by, 3, bz, 3).any(axis=(1, 3, 5))
            base = np.repeat(np.repeat(np.repeat(active, 3, 0), 3, 1), 3, 2).astype(np.float32)
        else:
            base = np.ones(shape, np.float32)
        rho_c, v_c = self.compact_macros()
        p = self._params
        rho_d = torch.from_numpy(base).to(dev)
        v_d = torch.zeros(shape + (3,), dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _lib.lib().call("lbm3d_scatter_dense", self._stream(), _ref(p), _ptr(self._dev["cells"]),
                            _ptr(rho_c), _ptr(v_c), 0, shape[0], shape[1], shape[2], _ptr(rho_d), _ptr(v_d))
        return rho_d.cpu().numpy(), v_d.cpu().numpy()

    def reduce(self, i_lo: int = 0, i_hi: Optional[int] = None):
        """FP64 sums over the fluid list: dict of sum_rho, sum_v (3), max_v,
        nonfinite, and over the slabs i_lo <= i <= i_hi slab_sum_vx,
        slab_count and slab_sum_rho."""
        torch = _torch()
        p = self._params
        dev = self._torch_device()
        # FP64 storage sums its FP64 fields; the others their FP32 fields
        native = self.dtype == "f64"
        rho_c, v_c = self.compact_macros(native=native)
        lib = _lib.lib()
        ws_bytes = lib.size("lbm3d_reduce_workspace_bytes", p.n_cells)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        out = torch.zeros(10, dtype=torch.float64, device=dev)
        with torch.cuda.device(dev):
            lib.call("lbm3d_reduce64" if native else "lbm3d_reduce", self._stream(), _ref(p),
                     _ptr(self._dev["cells"]), _ptr(rho_c), _ptr(v_c), int(i_lo),
                     int(self.nx - 1 if i_hi is None else i_hi), _ptr(out), _ptr(ws), ws_bytes)
        o = out.cpu().numpy()
        return {"sum_rho": o[0], "sum_v": o[1:4].copy(), "max_v": o[4], "nonfinite": int(o[5]),
                "slab_sum_vx": o[6], "slab_count": int(o[7]), "slab_sum_rho": o[8]}

    def get_max_v(self):
        """max |v| over the domain (cal_max_v, SP:394-402)."""
        r = self.reduce()
        val = float(r["max_v"]) if self._params.n_cells > 0 else 0.0
        self.max_v[None] = val
        return val

    # ---- permeability and steady state ------------------------------------------
    def effective_viscosity(self) -> float:
        """Kinematic viscosity the lattice realises, (tau - 1/2)/3."""
        tau, _, _ = relaxation_rates(float(self.niu), self.convention)
        return (tau - 0.5) / 3.0

    def permeability(self, method: str = "faces") -> float:
        """Darcy permeability along x in lattice units^2.

        Pressure drive (x faces in mode 1), method "faces":
            k = nu_eff * q * (nx - 1) / ((rho_left - rho_right) / 3),
        q the mean superficial x velocity over the slabs 1 <= i <= nx-2: the
        apparent permeability of the whole sample between the two face
        densities.  The reference's pressure faces hold their nodes at
        feq(rho_bc, 0) (SP:274-297), so part of the density drop sits at the
        faces themselves and this k includes their resistance.
        Method "gradient" takes the drop inside the sample instead:
            k = rho_m * nu_eff * q / G,  G = (rho_a - rho_b) / (3 (b - a)),
        rho_a, rho_b the mean density of the fluid cells of the slabs
        a = nx // 4 and b = nx - 1 - nx // 4, q the superficial mean x velocity
        and rho_m the mean density over the slabs a .. b (dynamic viscosity
        rho nu).
        Body force (fx != 0):
            k = nu_eff * q / F_eff,  F_eff = a * fx / 3 the momentum the
        force term injects per step (a = 1/3 for the reference convention),
        q over the whole box, measured with the momentum-consistent velocity
        (j + F_eff/2)/rho rather than the reported v = (j + fx/2)/rho.
        """
        p = self._params
        nu = self.effective_viscosity()
        if method not in ("faces", "gradient"):
            raise ValueError("method must be 'faces' or 'gradient'")
        if p.bc_mode[0] == 1 and p.bc_mode[1] == 1 and method == "gradient":
            a, b = self.nx // 4, self.nx - 1 - self.nx // 4
            if b - a < 1:
                raise ValueError("the gradient method needs nx >= 4")
            ra, rb = self.reduce(a, a), self.reduce(b, b)
            if ra["slab_count"] == 0 or rb["slab_count"] == 0:
                raise ValueError("a slab of the gradient method holds no fluid cell")
            r = self.reduce(a, b)
            q = r["slab_sum_vx"] / (self.ny * self.nz * (b - a + 1))
            rho_m = r["slab_sum_rho"] / r["slab_count"]
            g = (ra["slab_sum_rho"] / ra["slab_count"] - rb["slab_sum_rho"] / rb["slab_count"]) / (3.0 * (b - a))
            return rho_m * nu * q / g
        if method == "gradient":
            raise ValueError("the gradient method needs pressure faces on both x faces")
        if p.bc_mode[0] == 1 and p.bc_mode[1] == 1:
            r = self.reduce(1, self.nx - 2)
            q = r["slab_sum_vx"] / (self.ny * self.nz * max(self.nx - 2, 1))
            dp = (p.rho_bc[0] - p.rho_bc[1]) / 3.0
            return nu * q * (self.nx - 1) / dp
        if self.fx != 0:
            a, _ = CONVENTIONS[self.convention]
            r = self.reduce()
            f_eff = a * self.fx / 3.0
            q = (r["sum_v"][0] - 0.5 * (self.fx - f_eff) * p.n_cells) / (self.nx * self.ny * self.nz)
            return nu * q / f_eff
        raise ValueError("permeability needs pressure faces on x or a body force fx")

    def mean_vx(self) -> float:
        r = self.reduce()
        return float(r["sum_v"][0]) / max(self._params.n_cells, 1)

    def run_to_steady(self, tol: float = 1e-6, interval: int = 500, max_steps: int = 50_000):
        """Step until |d<vx>| / |<vx>| < tol between checks `interval` steps apart.

        Returns (steps, converged, history of <vx>).  Raises on non-finite fields.
        """
        hist = []
        prev = None
        while self._t < max_steps:
            self.run(min(interval, max_steps - self._t))
            r = self.reduce()
            if r["nonfinite"]:
                raise FloatingPointError(f"{r['nonfinite']} non-finite cells after {self._t} steps")
            cur = float(r["sum_v"][0]) / max(self._params.n_cells, 1)
            hist.append(cur)
            if prev is not None and cur != 0.0 and abs(cur - prev) / abs(cur) < tol:
                return self._t, True, hist
            prev = cur
        return self._t, False, hist

    # ---- checkpoint / export ----------------------------------------------------------
    def save_checkpoint(self, path: str):
        """Populations (in their current AA parity), face velocities and step count."""
        torch = _torch()
        torch.save({"g": self._dev["g"].cpu(), "face_u": self._dev["face_u"].cpu(), "t": self._t,
                    "dtype": self.dtype, "n_cells": self._params.n_cells, "band_rows": self._params.band_rows},
                   path)

    def load_checkpoint(self, path: str):
        torch = _torch()
        ck = torch.load(path, weights_only=True)
        if (ck["n_cells"] != self._params.n_cells or ck["dtype"] != self.dtype
                or ck.get("band_rows", 0) != self._params.band_rows):
            raise ValueError("checkpoint does not match this solver's geometry, list order or dtype")
        self._prepare_first_step()
        self._dev["g"].copy_(ck["g"].to(self._dev["g"].device))
        self._dev["face_u"].copy_(ck["face_u"].to(self._dev["face_u"].device))
        self._t = int(ck["t"])
        self._macro_cache = None

    def export_VTK(self, n):
        """VTK XML rectilinear grid ``./LB_SingelPhase_<n>.vtr`` with point data
        Solid, rho and velocity (SP:462-475)."""
        rho, v = self._dense_macros()
        nx, ny, nz = self.nx, self.ny, self.nz
        return _vtk.grid_to_vtk("./LB_SingelPhase_" + str(n), self.x, self.y, self.z,
                                pointData={"Solid": np.ascontiguousarray(self.solid.to_numpy()),
                                           "rho": np.ascontiguousarray(rho[:nx, :ny, :nz]),
                                           "velocity": (np.ascontiguousarray(v[:nx, :ny, :nz, 0]),
                                                        np.ascontiguousarray(v[:nx, :ny, :nz, 1]),
                                                        np.ascontiguousarray(v[:nx, :ny, :nz, 2]))})


def ctypes_void(ptr: int):
    import ctypes
    return ctypes.c_void_p(int(ptr))


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def _ref(p):
    import ctypes
    return ctypes.byref(p)
