"""CPU oracle for the single-phase D3Q19 MRT solver (test infrastructure).

Two independent restatements of the reference solver
``Single_phase/LBM_3D_SinglePhase_Solver.py`` (class ``LB3D_Solver_Single_Phase``):

* :class:`COracle` -- ctypes wrapper over ``lbm_oracle.c``: the reference's
  unfused four-kernel step (SP:477-481) on a dense AoS grid, FP32 or FP64,
  optionally over active 3^3 blocks only (sparse mode, SP:41-44), OpenMP
  parallel.  It is also the CPU baseline the benchmark times.
* :mod:`oracle.oracle_np` -- a vectorised numpy restatement (``np.roll``
  streaming), used to cross-check the C code on small grids.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package.  The product path never does.

Pinning of each function (what fails if the restatement is wrong) is listed
in ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

FACES = ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")
CONVENTIONS = {"reference": 0, "documented": 1, "textbook": 2}


def build(force: bool = False) -> None:
    """Compile the C oracle (gcc, via the Makefile beside this file)."""
    libs = [os.path.join(_HERE, n) for n in ("liblbm_oracle32.so", "liblbm_oracle64.so")]
    if force or not all(os.path.exists(p) for p in libs):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)


def build_native() -> str:
    """Compile the CPU baseline's build of the FP32 restatement (-O3
    -march=native) for the host this runs on; returns the compiler flags."""
    subprocess.run(["make", "-s", "-B", "native"], cwd=_HERE, check=True)
    out = subprocess.run(["make", "-s", "-p", "-n", "native"], cwd=_HERE, capture_output=True, text=True).stdout
    for line in out.splitlines():
        if line.startswith("CFLAGS_NATIVE"):
            return line.split("=", 1)[1].strip()
    return ""


_LIBS: dict = {}
# real -> (library file, symbol prefix); "f32native" is the CPU baseline's
# host-tuned FP32 build (build_native), never the checker
_FILES = {"f32": ("liblbm_oracle32.so", "orc32_"), "f64": ("liblbm_oracle64.so", "orc64_"),
          "f32native": ("liblbm_oracle32n.so", "orc32_")}


def _lib(real: str):
    if real not in _FILES:
        raise ValueError(f"real must be one of {sorted(_FILES)}")
    if real in _LIBS:
        return _LIBS[real]
    name, p = _FILES[real]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        if real == "f32native":
            raise FileNotFoundError(f"{path}: call oracle.build_native() on this host first")
        build()
    lib = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)

    def fn(n, res, args):
        f = getattr(lib, p + n)
        f.restype = res
        f.argtypes = args
        return f

    api = {
        "create": fn("create", vp, [ip, ip, ip, ip]),
        "destroy": fn("destroy", None, [vp]),
        "set_threads": fn("set_threads", None, [vp, ip]),
        "set_solid": fn("set_solid", None, [vp, ctypes.c_void_p]),
        "set_params": fn("set_params", None, [vp, ctypes.c_double, ip, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_void_p, dp, dp]),
        "set_compile_params": fn("set_compile_params", None, [vp, ctypes.c_double, ctypes.c_double,
                                                              ctypes.c_double, ctypes.c_void_p, dp]),
        "init_simulation": fn("init_simulation", None, [vp]),
        "step": fn("step", None, [vp, ip]),
        "get_max_v": fn("get_max_v", ctypes.c_double, [vp]),
        "get_shape": fn("get_shape", None, [vp, ctypes.c_void_p]),
        "get_rho": fn("get_rho", None, [vp, dp]),
        "get_v": fn("get_v", None, [vp, dp]),
        "get_F": fn("get_F", None, [vp, dp]),
        "set_F": fn("set_F", None, [vp, dp]),
        "set_macros": fn("set_macros", None, [vp, dp, dp]),
        "get_invM": fn("get_invM", None, [vp, dp]),
        "get_M": fn("get_M", None, [vp, dp]),
        "n_active_blocks": fn("n_active_blocks", ctypes.c_int64, [vp]),
    }
    _LIBS[real] = api
    return api


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class COracle:
    """The C restatement driven with the reference's attribute surface.

    Attributes mirror SP:17-28 (``niu``, ``fx/fy/fz``, ``bc_x_left`` ...,
    ``rho_bcxl`` ..., ``vx_bcxl`` ...).  They are frozen in two stages, as
    Taichi freezes the reference's: ``init_simulation()`` fixes S from niu,
    force_flag and the face velocities (SP:118-140); the first ``step()``
    fixes what the kernels capture when they are compiled, fx/fy/fz
    (SP:217-220) and the face modes and densities (SP:274-365).  Changes
    after the first step have no effect.
    """

    def __init__(self, nx, ny, nz, sparse_storage=False, real="f32", convention="reference", threads=1):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.sparse_storage = bool(sparse_storage)
        self.real = real
        self.convention = convention
        self._api = _lib(real)
        self._h = self._api["create"](self.nx, self.ny, self.nz, int(self.sparse_storage))
        self._api["set_threads"](self._h, int(threads))
        self.niu = 0.16667
        self.fx, self.fy, self.fz = 0.0, 0.0, 0.0
        self.bc = [0] * 6
        self.rho_bc = [1.0] * 6
        self.u_bc = [[0.0, 0.0, 0.0] for _ in range(6)]
        self._solid = np.zeros((self.nx, self.ny, self.nz), np.int8)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._api["destroy"](h)
            self._h = None

    def set_threads(self, n):
        self._api["set_threads"](self._h, int(n))

    def set_solid(self, solid):
        s = np.ascontiguousarray(np.asarray(solid).astype(np.int8))
        if s.shape != (self.nx, self.ny, self.nz):
            raise ValueError("solid shape mismatch")
        self._solid = s
        self._api["set_solid"](self._h, s.ctypes.data)

    def set_bc_vel(self, face, vel):
        i = FACES.index(face)
        self.bc[i] = 2
        self.u_bc[i] = [float(x) for x in vel]

    def set_bc_rho(self, face, rho):
        i = FACES.index(face)
        self.bc[i] = 1
        self.rho_bc[i] = float(rho)

    def init_simulation(self):
        mode = np.array(self.bc, np.int32)
        rho = np.array(self.rho_bc, np.float64)
        u = np.array(self.u_bc, np.float64).reshape(-1)
        self._api["set_params"](self._h, float(self.niu), CONVENTIONS[self.convention], float(self.fx),
                                float(self.fy), float(self.fz), mode.ctypes.data, _dp(rho), _dp(u))
        self._api["init_simulation"](self._h)
        self._compiled = False

    def step(self, n=1):
        if not getattr(self, "_compiled", True):
            mode = np.array(self.bc, np.int32)
            rho = np.array(self.rho_bc, np.float64)
            self._api["set_compile_params"](self._h, float(self.fx), float(self.fy), float(self.fz),
                                            mode.ctypes.data, _dp(rho))
            self._compiled = True
        self._api["step"](self._h, int(n))

    def shape(self):
        s = np.zeros(3, np.int32)
        self._api["get_shape"](self._h, s.ctypes.data)
        return tuple(int(x) for x in s)

    def rho(self):
        out = np.zeros(self.shape(), np.float64)
        self._api["get_rho"](self._h, _dp(out))
        return out

    def v(self):
        out = np.zeros(self.shape() + (3,), np.float64)
        self._api["get_v"](self._h, _dp(out))
        return out

    def F(self):
        out = np.zeros(self.shape() + (19,), np.float64)
        self._api["get_F"](self._h, _dp(out))
        return out

    def set_F(self, F):
        a = np.ascontiguousarray(F, np.float64)
        if a.shape != self.shape() + (19,):
            raise ValueError("F shape mismatch")
        self._api["set_F"](self._h, _dp(a))

    def set_macros(self, rho, v):
        r = np.ascontiguousarray(rho, np.float64)
        u = np.ascontiguousarray(v, np.float64)
        self._api["set_macros"](self._h, _dp(r), _dp(u))

    def get_max_v(self):
        return self._api["get_max_v"](self._h)

    def invM(self):
        out = np.zeros((19, 19), np.float64)
        self._api["get_invM"](self._h, _dp(out))
        return out

    def M(self):
        out = np.zeros((19, 19), np.float64)
        self._api["get_M"](self._h, _dp(out))
        return out

    def n_active_blocks(self):
        return int(self._api["n_active_blocks"](self._h))
