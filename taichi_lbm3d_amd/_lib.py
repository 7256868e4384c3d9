"""ctypes binding of liblbm3d.so (the C ABI declared in include/lbm3d.h).

The library is built in-tree by :func:`build` (hipcc, gfx950).  Loading fails
loudly when it is missing: there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_HERE, "liblbm3d.so")
SOURCES = [os.path.join(_HERE, "csrc", name) for name in ("lbm3d.hip", "conv3d.hip", "geo_io.hip")]
INCLUDE = os.path.join(_ROOT, "include")

ABI_VERSION = 14
F32, F16, BF16, F64 = 0, 1, 2, 3
FLAG_STOCHASTIC_ROUNDING = 1
FLAG_TILES = 2
# storage name -> storage type code; "...sr" names are the 16-bit types with
# stochastic rounding (the flag bits of STORAGE_FLAGS set in lbm3d_params.flags)
DTYPES = {"f32": F32, "f16": F16, "bf16": BF16, "f64": F64, "f16sr": F16, "bf16sr": BF16}
STORAGE_FLAGS = {"f16sr": FLAG_STOCHASTIC_ROUNDING, "bf16sr": FLAG_STOCHASTIC_ROUNDING}


def torch_storage_dtype(name: str):
    import torch
    return {F32: torch.float32, F16: torch.float16, BF16: torch.bfloat16, F64: torch.float64}[DTYPES[name]]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep FP32 arithmetic unpacked (no v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fuse a*b+c into an FMA only inside one source expression: the per-cell code
    # then rounds identically in every kernel that inlines it (single grid, batch)
    "-ffp-contract=on",
]


def build(verbose: bool = False) -> str:
    """Compile liblbm3d.so for gfx950 with hipcc; returns its path."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-I", INCLUDE, "-o", tmp, *SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(" ".join(cmd))
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class Params(ctypes.Structure):
    """Mirror of struct lbm3d_params."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("n_cells", ctypes.c_int32), ("force_on", ctypes.c_int32),
        ("bc_mode", ctypes.c_int32 * 6),
        ("face_u_off", ctypes.c_int32 * 6),
        ("flags", ctypes.c_int32),
        ("s_v", ctypes.c_double), ("s_o", ctypes.c_double),
        ("fx", ctypes.c_double), ("fy", ctypes.c_double), ("fz", ctypes.c_double),
        ("guo_a", ctypes.c_double), ("guo_b", ctypes.c_double),
        ("rho_bc", ctypes.c_double * 6),
        ("u_bc", (ctypes.c_double * 3) * 6),
        ("band_rows", ctypes.c_int32),
    ]


class Sample(ctypes.Structure):
    """Mirror of struct lbm3d_sample."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("n_cells", ctypes.c_int32), ("cell_begin", ctypes.c_int32), ("face_u_begin", ctypes.c_int32),
        ("rows_begin", ctypes.c_int64),
    ]


class BatchSizes(ctypes.Structure):
    """Mirror of struct lbm3d_batch_sizes."""
    _fields_ = [("cells", ctypes.c_int64), ("row_words", ctypes.c_int64), ("face_u", ctypes.c_int64),
                ("desc_bytes", ctypes.c_int64)]


class Launch(ctypes.Structure):
    """Mirror of struct lbm3d_launch."""
    _fields_ = [("block", ctypes.c_int32), ("grid_per_cu", ctypes.c_int32), ("reverse", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]




BATCH_ALIGN = 512  # LBM3D_BATCH_ALIGN
TILE_CELLS = 256   # cells per tile descriptor (lbm3d_tile_words)

# name -> (restype, argtypes); every symbol include/lbm3d.h declares
_vp, _i32, _i64, _sz, _int = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t, ctypes.c_int
_pp = ctypes.POINTER(Params)
SIGNATURES = {
    "lbm3d_abi_version": (_int, []),
    "lbm3d_last_error": (ctypes.c_char_p, []),
    "lbm3d_slot_stride": (_i64, [_i32]),
    "lbm3d_row_words": (_i64, [_pp]),
    "lbm3d_tile_words": (_i64, [_pp]),
    "lbm3d_face_bit_words": (_i64, [_i32]),
    "lbm3d_face_plane": (_i64, [_pp, _i32]),
    "lbm3d_compact_workspace_bytes": (_sz, [_i64]),
    "lbm3d_compact_count": (_int, [_vp, _pp, _vp, _vp, _vp, _sz]),
    "lbm3d_compact_fill": (_int, [_vp, _pp, _vp, _vp, _vp, _vp, _sz]),
    "lbm3d_build_rows": (_int, [_vp, _pp, _vp, _vp, _vp]),
    "lbm3d_build_tiles": (_int, [_vp, _pp, _vp, _vp]),
    "lbm3d_build_rows_at": (_int, [_vp, _pp, _vp, _vp, _i32, _vp]),
    "lbm3d_build_face_bits": (_int, [_vp, _pp, _vp, _vp, _vp]),
    "lbm3d_build_neighbors": (_int, [_vp, _pp, _vp, _vp, _vp]),
    "lbm3d_init": (_int, [_vp, _pp, _int, _vp, _vp, _vp, _vp, _vp]),
    "lbm3d_step": (_int, [_vp, _pp, _int, _vp, _vp, _vp, _vp, _vp, _i64]),
    "lbm3d_run": (_int, [_vp, _pp, _int, _vp, _vp, _vp, _vp, _vp, _i64, _i32]),
    "lbm3d_macros": (_int, [_vp, _pp, _int, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "lbm3d_scatter_dense": (_int, [_vp, _pp, _vp, _vp, _vp, _i64, _i32, _i32, _i32, _vp, _vp]),
    "lbm3d_reduce_workspace_bytes": (_sz, [_i32]),
    "lbm3d_reduce": (_int, [_vp, _pp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _sz]),
    "lbm3d_round16": (_int, [_vp, _i32, _i32, _vp, _vp, _vp, _i64]),
    "lbm3d_reduce64": (_int, [_vp, _pp, _vp, _vp, _vp, _i32, _i32, _vp, _vp, _sz]),
    "lbm3d_reduce_noalloc": (_int, [_vp, _pp, _vp, _vp, _vp, _i32, _i32, _vp]),
    "lbm3d_batch_layout": (_int, [_pp, ctypes.POINTER(Sample), _i32, ctypes.POINTER(BatchSizes)]),
    "lbm3d_batch_desc": (_int, [_pp, ctypes.POINTER(Sample), _i32, _vp, _sz]),
    "lbm3d_batch_init": (_int, [_vp, _pp, _int, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "lbm3d_batch_step": (_int, [_vp, _pp, _int, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _i64,
                                ctypes.POINTER(Launch)]),
    "lbm3d_batch_run": (_int, [_vp, _pp, _int, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _i64, _i32,
                               ctypes.POINTER(Launch)]),
    "lbm3d_batch_run_range": (_int, [_vp, _pp, _int, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _i64, _i32,
                                     _i32, _i32, ctypes.POINTER(Launch)]),
    "lbm3d_batch_macros": (_int, [_vp, _pp, _int, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "lbm3d_batch_reduce_workspace_bytes": (_sz, [_i64]),
    "lbm3d_batch_reduce": (_int, [_vp, _pp, _vp, _i32, _vp, _vp, _vp, _vp, _sz]),
    # include/lbm3d_conv.h
    "lbm3d_conv_last_error": (ctypes.c_char_p, []),
    "lbm3d_conv_kpad": (_i32, [_i32, _i32]),
    "lbm3d_conv3d_fwd": (_int, [_vp, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32]),
    "lbm3d_conv3d_wflip": (_int, [_vp, _vp, _vp, _i32, _i32, _i32]),
    "lbm3d_conv3d_dgrad": (_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32]),
    "lbm3d_conv3d_wgrad_workspace_bytes": (_sz, [_i32, _i32, _i32, _i32, _i32, _i32, _i32]),
    "lbm3d_conv3d_wgrad": (_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _sz]),
    "lbm3d_distance_periodic": (_int, [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32]),
    "lbm3d_relu_bias_grad_workspace_bytes": (_sz, [_i64, _i32]),
    "lbm3d_relu_bias_grad": (_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i32, _vp, _sz]),
    # include/lbm3d_io.h
    "lbm3d_io_last_error": (ctypes.c_char_p, []),
    "lbm3d_read_geo_ascii": (_i64, [ctypes.c_char_p, _i32, _i32, _i32, _vp]),
    "lbm3d_voxelize": (_int, [_vp, _vp, _i32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              _i32, _i32, _i32, _vp]),
}


class LbmError(RuntimeError):
    pass


class Lib:
    """Loaded liblbm3d.so with checked calls."""

    def __init__(self, path: str = LIB_PATH):
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build it with taichi_lbm3d_amd._lib.build() "
                "(hipcc --offload-arch=gfx950); the solver has no CPU fallback")
        self.path = path
        self.cdll = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(self.cdll, name)
            fn.restype = res
            fn.argtypes = args
        v = self.cdll.lbm3d_abi_version()
        if v != ABI_VERSION:
            raise ImportError(f"liblbm3d.so ABI {v}, expected {ABI_VERSION}: rebuild it")

    def last_error(self) -> str:
        msg = self.cdll.lbm3d_last_error()
        return msg.decode() if msg else ""

    def call(self, name: str, *args):
        rc = getattr(self.cdll, name)(*args)
        if rc != 0:
            conv = name.startswith("lbm3d_conv") or name == "lbm3d_distance_periodic"
            msg = self.cdll.lbm3d_conv_last_error() if conv else None
            detail = msg.decode() if msg else self.last_error()
            raise LbmError(f"{name} failed ({rc}): {detail}")
        return rc

    def size(self, name: str, *args) -> int:
        return int(getattr(self.cdll, name)(*args))


_LIB: Lib | None = None


def lib() -> Lib:
    global _LIB
    if _LIB is None:
        _LIB = Lib()
    return _LIB
