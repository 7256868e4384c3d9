"""taichi_lbm3d_amd: MI355X-native single-phase sparse D3Q19 MRT lattice Boltzmann.

The solver class keeps the surface of the reference's
``LB3D_Solver_Single_Phase`` (Single_phase/LBM_3D_SinglePhase_Solver.py) and
runs every step in hand-written HIP kernels (liblbm3d.so, C ABI in
include/lbm3d.h).
"""
from .solver import LB3D_Solver_Single_Phase, relaxation_rates, CONVENTIONS  # noqa: F401
from . import geometry, vtk  # noqa: F401

__all__ = ["LB3D_Solver_Single_Phase", "relaxation_rates", "CONVENTIONS", "geometry", "vtk"]
