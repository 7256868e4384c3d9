"""Vectorised numpy restatement of the reference single-phase solver (test infrastructure).

Follows ``Single_phase/LBM_3D_SinglePhase_Solver.py`` (SP) kernel by kernel,
in a different formulation from ``lbm_oracle.c`` (array-wide ``np.roll``
streaming and matrix products instead of per-cell loops), so that the two
restatements check each other.  Dense storage only; fields are C-order
``(nx, ny, nz[, q])`` like ``to_numpy()`` returns them.
"""
from __future__ import annotations

import numpy as np

# Direction set, SP:183-187.
E = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
              (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1),
              (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)], dtype=np.int64)
# Opposite directions, SP:85.
LR = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
# Weights, SP:195-197.
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
# Moment matrix, SP:64-82.
M = np.array([
    [1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1],
    [-1, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1],
    [1, -2, -2, -2, -2, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1],
    [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, -2, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, -2, 2, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1],
    [0, 0, 0, 0, 0, -2, 2, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1],
    [0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
    [0, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
    [0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
    [0, 0, 0, -1, -1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1]], dtype=np.float64)

GUO_SCALES = {"reference": (1 / 3, 1 / 9), "documented": (1.0, 1.0), "textbook": (3.0, 9.0)}


def relaxation(niu: float, convention: str = "reference") -> np.ndarray:
    """S_dig of init_simulation (SP:126-131); 'documented' uses tau=3*niu+0.5."""
    tau = niu / 3.0 + 0.5 if convention == "reference" else 3.0 * niu + 0.5
    s_v = 1.0 / tau
    s_o = 8.0 * (2.0 - s_v) / (8.0 - s_v)
    return np.array([0, s_v, s_v, 0, s_o, 0, s_o, 0, s_o, s_v, s_v, s_v, s_v, s_v, s_v, s_v, s_o, s_o, s_o])


def feq(rho, u):
    """feq for all 19 directions (SP:152-158); rho (...), u (...,3) -> (...,19)."""
    rho = np.asarray(rho)
    u = np.asarray(u)
    eu = u @ E.T.astype(u.dtype)
    uv = (u * u).sum(-1, keepdims=True)
    return W.astype(u.dtype) * rho[..., None] * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uv)


def meq_vec(rho, u):
    """Equilibrium moments, SP:209-215."""
    u = np.asarray(u)
    out = np.zeros(u.shape[:-1] + (19,), dtype=u.dtype)
    ux, uy, uz = u[..., 0], u[..., 1], u[..., 2]
    out[..., 0] = rho
    out[..., 1] = ux * ux + uy * uy + uz * uz
    out[..., 3], out[..., 5], out[..., 7] = ux, uy, uz
    out[..., 9] = 2 * ux * ux - uy * uy - uz * uz
    out[..., 11] = uy * uy - uz * uz
    out[..., 13], out[..., 14], out[..., 15] = ux * uy, uy * uz, ux * uz
    return out


def guo_moments(u, force, convention="reference"):
    """sum_l w_l [ a (e_l-u).F + b (e_l.u)(e_l.F) ] M[s,l]  (SP:231-236), per cell."""
    a, b = GUO_SCALES[convention]
    u = np.asarray(u)
    Fv = np.asarray(force, dtype=u.dtype)
    ef = E.astype(u.dtype)
    emu_f = (ef @ Fv)[None, :] - (u @ Fv)[..., None]  # (e_l - u).F
    eu = u @ ef.T
    e_f = ef @ Fv
    g = W * (a * emu_f + b * eu * e_f[None, :])
    return g @ M.T


def collide(F, rho, u, S, force=None, convention="reference"):
    """colission for one or many cells (SP:222-241): f = M^-1 [m - S(m-meq) + (1-S/2) G]."""
    m = F @ M.T
    m = m - S * (m - meq_vec(rho, u))
    if force is not None and np.any(np.asarray(force) != 0):
        m = m + (1.0 - 0.5 * S) * guo_moments(u, force, convention)
    return m @ np.linalg.inv(M).T


class NumpySolver:
    """Dense restatement of LB3D_Solver_Single_Phase (SP:9-481)."""

    def __init__(self, solid, niu=0.16667, force=(0.0, 0.0, 0.0), bc_mode=(0,) * 6,
                 rho_bc=(1.0,) * 6, u_bc=((0.0, 0.0, 0.0),) * 6, convention="reference"):
        self.solid = (np.asarray(solid) > 0)
        self.fluid = ~self.solid
        self.nx, self.ny, self.nz = self.solid.shape
        self.S = relaxation(niu, convention)
        self.force = np.array(force, dtype=np.float64)
        self.convention = convention
        self.bc_mode = list(bc_mode)
        self.rho_bc = list(rho_bc)
        self.u_bc = [np.array(u, dtype=np.float64) for u in u_bc]
        # init (SP:160-170), dense storage: every cell.
        shape = self.solid.shape
        self.rho = np.ones(shape)
        self.v = np.zeros(shape + (3,))
        self.F = feq(self.rho, self.v)
        self.f = self.F.copy()

    def colission(self):
        fl = self.fluid
        force = self.force if np.any(self.force != 0) else None  # force_flag, SP:137-140
        self.f[fl] = collide(self.F[fl], self.rho[fl], self.v[fl], self.S, force, self.convention)

    def streaming1(self):
        """Push with periodic wrap (SP:247-266); bounce-back into LR[s] (SP:267-268)."""
        newF = self.F.copy()
        fl = self.fluid
        for s in range(19):
            sh = tuple(int(x) for x in E[s])
            src_fluid_at_dest = np.roll(fl, sh, axis=(0, 1, 2))
            push = fl & src_fluid_at_dest
            moved = np.roll(self.f[..., s], sh, axis=(0, 1, 2))
            newF[..., s] = np.where(push, moved, newF[..., s])
            dest_solid = np.roll(self.solid, tuple(-x for x in sh), axis=(0, 1, 2))
            bb = fl & dest_solid
            newF[..., LR[s]] = np.where(bb, self.f[..., s], newF[..., LR[s]])
        self.F = newF

    def boundary_condition(self):
        """Faces in source order x-left, x-right, y-left, y-right, z-left, z-right (SP:274-370)."""
        n = (self.nx, self.ny, self.nz)
        for face in range(6):
            mode = self.bc_mode[face]
            if mode not in (1, 2):
                continue
            axis, right = divmod(face, 2)
            pos = n[axis] - 1 if right else 0
            inw = n[axis] - 2 if right else 1
            sl = [slice(None)] * 3
            sl[axis] = pos
            sl = tuple(sl)
            sq = [slice(None)] * 3
            sq[axis] = inw
            sq = tuple(sq)
            fluid_face = self.fluid[sl]
            if mode == 1:
                u = np.where(self.solid[sq][..., None], self.v[sq], self.v[sl])
                new = feq(np.full(fluid_face.shape, self.rho_bc[face]), u)
            else:
                new = feq(np.ones(fluid_face.shape), np.broadcast_to(self.u_bc[face], fluid_face.shape + (3,)))
            Fs = self.F[sl]
            Fs[fluid_face] = new[fluid_face]

    def streaming3(self):
        """Copy F->f and recompute rho, v with the half-force shift (SP:372-392)."""
        fl = self.fluid
        self.f[fl] = self.F[fl]
        rho = self.f[fl].sum(-1)
        j = self.f[fl] @ E.astype(np.float64)
        v = j / rho[:, None] + (self.force / 2)[None, :] / rho[:, None]
        self.rho[fl] = rho
        self.v[fl] = v
        self.rho[self.solid] = 1.0
        self.v[self.solid] = 0.0

    def step(self, n=1):
        for _ in range(n):
            self.colission()
            self.streaming1()
            self.boundary_condition()
            self.streaming3()

    def get_max_v(self):
        return float(np.sqrt((self.v ** 2).sum(-1)).max())
