"""Pins of the oracle's dynamics against closed forms and invariants.

Each test names the reference lines it pins and the mistake it would catch.
All runs use the FP64 C oracle unless stated; the FP32 build and the numpy
restatement are cross-checked at the end.
"""
import numpy as np
import pytest

from oracle import COracle
from oracle import oracle_np as onp

GUO = {"reference": (1 / 3, 1 / 9), "documented": (1.0, 1.0), "textbook": (3.0, 9.0)}
FACES = ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")


def nu_eff(niu, conv):
    tau = niu / 3 + 0.5 if conv == "reference" else 3 * niu + 0.5
    return (tau - 0.5) / 3


def oracle(shape, conv="reference", real="f64", solid=None, niu=0.1, force=(0, 0, 0), sparse=False):
    o = COracle(*shape, sparse_storage=sparse, real=real, convention=conv)
    if solid is not None:
        o.set_solid(solid)
    o.niu = niu
    o.fx, o.fy, o.fz = force
    return o


@pytest.mark.parametrize("conv", ["reference", "documented", "textbook"])
def test_single_cell_collision_moments(conv):
    """A 1x1x1 periodic box streams every population onto itself, so one step is
    colission (SP:222-241) followed by streaming3 (SP:372-392).  The moments of
    the result obey the MRT update with the Guo term in closed form:
    m0 += -(3a-b)/3 u.F, j += a F/3, and each non-conserved moment relaxes with
    its S towards meq (SP:131, SP:209-215)."""
    a, b = GUO[conv]
    rng = np.random.default_rng(5)
    for trial in range(6):
        F = 1e-3 * rng.standard_normal(3)
        o = oracle((1, 1, 1), conv, force=tuple(F), niu=0.05 + 0.1 * trial)
        o.init_simulation()
        rho0 = 1 + 0.01 * rng.standard_normal()
        u = 0.05 * rng.standard_normal(3)
        Fin = onp.feq(np.array(rho0), u) + 1e-3 * rng.standard_normal(19)
        o.set_F(Fin.reshape(1, 1, 1, 19))
        o.set_macros(np.full((1, 1, 1), rho0), u.reshape(1, 1, 1, 3))
        o.step(1)
        Fout = o.F().reshape(19)
        m_in, m_out = onp.M @ Fin, onp.M @ Fout
        S = onp.relaxation(o.niu, conv)
        uF = u @ F
        G = np.zeros(19)
        G[0] = -(3 * a - b) / 3 * uF
        G[1] = 2 * b / 9 * uF
        G[3], G[5], G[7] = a * F / 3
        G[9] = 2 * b / 9 * (2 * F[0] * u[0] - F[1] * u[1] - F[2] * u[2])
        G[11] = 2 * b / 9 * (F[1] * u[1] - F[2] * u[2])
        G[13] = b / 9 * (F[0] * u[1] + F[1] * u[0])
        G[14] = b / 9 * (F[1] * u[2] + F[2] * u[1])
        G[15] = b / 9 * (F[0] * u[2] + F[2] * u[0])
        meq = onp.meq_vec(rho0, u[None])[0]
        want = m_in - S * (m_in - meq) + (1 - S / 2) * G
        np.testing.assert_allclose(m_out, want, atol=1e-14)
        # conserved up to the force term
        assert abs(m_out[0] - m_in[0] - G[0]) < 1e-15
        np.testing.assert_allclose(m_out[[3, 5, 7]] - m_in[[3, 5, 7]], a * F / 3, atol=1e-16)
        # streaming3's macros from the new populations, v = j/rho + F/(2 rho) (SP:380-388)
        r = o.rho().reshape(())
        np.testing.assert_allclose(o.v().reshape(3), (onp.E.T @ Fout) / r + F / 2 / r, atol=1e-16)


def test_rest_state_is_fixed_point():
    """f = w, rho = 1, v = 0 with no force is stationary through collision,
    streaming with bounce-back and pressure faces at rho 1 (SP:160-170)."""
    rng = np.random.default_rng(2)
    solid = (rng.random((9, 8, 7)) < 0.3).astype(np.int8)
    o = oracle((9, 8, 7), solid=solid)
    o.set_bc_rho("x_left", 1.0)
    o.set_bc_rho("x_right", 1.0)
    o.init_simulation()
    o.step(30)
    fl = solid == 0
    np.testing.assert_allclose(o.F()[fl], np.broadcast_to(onp.W, (fl.sum(), 19)), atol=1e-15)
    assert np.abs(o.v()).max() < 1e-15 and np.abs(o.rho()[fl] - 1).max() < 1e-15


def _perturbed(o, shape, solid, rng, amp=1e-3):
    rho = 1 + amp * rng.standard_normal(shape)
    u = amp * rng.standard_normal(shape + (3,))
    F = onp.feq(rho, u) + amp * 0.1 * rng.standard_normal(shape + (19,))
    F[solid != 0] = 0
    rho = F.sum(-1)
    u = (F @ onp.E.astype(float)) / np.where(rho == 0, 1, rho)[..., None]
    rho[solid != 0] = 1
    u[solid != 0] = 0
    o.set_F(F)
    o.set_macros(rho, u)
    return F


def test_mass_conserved_periodic_with_bounce_back():
    """No force, all faces periodic: collision conserves m0 and streaming1
    (push + halfway bounce-back, SP:259-269) moves every population exactly
    once, so total mass is invariant."""
    rng = np.random.default_rng(3)
    shape = (10, 9, 8)
    solid = (rng.random(shape) < 0.3).astype(np.int8)
    o = oracle(shape, solid=solid)
    o.init_simulation()
    F = _perturbed(o, shape, solid, rng)
    fl = solid == 0
    m0 = F[fl].sum()
    o.step(50)
    assert abs(o.F()[fl].sum() - m0) < 1e-12 * m0
    assert abs(o.rho()[fl].sum() - m0) < 1e-12 * m0


@pytest.mark.parametrize("conv", ["reference", "documented", "textbook"])
def test_body_force_injection_closed_form(conv):
    """All-fluid periodic box: each step adds exactly a F/3 of momentum per cell
    and -(3a-b)/3 u.F of mass per cell (SP:230-238; a, b = 1/3, 1/9 in SP)."""
    a, b = GUO[conv]
    rng = np.random.default_rng(4)
    shape = (6, 5, 4)
    Fv = np.array([2e-4, -1e-4, 3e-5])
    o = oracle(shape, conv, force=tuple(Fv))
    o.init_simulation()
    _perturbed(o, shape, np.zeros(shape, np.int8), rng)
    for _ in range(5):
        F0, v0 = o.F(), o.v()
        o.step(1)
        F1 = o.F()
        dj = (F1.reshape(-1, 19) @ onp.E.astype(float)).sum(0) - (F0.reshape(-1, 19) @ onp.E.astype(float)).sum(0)
        np.testing.assert_allclose(dj, a * Fv / 3 * np.prod(shape), rtol=1e-9)
        dm = F1.sum() - F0.sum()
        want = -(3 * a - b) / 3 * (v0.reshape(-1, 3) @ Fv).sum()
        assert abs(dm - want) < 1e-13 + 1e-9 * abs(want)  # sums of ~2000 O(1) terms


@pytest.mark.parametrize("conv,niu", [("reference", 0.1), ("reference", 0.5), ("documented", 0.1),
                                      ("documented", 0.5), ("textbook", 0.1)])
def test_shear_wave_viscosity(conv, niu):
    """A decaying shear wave u_x = U sin(2 pi z / N) decays at nu k^2 with
    nu = (tau - 1/2)/3: tau = niu/3 + 1/2 in SP (SP:127), so nu = niu/9;
    tau = 3 niu + 1/2 in the documented form.  Within 1 %."""
    N, U = 64, 1e-3
    o = oracle((1, 1, N), conv, niu=niu)
    o.init_simulation()
    z = np.arange(N)
    u = np.zeros((1, 1, N, 3))
    u[..., 0] = U * np.sin(2 * np.pi * z / N)
    rho = np.ones((1, 1, N))
    o.set_F(onp.feq(rho, u))
    o.set_macros(rho, u)
    s = np.sin(2 * np.pi * z / N)
    o.step(100)
    a1 = o.v()[..., 0].reshape(-1) @ s
    o.step(1000)
    a2 = o.v()[..., 0].reshape(-1) @ s
    nu = np.log(a1 / a2) / ((2 * np.pi / N) ** 2 * 1000)
    assert abs(nu / nu_eff(niu, conv) - 1) < 0.01


@pytest.mark.parametrize("conv", ["reference", "documented", "textbook"])
def test_poiseuille_slit(conv):
    """Slit between solid planes z = 0 and z = NZ-1 (example_poiseuille_flow.py:20-31),
    body force along y.  Halfway bounce-back puts the walls at z = 1/2 and
    NZ - 3/2, and with (1/s_v - 1/2)(1/s_o - 1/2) = 3/16 the momentum profile is
    the exact parabola  j + F_eff/2 = F_eff/(2 nu) (z - 1/2)(NZ - 3/2 - z),
    F_eff = a F/3.  Exact to round-off for the mass-conserving textbook scaling;
    SP's scaling loses mass at -(8/27) u.F per step, which bounds the rest."""
    NZ, niu, Fy, steps = 16, 3.0, 1e-5, 20000
    solid = np.zeros((1, 1, NZ), np.int8)
    solid[..., 0] = solid[..., -1] = 1
    o = oracle((1, 1, NZ), conv, solid=solid, niu=niu, force=(0, Fy, 0))
    o.init_simulation()
    o.step(steps)
    a, b = GUO[conv]
    f_eff, nu = a * Fy / 3, nu_eff(niu, conv)
    z = np.arange(NZ)
    ua = f_eff / (2 * nu) * (z - 0.5) * (NZ - 1.5 - z)
    rho, v = o.rho()[0, 0], o.v()[0, 0, :, 1]
    j = rho * v - Fy / 2
    fl = slice(1, NZ - 1)
    err = np.abs(j[fl] + f_eff / 2 - ua[fl]).max() / ua.max()
    drift = abs(1 - rho[fl].mean())
    if conv == "textbook":
        assert err < 1e-11 and drift < 1e-12
    else:
        pred = (3 * a - b) / 3 * Fy * ua[fl].mean() * steps
        assert abs(drift / pred - 1) < 0.2  # the mass loss the force term predicts
        assert err < 1e-8


def test_pressure_face_invariants():
    """No force: a pressure face node gets feq(rho_bc, v_prev) (SP:274-281), and
    v_prev starts at 0 and stays exactly 0 (the feq(rho,0) populations cancel in
    pairs), so rho = rho_bc and v = 0 on every fluid face node."""
    rng = np.random.default_rng(6)
    shape = (10, 8, 7)
    solid = (rng.random(shape) < 0.25).astype(np.int8)
    o = oracle(shape, solid=solid)
    o.set_bc_rho("x_left", 1.01)
    o.set_bc_rho("x_right", 0.99)
    o.init_simulation()
    o.step(60)
    rho, v = o.rho(), o.v()
    for i, r in ((0, 1.01), (shape[0] - 1, 0.99)):
        fl = solid[i] == 0
        assert np.all(v[i][fl] == 0.0)
        assert np.abs(rho[i][fl] - r).max() < 1e-15
    assert np.abs(v[1:-1][solid[1:-1] == 0]).max() > 1e-6  # the interior does flow


def test_velocity_face_and_edge_precedence():
    """All six faces in velocity mode with distinct velocities: each face node
    carries feq(1, u_bc) of the last face that contains it (z-right > z-left >
    y-right > y-left > x-right > x-left, SP:283-370), so rho = 1 and
    v = u_bc + F/2 there."""
    shape = (6, 5, 4)
    Fv = np.array([1e-4, 0, -2e-4])
    o = oracle(shape, force=tuple(Fv))
    vel = [np.array([0.01 * (f + 1), -0.005 * f, 0.002 * f]) for f in range(6)]
    for f in range(6):
        o.set_bc_vel(FACES[f], vel[f])
    o.init_simulation()
    o.step(3)
    rho, v = o.rho(), o.v()
    n = shape
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                owner = -1
                for f, pos in enumerate((i, i, j, j, k, k)):
                    if pos == (n[f // 2] - 1 if f % 2 else 0):
                        owner = f
                if owner < 0:
                    continue
                assert abs(rho[i, j, k] - 1) < 1e-15
                np.testing.assert_allclose(v[i, j, k], vel[owner] + Fv / 2, atol=1e-16)


def test_sparse_equals_dense_on_fluid():
    """sparse_storage only changes the SNode layout (SP:36-44): fluid values are
    identical; the padded extent is 3*(n//3+1) and cells outside active blocks
    read 0."""
    rng = np.random.default_rng(8)
    shape = (10, 8, 7)
    solid = np.ones(shape, np.int8)
    solid[1:6, 1:7, 2:6] = (rng.random((5, 6, 4)) < 0.3)
    kw = dict(solid=solid, force=(1e-4, 2e-5, 0))
    d = oracle(shape, **kw)
    s = oracle(shape, sparse=True, **kw)
    for o in (d, s):
        o.set_bc_rho("y_left", 1.002)
        o.init_simulation()
        o.step(20)
    fl = solid == 0
    assert s.shape() == (12, 9, 9)
    np.testing.assert_array_equal(s.rho()[:10, :8, :7][fl], d.rho()[fl])
    np.testing.assert_array_equal(s.v()[:10, :8, :7][fl], d.v()[fl])
    act = np.zeros((4, 3, 3), bool)
    pad = np.zeros(s.shape(), bool)
    pad[:10, :8, :7] = fl
    act = pad.reshape(4, 3, 3, 3, 3, 3).any(axis=(1, 3, 5))
    expect = np.repeat(np.repeat(np.repeat(act, 3, 0), 3, 1), 3, 2).astype(float)
    np.testing.assert_array_equal(s.rho()[~pad], expect[~pad])
    assert s.n_active_blocks() == int(act.sum())


@pytest.mark.parametrize("conv", ["reference", "documented", "textbook"])
def test_c_and_numpy_restatements_agree(conv):
    """Two independent restatements (per-cell C loops, array-wide numpy) of
    SP:222-392 agree in FP64 on mixed boundaries and a body force."""
    rng = np.random.default_rng(9)
    shape = (9, 8, 7)
    solid = (rng.random(shape) < 0.25).astype(np.int8)
    modes = [1, 1, 2, 0, 2, 2]
    rho_bc = [1.01, 0.99, 1, 1, 1, 1]
    u_bc = [[0, 0, 0], [0, 0, 0], [0.01, 0.02, 0], [0, 0, 0], [0, 0.01, 0], [0.03, 0, 0.01]]
    force = (1e-3, 2e-4, -3e-4)
    o = oracle(shape, conv, solid=solid, force=force)
    for f in range(6):
        if modes[f] == 1:
            o.set_bc_rho(FACES[f], rho_bc[f])
        elif modes[f] == 2:
            o.set_bc_vel(FACES[f], u_bc[f])
    o.init_simulation()
    o.step(40)
    n = onp.NumpySolver(solid, niu=0.1, force=force, bc_mode=modes, rho_bc=rho_bc, u_bc=u_bc, convention=conv)
    n.step(40)
    assert np.abs(o.rho() - n.rho).max() < 1e-13
    assert np.abs(o.v() - n.v).max() < 1e-13
    assert abs(o.get_max_v() - n.get_max_v()) < 1e-13


def test_fp32_build_tracks_fp64():
    rng = np.random.default_rng(10)
    shape = (8, 8, 8)
    solid = (rng.random(shape) < 0.2).astype(np.int8)
    runs = {}
    for real in ("f32", "f64"):
        o = oracle(shape, real=real, solid=solid, force=(1e-4, 0, 0))
        o.set_bc_vel("z_right", [0.05, 0, 0])
        o.init_simulation()
        o.step(100)
        runs[real] = (o.rho(), o.v())
    fl = solid == 0
    dv = np.linalg.norm(runs["f32"][1][fl] - runs["f64"][1][fl]) / np.linalg.norm(runs["f64"][1][fl])
    assert 0 < dv < 1e-4
