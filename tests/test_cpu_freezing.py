"""When parameters take effect, as Taichi freezes them (CPU).

The reference reads niu, the face velocities and force_flag in
init_simulation() (SP:118-140), while its kernels capture fx/fy/fz
(cal_local_force, SP:217-220), force_flag (SP:230) and the face modes and
densities (ti.static in Boundary_condition, SP:274-365) when Taichi compiles
them, at the first step().  Nothing changes after that.  These tests check the
oracle's two-stage freezing against runs where the same values were set at the
stage that counts, and the GPU solver's host-side freezing of its parameter
block (no GPU needed).
"""
import numpy as np
import pytest

from oracle import COracle

SHAPE = (8, 7, 6)


def solid():
    return (np.random.default_rng(4).random(SHAPE) < 0.25).astype(np.int8)


def oracle(niu=0.1, fx=0.0, rho=None, vel=None):
    o = COracle(*SHAPE, real="f64")
    o.set_solid(solid())
    o.niu, o.fx = niu, fx
    for face, r in (rho or {}).items():
        o.set_bc_rho(face, r)
    for face, u in (vel or {}).items():
        o.set_bc_vel(face, u)
    return o


def fields(o):
    return o.rho(), o.v()


def test_niu_after_init_simulation_is_ignored():
    a, b = oracle(), oracle()
    a.init_simulation()
    b.init_simulation()
    b.niu = 0.4
    a.step(9)
    b.step(9)
    for x, y in zip(fields(a), fields(b)):
        np.testing.assert_array_equal(x, y)
    c = oracle(niu=0.4)
    c.init_simulation()
    c.step(9)
    assert not np.array_equal(a.rho(), c.rho())


def test_force_after_init_simulation_only_shifts_reported_velocity():
    """force_flag stays 0, so the collision adds no force term; streaming3
    still reports v = j/rho + (f/2)/rho with the compiled-in fx.  After one
    step (which collides with init's v = 0) the populations and rho are those
    of the unforced run and v is shifted by exactly fx/(2 rho); later steps
    collide with the shifted v."""
    a, b = oracle(rho={"x_left": 1.003}), oracle(rho={"x_left": 1.003})
    a.init_simulation()
    b.init_simulation()
    b.fx = 2e-4
    a.step(1)
    b.step(1)
    np.testing.assert_array_equal(a.F(), b.F())
    fl = solid() == 0
    np.testing.assert_array_equal(a.rho(), b.rho())
    dv = b.v()[..., 0][fl] - a.v()[..., 0][fl]
    np.testing.assert_allclose(dv, 1e-4 / a.rho()[fl], rtol=1e-9)
    np.testing.assert_array_equal(a.v()[..., 1:], b.v()[..., 1:])
    c = oracle(fx=2e-4, rho={"x_left": 1.003})
    c.init_simulation()
    c.step(1)
    assert not np.array_equal(c.F(), b.F())  # a force known at init_simulation adds the force term


def test_face_mode_and_density_before_first_step_count():
    a = oracle(rho={"x_left": 1.002, "x_right": 0.997})
    a.init_simulation()
    b = oracle(rho={"x_left": 1.004})
    b.init_simulation()
    b.set_bc_rho("x_left", 1.002)
    b.set_bc_rho("x_right", 0.997)
    a.step(8)
    b.step(8)
    for x, y in zip(fields(a), fields(b)):
        np.testing.assert_array_equal(x, y)


def test_face_velocity_after_init_simulation_is_ignored():
    a = oracle(vel={"z_right": [0.02, 0.0, 0.0]})
    a.init_simulation()
    b = oracle(vel={"z_right": [0.02, 0.0, 0.0]})
    b.init_simulation()
    b.u_bc[5] = [0.09, 0.0, 0.0]
    a.step(8)
    b.step(8)
    for x, y in zip(fields(a), fields(b)):
        np.testing.assert_array_equal(x, y)


def test_changes_after_first_step_are_ignored():
    a = oracle(fx=1e-4, rho={"x_left": 1.002})
    b = oracle(fx=1e-4, rho={"x_left": 1.002})
    a.init_simulation()
    b.init_simulation()
    a.step(1)
    b.step(1)
    b.niu, b.fx = 0.3, 0.0
    b.set_bc_rho("y_left", 1.01)
    b.set_bc_rho("x_left", 1.0)
    a.step(7)
    b.step(7)
    for x, y in zip(fields(a), fields(b)):
        np.testing.assert_array_equal(x, y)


def test_planted_late_freeze_is_caught():
    """An oracle that re-read the face densities at every step would differ:
    the comparison above is sensitive to the change it claims is ignored."""
    a = oracle(rho={"x_left": 1.002})
    a.init_simulation()
    a.step(1)
    a.set_bc_rho("x_left", 1.0)
    a._compiled = False  # planted: freeze again on the next step
    b = oracle(rho={"x_left": 1.002})
    b.init_simulation()
    b.step(1)
    a.step(7)
    b.step(7)
    assert not np.array_equal(a.rho(), b.rho())


def test_solver_parameter_block_freezes_in_two_stages():
    """The GPU solver's parameter block, filled on the host: init_simulation's
    stage fixes the rates, force_flag and face velocities; the first step's
    stage takes the force, face modes and densities."""
    from taichi_lbm3d_amd import LB3D_Solver_Single_Phase
    from taichi_lbm3d_amd.solver import relaxation_rates
    s = LB3D_Solver_Single_Phase(*SHAPE)
    s.niu = 0.1
    s.set_bc_vel_z1([0.02, 0.0, 0.0])
    s.set_bc_rho_x0(1.004)
    s._freeze_init()
    p = s._params
    _, sv, so = relaxation_rates(0.1)
    assert p.s_v == sv and p.s_o == so and p.force_on == 0
    assert p.bc_mode[0] == 1 and p.rho_bc[0] == pytest.approx(1.004) and p.u_bc[5][0] == pytest.approx(0.02)
    s.niu = 0.4
    s.fx = 2e-4
    s.vx_bczr = 0.09
    s.set_bc_rho_x1(0.997)
    s.set_bc_rho_x0(1.002)
    s._freeze_compile()
    assert p.s_v == sv and p.force_on == 0 and p.fx == pytest.approx(2e-4)
    assert p.u_bc[5][0] == pytest.approx(0.02)
    assert p.bc_mode[1] == 1 and p.rho_bc[1] == pytest.approx(0.997) and p.rho_bc[0] == pytest.approx(1.002)
    assert p.face_u_off[0] == 0 and p.face_u_off[1] == 3 * SHAPE[1] * SHAPE[2]
