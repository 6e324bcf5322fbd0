"""Pins the node passes of the oracle: predict (predict.wgsl:19-36), the hard
and compliant drives (driveHard.wgsl:15-35, driveCompliant.wgsl:24-41) and
finalize (finalize.wgsl:29-74) against closed forms and the known answers of
the reference's kinematicDrive.test.js and native-compute/page.js."""
import math

import numpy as np
import pytest

from _raw import Raw

DT = 1 / 1200


def test_one_free_fall_step_drops_g_dt2():
    # native-compute/page.js:38: a free node falls 9.81 dt^2 in one substep.
    dt = 1 / 600
    r = Raw([[0.0, 1.0, 0.0]], precision="f32", dt=dt)
    r.sc.gravity[:] = (0.0, -9.81, 0.0)
    r.call("o_predict")
    assert abs((1.0 - r.pos[0, 1]) - 9.81 * dt * dt) < 2e-7


def test_predict_skips_inert_nodes():
    r = Raw([[0.0, 1.0, 0.0]], w=[0.0], vel=[[1.0, 2.0, 3.0]])
    r.sc.gravity[:] = (0.0, -9.81, 0.0)
    r.call("o_predict")
    assert np.array_equal(r.pos, [[0.0, 1.0, 0.0]]) and np.array_equal(r.v.T, [[1.0, 2.0, 3.0]])


def hard(here, goal, vel, dt=DT):
    r = Raw([here], w=[0.0], dt=dt).drive([0], [[0, 0, 0]], goal, vel, law=0)
    r.call("o_drive")
    return r.pos[0], r.v.T[0], r.q.T[0]


def test_hard_drive_advances_by_target_velocity():
    pos, _, prev = hard([0, 0, 0], [10, 0, 0], [1, 0, 0])
    assert np.array_equal(pos, [DT, 0, 0]) and np.array_equal(prev, [0, 0, 0])


def test_hard_drive_lands_exactly_on_goal():
    goal = np.array([0.001, -0.0002, 0.0])
    pos, _, _ = hard([0, 0, 0], goal, goal * 1e6)
    assert np.array_equal(pos, goal)


def test_hard_drive_reaches_goal_in_the_implied_substeps():
    goal = np.array([0.012, -0.02, 0.005])
    vel = goal / (20 * DT)
    r = Raw([[0, 0, 0]], w=[0.0]).drive([0], [[0, 0, 0]], goal, vel, law=0)
    for _ in range(20):
        r.call("o_drive")
    assert np.all(np.abs(r.pos[0] - goal) < 1e-12)


def test_hard_drive_velocity_and_rest():
    _, v, _ = hard([0, 0, 0], [10, 10, 10], [0.3, -0.1, 0.2])
    assert np.allclose(v, [0.3, -0.1, 0.2], atol=1e-9)
    pos, v, _ = hard([1, 2, 3], [1, 2, 3], [0, 0, 0])
    assert np.array_equal(pos, [1, 2, 3]) and np.array_equal(v, [0, 0, 0])


def test_hard_drive_keeps_a_held_set_rigid():
    offa, offb = np.array([0.01, 0, 0]), np.array([-0.02, 0.03, 0])
    t1, t2 = np.zeros(3), np.array([0.05, -0.01, 0.02])
    r = Raw([t1 + offa, t1 + offb], w=[0.0, 0.0]).drive([0, 1], [offa, offb], t2, (t2 - t1) / DT)
    r.call("o_drive")
    assert np.allclose(r.pos[0] - r.pos[1], offa - offb, atol=1e-12)


def test_compliant_drive_closed_form_and_limits():
    tau = 0.05
    goal = np.array([1.0, -0.5, 0.25])
    r = Raw([[0, 0, 0]]).drive([0], [[0, 0, 0]], goal, [0, 0, 0], law=1, response_time=tau)
    r.call("o_drive")
    frac = DT * DT / (DT * DT + tau * tau)
    assert np.allclose(r.pos[0], goal * frac, rtol=1e-12)
    r0 = Raw([[0, 0, 0]]).drive([0], [[0, 0, 0]], [1, 2, 3], [0, 0, 0], law=1, response_time=0.0)
    r0.call("o_drive")
    assert np.array_equal(r0.pos[0], [1, 2, 3])
    rinf = Raw([[1, 1, 1]]).drive([0], [[0, 0, 0]], [5, 5, 5], [0, 0, 0], law=1, response_time=1e6)
    rinf.call("o_drive")
    assert np.allclose(rinf.pos[0], [1, 1, 1], atol=1e-6)


def test_floor_plane_projects_and_applies_coulomb_friction():
    # A node 3 mm under the floor, sliding 1 mm sideways since the substep
    # began: depth 3 mm, friction 0.1 -> at most 0.3 mm of the slide is taken.
    dt = 1 / 600
    r = Raw([[0.001, -0.003, 0.0]], prev=[[0.0, 0.0, 0.0]], bodies=((0.1, 0.0),), dt=dt)
    r.sc.tank_min[:] = (-1.0, 0.0, -1.0)
    r.sc.tank_max[:] = (1.0, 1.0, 1.0)
    r.call("o_finalize")
    x, y, z = r.pos[0]
    assert y == 0.0
    assert x == pytest.approx(0.001 - 0.1 * 0.003, rel=1e-12)
    # Sticking: friction 1 cancels the whole 1 mm slide.
    r2 = Raw([[0.001, -0.003, 0.0]], prev=[[0.0, 0.0, 0.0]], bodies=((1.0, 0.0),), dt=dt)
    r2.sc.tank_min[:] = (-1.0, 0.0, -1.0)
    r2.sc.tank_max[:] = (1.0, 1.0, 1.0)
    r2.call("o_finalize")
    assert r2.pos[0, 0] == pytest.approx(0.0, abs=1e-15)


def test_tank_is_open_at_the_top_and_walls_push_in():
    r = Raw([[0.0, 5.0, 0.0], [1.2, 0.5, -1.3]], prev=[[0.0, 5.0, 0.0], [1.2, 0.5, -1.3]], dt=1 / 600)
    r.sc.tank_min[:] = (-1.0, 0.0, -1.0)
    r.sc.tank_max[:] = (1.0, 1.0, 1.0)
    r.call("o_finalize")
    assert r.pos[0, 1] == 5.0
    assert r.pos[1, 0] == 1.0 and r.pos[1, 2] == -1.0


def test_finalize_velocity_is_damped_difference():
    dt = 1 / 600
    r = Raw([[0.01, 0.5, 0.0]], prev=[[0.0, 0.5, 0.0]], bodies=((0.0, 0.6),), dt=dt)
    r.sc.tank_min[:] = (-1.0, 0.0, -1.0)
    r.sc.tank_max[:] = (1.0, 1.0, 1.0)
    r.call("o_finalize")
    assert r.v.T[0, 0] == pytest.approx(0.01 / dt * math.exp(-0.6 * dt), rel=1e-12)
