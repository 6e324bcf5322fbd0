"""Pins the oracle's sphere collider (finalize's extension for the 64^3 drag
workload: no reference counterpart) by its closed forms. A node inside the
ball is moved out along the radius to the surface, then resolvePlane's
position-level Coulomb friction (finalize.wgsl:29-42) acts in the tangent
plane there: the tangential motion since the substep began is cancelled, but
by no more than friction x depth.

Driven through the oracle's own finalize on a one-cell body whose other
corners sit far outside the ball and the tank, so only the node under test
meets the collider.
"""
import numpy as np
import pytest

from vsim import scenes

C = np.array([0.3125, -0.1875, 0.5])   # exact in f32, as the collider's uniforms are
R = 0.25
MU = 0.5


def finalize(oracle, p, prev, sphere=(*C, R), precision="f64"):
    mat = dict(youngs_modulus=1e5, poisson_ratio=0.3, density=1000.0, friction=MU)
    sc = scenes.jelly_block(1, h=1.0, pinned=False, material=mat, tank_margin=100.0)
    sc.sphere = sphere
    o = oracle.OracleScene(sc, precision)
    o.p[:] = 50.0
    o.q[:] = 50.0
    o.p[:, 0] = p
    o.q[:, 0] = prev
    o.finalize()
    assert np.all(o.p[:, 1:] == 50.0)
    return o.p[:, 0].astype(float), o.v[:, 0].astype(float)


def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def test_a_node_at_rest_inside_moves_out_along_the_radius(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = unit(rng.normal(size=3))
        p = C + n * R * rng.uniform(0.05, 0.95)
        q, _ = finalize(oracle, p, p)
        assert np.linalg.norm(q - C) == pytest.approx(R, abs=1e-12)
        assert unit(q - C) == pytest.approx(n, abs=1e-12)


def test_sliding_is_capped_at_friction_times_depth(oracle):
    # tangential motion L far above MU * depth: the tangent component loses
    # exactly MU * depth, the normal push is untouched.
    n = unit([0.2, 1.0, -0.4])
    t = unit(np.cross(n, [1.0, 0.0, 0.0]))
    depth, L = 0.03, 0.2
    p = C + n * (R - depth)
    q, _ = finalize(oracle, p, p - L * t)
    q0 = p + n * depth
    assert np.dot(q - q0, n) == pytest.approx(0, abs=1e-12)
    assert q - q0 == pytest.approx(-MU * depth * t, abs=1e-12)


def test_sticking_cancels_the_tangential_motion(oracle):
    n = unit([-0.5, 0.3, 0.8])
    t = unit(np.cross(n, [0.0, 1.0, 0.0]))
    depth, L = 0.04, 0.01                                  # L < MU * depth: sticks
    p = C + n * (R - depth)
    prev = p - L * t
    q, v = finalize(oracle, p, prev)
    moved = q - prev
    assert moved - np.dot(moved, n) * n == pytest.approx(np.zeros(3), abs=1e-12)
    assert q == pytest.approx(p + n * depth - L * t, abs=1e-12)


def test_outside_and_on_the_surface_nothing_moves(oracle):
    for r in (R * 1.5, R * 1.0000001):
        p = C + unit([1.0, 2.0, 3.0]) * r
        q, _ = finalize(oracle, p, p - 0.01)
        assert np.array_equal(q, p)


def test_a_node_on_the_centre_leaves_along_plus_y(oracle):
    q, _ = finalize(oracle, C, C)
    assert q == pytest.approx(C + [0.0, R, 0.0], abs=1e-12)


def test_no_sphere_without_a_radius(oracle):
    p = C + 0.1
    for sphere in (None, (*C, 0.0), (*C, -1.0)):
        q, _ = finalize(oracle, p, p, sphere=sphere)
        assert np.array_equal(q, p)


def test_f32_build_lands_on_the_surface(oracle):
    rng = np.random.default_rng(8)
    for _ in range(20):
        n = unit(rng.normal(size=3))
        p = C + n * R * rng.uniform(0.05, 0.95)
        q, _ = finalize(oracle, p, p, precision="f32")
        assert np.linalg.norm(q - C) == pytest.approx(R, abs=2e-7)
