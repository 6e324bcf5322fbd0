"""Pins the oracle's splat deformation (sample.wgsl:40-80 + covariance.wgsl:
30-42) the way the reference's sample.test.js and covariance.test.js do: the
trilinear field reproduces an affine map exactly (position and gradient), and
the deformed covariance equals F R diag(s)^2 R^T F^T computed independently
(scipy's rotation of an xyzw quaternion), with the diagonal known answer."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle.oracle import OracleScene
from vsim import scenes


def affine_scene(A, b, splats, side=3, h=0.0625, seed=0):
    # h and the origin are exact in f32, so the rest lattice is exactly regular.
    sc = scenes.jelly_block(side, h=h, splats=splats, seed=seed, pinned=False, origin=(0.125, -0.25, 0.375))
    o = OracleScene(sc, "f64")
    rest = o.p.copy()
    o.p[:] = (A @ rest) + np.asarray(b, float)[:, None]
    return sc, o


def test_affine_map_is_reproduced_in_position_and_gradient():
    rng = np.random.default_rng(1)
    A = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
    b = np.array([0.01, -0.02, 0.03])
    sc, o = affine_scene(A, b, 500)
    body = sc.bodies[0]
    q = body.splat_quat.astype(float)
    q /= np.linalg.norm(q, axis=1, keepdims=True)   # unit in f64, as scipy assumes
    o.sq[:] = q.reshape(-1)
    mean, cov = o.deform_splats()
    want = A @ body.splat_mean.astype(float).T + b[:, None]
    assert np.max(np.abs(mean - want)) < 1e-12
    # gradient: the covariance of a unit isotropic splat is F F^T s^2
    R = Rotation.from_quat(q).as_matrix()
    S = body.splat_scale.astype(float)
    ref = np.einsum("ij,njk,nk,nlk,ml->nim", A, R, S ** 2, R, A)
    got = np.stack([cov[0], cov[1], cov[2], cov[1], cov[3], cov[4], cov[2], cov[4], cov[5]], axis=1).reshape(-1, 3, 3)
    rel = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert rel < 1e-9


def test_translation_moves_every_splat_by_the_same_amount():
    sc, o = affine_scene(np.eye(3), [0.5, 0.25, -0.125], 300)
    mean, _ = o.deform_splats()
    d = mean - sc.bodies[0].splat_mean.astype(float).T
    assert np.max(np.abs(d - np.array([[0.5], [0.25], [-0.125]]))) < 1e-12


def test_diagonal_known_answer():
    # covariance.test.js:118-130: F = diag(2, 3, 0.5), no rotation, s = (0.01,
    # 0.02, 0.04) -> diag(0.02^2, 0.06^2, 0.02^2).
    sc, o = affine_scene(np.diag([2.0, 3.0, 0.5]), [0, 0, 0], 4)
    o.sq[:] = np.tile([0.0, 0.0, 0.0, 1.0], o.G)
    o.ss[0][:] = 0.01
    o.ss[1][:] = 0.02
    o.ss[2][:] = 0.04
    _, cov = o.deform_splats()
    want = [0.02 ** 2, 0, 0, 0.06 ** 2, 0, 0.02 ** 2]
    for k in range(6):
        assert np.allclose(cov[k], want[k], rtol=1e-12, atol=1e-18)


def test_outside_points_extrapolate_from_the_outermost_cell():
    # A point past the lattice keeps the affine map (extrapolation, not a clamp
    # to the boundary): sample.js:42-59.
    A = np.diag([1.5, 1.0, 0.8])
    sc, o = affine_scene(A, [0, 0, 0], 2)
    o.sm[:, 0] = [0.125 - 0.02, -0.25 + 0.25, 0.375 + 0.1]    # below x min, above y max
    mean, _ = o.deform_splats()
    assert np.max(np.abs(mean[:, 0] - A @ o.sm[:, 0])) < 1e-12


def test_f32_covariance_is_symmetric_and_positive():
    sc = scenes.c1_scene(splats=2000)
    o = OracleScene(sc, "f32")
    o.frame(10, splats=True)
    cov = o.out_cov.astype(np.float64)
    M = np.stack([cov[0], cov[1], cov[2], cov[1], cov[3], cov[4], cov[2], cov[4], cov[5]], axis=1).reshape(-1, 3, 3)
    assert np.all(np.linalg.eigvalsh(M) > -1e-12)
