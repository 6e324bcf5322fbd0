"""The oracle's covariance -> engine-axes decomposition (four-sweep cyclic
Jacobi, apps/simulator/src/render/ellipsoid.js:40-149 and ellipsoid.wgsl:32-91),
pinned by the properties apps/simulator/test/unit/ellipsoid.test.js:116-191
asserts, plus numpy's symmetric eigensolver on the same matrices.

Every check rebuilds the covariance the answer describes, R diag(s)^2 R^T, and
compares it with the input: the rotation and the axis order are not unique,
the ellipsoid is.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def stream(seed):
    """The reference test's linear congruential stream, in the same f64
    arithmetic JS uses (ellipsoid.test.js:87-90)."""
    state = float(seed)

    def nxt():
        nonlocal state
        state = math.fmod(state * 1103515245.0 + 12345.0, 2147483648.0)
        return state / 2147483648.0
    return nxt


def some_rotation(nxt):
    q = np.array([nxt() * 2 - 1 for _ in range(4)])
    return q / np.linalg.norm(q)


def rotation_matrix(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def covariance(q, s):
    r = rotation_matrix(np.asarray(q, np.float64))
    return r @ np.diag(np.asarray(s, np.float64) ** 2) @ r.T


def carried(gradient, q, s):
    return gradient @ covariance(q, s) @ gradient.T


def six(c):
    return np.array([c[0, 0], c[0, 1], c[0, 2], c[1, 1], c[1, 2], c[2, 2]])


def rebuilt_error(c, precision="f64"):
    q, s = O.ellipsoid_axes(six(c), precision)
    back = covariance(np.asarray(q, np.float64), np.asarray(s, np.float64))
    return np.abs(back - c).max() / np.abs(c).max(), q, s


def test_identity_is_a_no_op():
    rest = some_rotation(stream(11))
    c = covariance(rest, [0.02, 0.006, 0.0004])
    err, _, _ = rebuilt_error(c)
    assert err < 1e-9


def test_diagonal_covariance_reads_straight_off():
    q, s = O.ellipsoid_axes([0.0004, 0, 0, 0.0036, 0, 0.0004])
    assert np.allclose(sorted(s), [0.02, 0.02, 0.06], atol=1e-9, rtol=0)
    assert np.array_equal(q, [0, 0, 0, 1])  # nothing off the diagonal: no rotation at all


def test_describes_the_same_ellipsoid_2000_trials():
    nxt = stream(20260810)
    worst = 0.0
    for trial in range(2000):
        amp = [0.05, 0.4, 1.5][trial % 3]
        g = np.eye(3) + np.array([nxt() * 2 - 1 for _ in range(9)]).reshape(3, 3).T * amp
        rot = some_rotation(nxt)
        flat = trial % 4 == 0
        s = np.array([1, 0.3, 0.003 if flat else 0.1]) * np.array([(0.5 + nxt()) * 0.02 for _ in range(3)])
        err, _, _ = rebuilt_error(carried(g, rot, s))
        worst = max(worst, err)
    assert worst < 1e-9, worst


def test_unit_quaternion_500_trials():
    nxt = stream(7)
    for trial in range(500):
        g = np.eye(3) + np.array([nxt() * 2 - 1 for _ in range(9)]).reshape(3, 3) * 0.8
        c = carried(g, some_rotation(nxt), [0.02, 0.01, 0.0002])
        q, _ = O.ellipsoid_axes(six(c))
        assert abs(np.linalg.norm(q) - 1) < 1e-9, trial


def test_degenerate_flattens_instead_of_failing():
    flat = np.diag([1.0, 0.0, 1.0])
    c = carried(flat, some_rotation(stream(3)), [0.02, 0.01, 0.005])
    q, s = O.ellipsoid_axes(six(c))
    assert np.all(np.isfinite(q)) and np.all(np.isfinite(s))
    ss = sorted(s)
    assert ss[0] < 1e-9 and ss[2] > 1e-3


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_eigenvalues_equal_numpy_eigh(seed):
    rng = np.random.default_rng(seed)
    for _ in range(300):
        a = rng.normal(size=(3, 3)) * rng.uniform(1e-3, 1.0, size=3)
        c = a @ a.T
        _, s = O.ellipsoid_axes(six(c))
        ref = np.linalg.eigvalsh(c)
        assert np.allclose(np.sort(s ** 2), ref, rtol=1e-10, atol=1e-14 * ref.max())


def test_axes_are_eigenvectors():
    # R^T C R is diagonal with the squared axis lengths on it, in order.
    rng = np.random.default_rng(9)
    for _ in range(200):
        a = rng.normal(size=(3, 3))
        c = a @ a.T
        q, s = O.ellipsoid_axes(six(c))
        r = rotation_matrix(q)
        d = r.T @ c @ r
        assert np.abs(d - np.diag(s ** 2)).max() < 1e-12 * np.abs(c).max()


def test_f32_form_tracks_f64():
    # The f32 restatement (what the GPU reproduces bit for bit) rebuilds the
    # covariance to f32 precision over the reference test's input family.
    nxt = stream(20260810)
    worst = 0.0
    for trial in range(600):
        amp = [0.05, 0.4, 1.5][trial % 3]
        g = np.eye(3) + np.array([nxt() * 2 - 1 for _ in range(9)]).reshape(3, 3).T * amp
        rot = some_rotation(nxt)
        s = np.array([1, 0.3, 0.003 if trial % 4 == 0 else 0.1]) * np.array([(0.5 + nxt()) * 0.02 for _ in range(3)])
        c = carried(g, rot, s).astype(np.float32).astype(np.float64)
        err, q, _ = rebuilt_error(c, "f32")
        worst = max(worst, err)
        assert abs(np.linalg.norm(np.asarray(q, np.float64)) - 1) < 1e-6
    assert worst < 2e-6, worst


def test_zero_off_diagonal_is_skipped_exactly():
    # An exactly-zero off-diagonal takes no rotation (ellipsoid.js:93-95):
    # only the (0, 1) plane turns here, about +z.
    q, s = O.ellipsoid_axes([2.0, 1.0, 0.0, 2.0, 0.0, 3.0])
    assert q[0] == 0 and q[1] == 0
    assert np.allclose(np.sort(s ** 2), [1, 3, 3], rtol=1e-14)
    assert np.isclose(abs(q[2]), math.sin(math.pi / 8), rtol=1e-14)
