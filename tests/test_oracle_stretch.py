"""Pins the oracle's tensile limit (the stretch sweeps of project.wgsl:162-179,
restated in oracle.c's cell projection) with the known answers of the
reference's stretch.test.js:5-31, driven through the oracle's own cell path:
a one-cell lattice (h = 1 m, so the limits are the unit-cell lengths times
1.05) swept with the limit-only flag, corners placed so that exactly one
corner pair is past its limit.

* 0.475: two free ends 2 m apart on an axis edge each move 0.475 m inward;
* a prescribed end gives the free end the whole correction, and two
  prescribed ends none;
* unequal weights keep the centre of mass (masses 1/w) and leave the pair at
  exactly the limit;
* compression and rigid motion are left alone;
* the 28 corner pairs are the cube's 12 edges, 12 face and 4 body diagonals
  with lengths 1, sqrt 2 and sqrt 3, and every one of them limits at its own
  length (a wrong length in any entry moves that pair to the wrong spot).
"""
import itertools
import math

import numpy as np
import pytest

from vsim import scenes

LIMIT = 1.05
CORNERS = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], float)


def one_cell(oracle, precision="f64", max_stretch=LIMIT):
    mat = dict(youngs_modulus=1e5, poisson_ratio=0.3, density=1000.0, max_stretch=max_stretch)
    sc = scenes.jelly_block(1, h=1.0, pinned=False, material=mat, tank_margin=10.0)
    return oracle.OracleScene(sc, precision)


def sweep(o, pos, w):
    """Corner positions (8, 3) and inverse masses through one limit-only
    sweep of the cell (colour 0 holds the lone cell): corner c is node c."""
    o.p[:] = np.asarray(pos, float).T
    o.w[:] = w
    o.elastic(0, limit_only=1)
    return o.p.T.astype(float).copy()


def placed(a, b, pa, pb, rest):
    """Corners a and b at pa, pb; every other corner at their midpoint, where
    it is within 0.6 of a unit length of both (inside every limit)."""
    pos = np.tile((np.asarray(pa, float) + np.asarray(pb, float)) / 2, (8, 1))
    pos[a], pos[b] = pa, pb
    return pos


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_lengthening_is_corrected_and_compression_is_not(oracle, precision):
    # stretch.test.js:6-11: stretchCorrection([2,3,4], [2,5,4], 1, 1, 1.05)[1] = 0.475
    o = one_cell(oracle, precision)
    w = np.array([1, 0, 1, 0, 0, 0, 0, 0], float)          # corners 0 and 2: the y edge
    out = sweep(o, placed(0, 2, [2, 3, 4], [2, 5, 4], 1.0), w)
    tol = 1e-12 if precision == "f64" else 2e-6
    assert out[0] == pytest.approx([2, 3.475, 4], abs=tol)
    assert out[2] == pytest.approx([2, 4.525, 4], abs=tol)
    for far in ([2, 4, 4], [2, 3.5, 4]):                     # at and below rest length: no force
        pos = placed(0, 2, [2, 3, 4], far, 1.0)
        assert np.array_equal(sweep(o, pos, w), pos)


def test_prescribed_end_gives_the_free_end_the_whole_correction(oracle):
    # stretch.test.js:13-18: a prescribed, b free with w = 4: b lands at 1.05.
    o = one_cell(oracle)
    pos = placed(0, 1, [0, 0, 0], [2, 0, 0], 1.0)
    out = sweep(o, pos, np.array([0, 4, 0, 0, 0, 0, 0, 0], float))
    assert out[1] == pytest.approx([1.05, 0, 0], abs=1e-12)
    assert np.array_equal(out[0], pos[0])
    # both ends prescribed: nothing moves
    assert np.array_equal(sweep(o, pos, np.zeros(8)), pos)


def test_unequal_weights_keep_the_centre_of_mass(oracle):
    # stretch.test.js:20-25: w = 2 and 1; the pair ends 1.05 apart and the
    # mass-weighted centre (a/2 + b)/1.5 stays at 2/1.5.
    o = one_cell(oracle)
    out = sweep(o, placed(0, 1, [0, 0, 0], [2, 0, 0], 1.0), np.array([2, 1, 0, 0, 0, 0, 0, 0], float))
    a, b = out[0, 0], out[1, 0]
    assert b - a == pytest.approx(1.05, abs=1e-12)
    assert (a / 2 + b) / 1.5 == pytest.approx(2 / 1.5, abs=1e-12)
    assert a > 0 and b < 2                                   # pulled together, never pushed apart


def test_rigid_motion_and_uniform_compression_are_untouched(oracle):
    o = one_cell(oracle)
    rng = np.random.default_rng(7)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, s = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - s * z), 2 * (x * z + s * y)],
                  [2 * (x * y + s * z), 1 - 2 * (x * x + z * z), 2 * (y * z - s * x)],
                  [2 * (x * z - s * y), 2 * (y * z + s * x), 1 - 2 * (x * x + y * y)]])
    w = np.ones(8)
    for pos in (CORNERS @ R.T + [3.0, -1.0, 2.0], CORNERS * 0.7):
        assert np.array_equal(sweep(o, pos, w), pos)


def test_the_edge_table_is_every_corner_pair_with_its_length(oracle):
    # stretch.test.js:27-31: 28 distinct pairs, each at hypot(corner a - corner b).
    a, b, length = oracle.stretch_edges()
    pairs = set(zip(a.tolist(), b.tolist()))
    assert len(pairs) == 28 and pairs == set(itertools.combinations(range(8), 2))
    for i, j, L in zip(a, b, length):
        assert L == pytest.approx(math.dist(CORNERS[i], CORNERS[j]), abs=1e-7)
    counts = {k: int(np.sum(np.abs(length - math.sqrt(k)) < 1e-6)) for k in (1, 2, 3)}
    assert counts == {1: 12, 2: 12, 3: 4}


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_every_pair_limits_at_its_own_length(oracle, precision):
    # Each of the 28 pairs, stretched to 1.2 x its unit length along a seeded
    # direction with both ends free: both move inward by (1.2 - 1.05) L / 2,
    # along the pair, and the pair ends at 1.05 L.
    o = one_cell(oracle, precision)
    a, b, _ = oracle.stretch_edges()
    rng = np.random.default_rng(31)
    # the table holds the lengths as f32 literals (sqrt 2 and sqrt 3 rounded),
    # as the kernel's constant table does: 1e-7 in the f64 build
    tol = 1e-7 if precision == "f64" else 4e-6
    for i, j in zip(a, b):
        L = math.dist(CORNERS[i], CORNERS[j])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pa = rng.uniform(-1, 1, 3)
        pb = pa + 1.2 * L * u
        w = np.zeros(8)
        w[i] = w[j] = 1.0
        pos = placed(i, j, pa, pb, L).astype(o.dtype).astype(float)
        out = sweep(o, pos, w)
        assert out[i] == pytest.approx(pa + 0.075 * L * u, abs=tol), (i, j)
        assert out[j] == pytest.approx(pb - 0.075 * L * u, abs=tol), (i, j)
        assert np.linalg.norm(out[j] - out[i]) == pytest.approx(LIMIT * L, abs=tol)
        rest = [k for k in range(8) if k not in (i, j)]
        assert np.array_equal(out[rest], pos[rest])


def test_limit_off_at_or_below_one(oracle):
    # project.wgsl:162: the sweeps run only when maxStretch > 1.
    for m in (0.0, 1.0):
        o = one_cell(oracle, max_stretch=m)
        pos = placed(0, 1, [0, 0, 0], [2, 0, 0], 1.0)
        assert np.array_equal(sweep(o, pos, np.ones(8)), pos)
