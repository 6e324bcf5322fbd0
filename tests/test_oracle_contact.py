"""Pins the oracle's contact passes (bin.wgsl, sort.wgsl, collide.wgsl,
settle.wgsl) against the known answers of the reference's contact.test.js
(one pair: equal split, mass ratio, inert node, coincident nodes, Coulomb
stick/slide, the lower friction) and against brute force: the grid must find
exactly the pairs an all-pairs reach test finds, and an overfull bucket keeps
the lowest record ids."""
import numpy as np
import pytest

from _raw import Raw
from vsim import scenes

R = 0.01


def pair(a_pos, b_pos, a_prev=None, b_prev=None, wa=1.0, wb=1.0, fa=0.0, fb=0.0, precision="f64"):
    pos = [a_pos, b_pos]
    prev = [a_prev if a_prev is not None else a_pos, b_prev if b_prev is not None else b_pos]
    r = Raw(pos, prev=prev, w=[wa, wb], bodies=((fa, 0.0), (fb, 0.0)), body_of=[0, 1], precision=precision)
    r.surface([0, 1], R)
    before = r.pos.copy()
    r.call("o_contact")
    return r.pos - before, r


def test_equal_nodes_split_the_overlap():
    d, _ = pair([0, 0, 0], [0.02 - 0.004, 0, 0])
    assert d[0, 0] == pytest.approx(-0.002, abs=1e-12)
    assert d[1, 0] == pytest.approx(0.002, abs=1e-12)
    assert np.all(np.abs(d[0] + d[1]) < 1e-15)


def test_lighter_node_gives_way():
    d, _ = pair([0, 0, 0], [0.016, 0, 0], wa=1.0, wb=0.1)
    assert d[0, 0] == pytest.approx(-0.004 / 1.1, abs=1e-12)
    assert d[1, 0] == pytest.approx(0.004 * 0.1 / 1.1, abs=1e-12)


def test_inert_node_stays():
    d, _ = pair([0.016, 0, 0], [0, 0, 0], wa=0.0, wb=1.0)
    assert np.all(d[0] == 0)
    assert d[1, 0] == pytest.approx(-0.004, abs=1e-12)


def test_merely_close_pairs_do_nothing():
    for x in (0.02, 0.05):
        d, r = pair([0, 0, 0], [x, 0, 0])
        assert np.all(d == 0) and np.all(r.pair_count == 0)


def test_coincident_nodes_separate_along_y_by_index():
    d, _ = pair([0, 0, 0], [0, 0, 0])
    assert np.all(np.isfinite(d))
    assert d[0, 1] == pytest.approx(0.01, abs=1e-12)    # lower node index -> +y
    assert d[1, 1] == pytest.approx(-0.01, abs=1e-12)


def test_coulomb_stick_slide_and_frictionless():
    ov = 0.004

    def slide(s, mu):
        d, _ = pair([0, 0, 0], [0.02 - ov, 0, 0], a_prev=[0, -s, 0], fa=mu, fb=mu)
        return d[0]

    assert slide(0.0002, 0.5)[1] == pytest.approx(-0.0002 / 2, abs=1e-12)
    assert slide(0.01, 0.5)[1] == pytest.approx(-(0.5 * ov) / 2, abs=1e-12)
    free = slide(0.01, 0.0)
    assert free[1] == pytest.approx(0, abs=1e-12) and free[0] == pytest.approx(-ov / 2, abs=1e-12)


def test_lower_friction_wins():
    d, _ = pair([0, 0, 0], [0.016, 0, 0], a_prev=[0, -0.01, 0], fa=0.9, fb=0.1)
    assert d[0, 1] == pytest.approx(-(0.1 * 0.004) / 2, abs=1e-12)


def test_bucket_clamps_outside_points():
    r = Raw([[5.0, -5.0, 0.0], [-5.0, 5.0, 0.0]], body_of=[0, 1], bodies=((0, 0), (0, 0)))
    r.surface([0, 1], R, origin=(-1.0, -1.0, -1.0), size=0.02, dims=(10, 10, 10))
    r.call("o_contact")
    # (clamp(300)=9, clamp(-200)=0, 50->9) and (0, 9, 9)
    assert list(r.keys) == [9 + 10 * (0 + 10 * 9), 0 + 10 * (9 + 10 * 9)]


def test_points_within_a_bucket_edge_are_at_most_one_bucket_apart():
    rng = np.random.default_rng(7)
    size = 0.02
    pts = []
    for _ in range(2000):
        p = rng.uniform(-0.5, 0.5, 3)
        q = p + rng.normal(size=3) * size / 3
        if np.linalg.norm(q - p) < size:
            pts += [p, q]
    pts = np.array(pts)
    n = len(pts)
    r = Raw(pts, body_of=np.arange(n) % 2, bodies=((0, 0), (0, 0)))
    r.surface(list(range(n)), 0.0, origin=(-1.0, -1.0, -1.0), size=size, dims=(100, 100, 100), pair_cap=1)
    r.call("o_contact")
    k = r.keys.astype(np.int64)
    c = np.stack([k % 100, (k // 100) % 100, k // 10000], axis=1)
    assert np.all(np.abs(c[0::2] - c[1::2]) <= 1)


def brute_pairs(pos, body, radius):
    p = pos.astype(np.float32)
    out = set()
    n = len(p)
    for s in range(n):
        d = p[s] - p
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        reach = (radius[s] + radius).astype(np.float32)
        hit = (d2 < reach * reach) & (body != body[s])
        out |= {(s, int(o)) for o in np.nonzero(hit)[0]}
    return out


def test_grid_finds_exactly_the_brute_force_pairs():
    rng = np.random.default_rng(3)
    n = 400
    pos = rng.uniform(0, 0.12, (n, 3))
    body = rng.integers(0, 4, n)
    order = np.argsort(body, kind="stable")
    pos, body = pos[order], body[order]
    r = Raw(pos, body_of=body, bodies=((0.4, 0),) * 4, precision="f32")
    r.surface(list(range(n)), 0.005, origin=(-0.01, -0.01, -0.01), size=0.01, dims=(15, 15, 15), pair_cap=128)
    before = r.pos
    r.call("o_contact")
    assert np.bincount(r.keys).max() <= 32
    got = {(s, int(r.pairs[s, k])) for s in range(n) for k in range(r.pair_count[s])}
    want = brute_pairs(before, body, np.full(n, 0.005, np.float32))
    assert len(want) > 50
    assert got == want


def test_overfull_bucket_keeps_the_lowest_record_ids():
    # 40 nodes of body 1 packed into one bucket with one node of body 0: the
    # bucket keeps records 0..31 (sort.wgsl:27-37), so the body-0 record sees
    # exactly 1..31 -- never one of the ten highest ids.
    rng = np.random.default_rng(5)
    pts = np.concatenate([[[0.005, 0.005, 0.005]], 0.001 + rng.uniform(0, 0.008, (40, 3))])
    body = np.array([0] + [1] * 40)
    r = Raw(pts, body_of=body, bodies=((0, 0), (0, 0)))
    r.surface(list(range(41)), 0.05, origin=(0.0, 0.0, 0.0), size=0.01, dims=(3, 3, 3), pair_cap=64)
    r.call("o_contact")
    assert r.pair_count[0] == 31
    assert sorted(r.pairs[0, :31]) == list(range(1, 32))


def test_worst_case_occupancy_known_answers():
    assert scenes.worst_case_bucket_occupancy(0.01, 0.01) == 8
    assert scenes.worst_case_bucket_occupancy(0.05, 0.005) > 32
    assert scenes.worst_case_bucket_occupancy(0.0142, 0.0112) <= 32


def test_two_body_scene_contact_pushes_the_bodies_apart():
    # Two equal blocks overlapping by half a cell: one contact pass moves the
    # centroids apart along x and nothing else of note.
    a = scenes.block_lattice((3, 3, 3), (0.0, 0.0, 0.0), 0.01, 1000.0)
    b = scenes.block_lattice((3, 3, 3), (0.025, 0.0, 0.0), 0.01, 1000.0)
    mat = scenes.Material(friction=0.0)
    sc = scenes.Scene([scenes.Body(a, mat), scenes.Body(b, mat)], dt=1 / 600, gravity=(0, 0, 0),
                      tank_min=(-1, -1, -1), tank_max=(1, 1, 1), contact=1)
    from oracle.oracle import OracleScene
    o = OracleScene(sc, "f64")
    before = o.p.copy()
    o.contact()
    d = o.p - before
    assert d[0, :a.node_count].sum() < -1e-4 and d[0, a.node_count:].sum() > 1e-4
    assert np.all(d[0, :a.node_count][before[0, :a.node_count] < 0.0149] == 0)
    assert np.max(np.abs(d)) > 1e-4
