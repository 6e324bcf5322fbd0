"""The oracle's pick and grip (nearestNodeToRay, pick.js:41-68; withinRadius,
grip.js:99-138), pinned by the known answers of
packages/visionary-sim/test/unit/pick.test.js and grip.test.js, and by brute
force over the definitions."""
import numpy as np

from oracle import oracle as O

FORWARD = ([0.0, 0.0, 0.0], [0.0, 0.0, 1.0])


def planes(points):
    return np.asarray(points, np.float32).reshape(-1, 3).T


def pick(points, masses=None, radius=0.1):
    masses = np.ones(len(points), np.float32) if masses is None else np.asarray(masses, np.float32)
    return O.pick_ray(planes(points), masses, *FORWARD, radius)[0]


def test_pick_known_answers():
    assert pick([[1, 0, 1], [0.02, 0, 1], [0, 0.5, 1]]) == 1
    assert pick([[0, 0, 5], [0.05, 0, 1]]) == 1           # the nearer along the ray
    assert pick([[0, 0, 1]], [0]) == -1                    # no material
    assert pick([[0, 0, 1], [0, 0, 2]], [0, 1]) == 1
    assert pick([[0, 0, -1]]) == -1                        # behind the pointer
    assert pick([[0.4, 0, 1], [0, 0.4, 2]]) == -1          # a miss, not the least bad


def test_pick_looks_past_static_material():
    pos = planes([[0, 0, 1], [0.05, 0, 2], [0.2, 0, 1.5]])
    grab = np.array([0, 3, 4], np.float32)  # node 0 static
    assert O.pick_ray(pos, grab, *FORWARD, 0.1)[0] == 1
    grab = np.array([0, 0, 4], np.float32)  # nodes 0 and 1 static
    assert O.pick_ray(pos, grab, *FORWARD, 0.1)[0] == -1


def test_pick_matches_brute_force_argmin():
    rng = np.random.default_rng(3)
    for trial in range(200):
        n = int(rng.integers(1, 400))
        pts = rng.normal(size=(n, 3)).astype(np.float32)
        if trial % 3 == 0:  # lattice-like: exact depth ties
            pts = np.round(pts * 4) / 4
        mass = (rng.random(n) < 0.8).astype(np.float32)
        o = rng.normal(size=3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = float(rng.uniform(0.05, 1.0))
        got, depth = O.pick_ray(pts.T, mass, o, d, r)
        best = (np.inf, -1)
        for i in range(n):
            if mass[i] <= 0:
                continue
            dx, dy, dz = (float(pts[i, a]) - o[a] for a in range(3))
            dep = dx * d[0] + dy * d[1] + dz * d[2]
            if dep <= 0 or dx * dx + dy * dy + dz * dz - dep * dep > r * r:
                continue
            best = min(best, (dep, i))
        assert got == best[1]
        if got >= 0:
            assert depth == best[0]


def lattice(h, extent=0.12):
    n = round(extent / h)
    per = n + 1
    i, j, k = np.meshgrid(np.arange(per), np.arange(per), np.arange(per), indexing="ij")
    pos = np.stack([(i - n / 2) * h, (j - n / 2) * h, (k - n / 2) * h]).reshape(3, -1).astype(np.float32)
    half = n // 2
    centre = (half * per + half) * per + half
    return pos, np.ones(pos.shape[1], np.float32), centre


def grip(pos, mass, anchor, radius, limit=1 << 20, begin=0, end=None):
    end = pos.shape[1] if end is None else end
    return O.grip_radius(pos, mass, anchor, begin, end, radius, limit)


def test_grip_resolution_independent_material():
    coarse = lattice(0.01)
    fine = lattice(0.005)
    nc = grip(*coarse, 0.04)[0]
    nf = grip(*fine, 0.04)[0]
    ratio = (len(nf) * 0.005 ** 3) / (len(nc) * 0.01 ** 3)
    assert 0.9 < ratio < 1.1


def test_grip_hand_is_a_size_in_metres():
    for h in (0.01, 0.005):
        nodes, offs, _ = grip(*lattice(h), 0.04)
        far = np.max(np.linalg.norm(offs, axis=1))
        assert 0.04 - h < far <= 0.04


def test_grip_filters_static_before_counting_and_stays_in_its_body():
    pos = planes([[0, 0, 0], [0.01, 0, 0], [0.02, 0, 0], [-0.02, 0, 0], [0.025, 0, 0], [0.04, 0, 0], [0, 0, 0]])
    mass = np.ones(7, np.float32)
    mass[1] = 0  # static
    nodes, _, found = O.grip_radius(pos, mass, 0, 0, 6, 0.03, 2)
    assert list(nodes) == [0, 2] and found == 4
    nodes, _, _ = O.grip_radius(pos, mass, 0, 0, 6, 0.03, 9)
    assert list(nodes) == [0, 2, 3, 4]


def test_grip_rejects_static_or_empty_anchor():
    pos = np.zeros((3, 2), np.float32)
    mass = np.array([0, 0], np.float32)
    for anchor in (0, 1):
        nodes, offs, found = O.grip_radius(pos, mass, anchor, 0, 2, 1.0, 9)
        assert len(nodes) == 0 and found == 0


def test_grip_anchor_first_then_nearest_with_index_ties():
    pos, mass, centre = lattice(0.01)
    nodes, offs, _ = grip(pos, mass, centre, 0.015)
    assert nodes[0] == centre and np.all(offs[0] == 0)
    dist = np.linalg.norm(offs, axis=1)
    assert np.all(np.diff(dist) >= -1e-12)
    face = list(nodes[1:7])
    assert face == sorted(face)


def test_grip_truncation_and_found():
    pos, mass, centre = lattice(0.01)
    full = grip(pos, mass, centre, 0.03)
    lim = grip(pos, mass, centre, 0.03, limit=5)
    assert list(lim[0]) == list(full[0][:5]) and lim[2] == len(full[0]) and len(lim[0]) == 5


def test_grip_radius_zero_is_single_node():
    pos, mass, centre = lattice(0.01)
    nodes, offs, found = grip(pos, mass, centre, 0.0)
    assert list(nodes) == [centre] and found == 1


def test_grip_never_crosses_bodies():
    pos, mass, centre = lattice(0.01)
    extra = np.array([[0.07], [0], [0]], np.float32)
    pos2 = np.concatenate([pos, extra], axis=1)
    mass2 = np.ones(pos2.shape[1], np.float32)
    nodes, _, _ = O.grip_radius(pos2, mass2, centre, 0, pos.shape[1], 1.0, 1 << 20)
    assert pos.shape[1] not in set(nodes)


def test_grip_matches_brute_force_sort():
    rng = np.random.default_rng(8)
    for trial in range(60):
        n = int(rng.integers(2, 3000))
        pts = (np.round(rng.normal(size=(3, n)) * 8) / 8).astype(np.float32)  # many exact ties
        mass = (rng.random(n) < 0.85).astype(np.float32)
        begin = int(rng.integers(0, n // 2))
        end = int(rng.integers(begin + 1, n + 1))
        anchor = int(rng.integers(begin, end))
        r = float(rng.uniform(0.0, 2.0))
        limit = int(rng.integers(1, 400))
        nodes, offs, found = O.grip_radius(pts, mass, anchor, begin, end, r, limit)
        if mass[anchor] <= 0:
            assert found == 0 and len(nodes) == 0
            continue
        a = pts[:, anchor].astype(np.float64)
        cand = []
        for i in range(begin, end):
            if i == anchor or mass[i] <= 0:
                continue
            dx, dy, dz = (float(pts[k, i]) - a[k] for k in range(3))
            d2 = dx * dx + dy * dy + dz * dz
            if d2 <= r * r:
                cand.append((d2, i))
        cand.sort()
        assert found == len(cand) + 1
        assert list(nodes) == [anchor] + [i for _, i in cand[:limit - 1]]
