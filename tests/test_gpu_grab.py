"""Pick and grip on the device (sim-pick-ray, sim-grip-radius) through the
layer-1 C ABI, against the oracle's nearestNodeToRay / withinRadius
(pick.js:41-68, grip.js:99-138). Both compute in f64 like the JS, so the
answers must be identical: node ids, depths, found counts and offsets."""
import numpy as np
import pytest

from oracle import oracle as O
from vsim import native

pytestmark = pytest.mark.gpu

OUT_PICK_GRID = 64


class Grab:
    def __init__(self, pos, mass, limit=4096):
        self.b = native.Backend(0)
        self.n = pos.shape[1]
        self.limit = limit
        self.pos = self.b.alloc(12 * self.n)
        self.b.write(self.pos, np.ascontiguousarray(pos, np.float32))
        self.mass = self.b.alloc(4 * self.n)
        self.b.write(self.mass, np.ascontiguousarray(mass, np.float32))
        self.params = self.b.alloc(C_SIZE)
        self.pick_out = self.b.alloc(16 + 16 * OUT_PICK_GRID)
        self.cand = self.b.alloc(16 * self.n)
        self.grip_out = self.b.alloc(32 + 4 * (limit + (limit & 1)) + 24 * limit)

    def pick(self, origin, direction, radius, grid=OUT_PICK_GRID):
        p = native.VsGrabParams()
        p.origin[:] = origin
        p.direction[:] = direction
        p.radius = radius
        self.b.write(self.params, p)
        self.b.dispatch(self.b.launch("sim-pick-ray", grid, plane=self.n, nodePos=self.pos, nodeInvMass=self.mass,
                                      params=self.params, out=self.pick_out))
        raw = self.b.read(self.pick_out, 16)
        return int(np.frombuffer(raw[4:8], np.int32)[0]), float(np.frombuffer(raw[8:16], np.float64)[0])

    def grip(self, anchor, begin, end, radius, limit=None, grid=32):
        limit = self.limit if limit is None else limit
        p = native.VsGrabParams()
        p.radius = radius
        p.anchor, p.begin, p.end = anchor, begin, end
        self.b.write(self.params, p)
        launch = self.b.launch("sim-grip-radius", grid, plane=self.n, items=self.n, pair_cap=limit,
                               nodePos=self.pos, nodeInvMass=self.mass, params=self.params, out=self.grip_out)
        launch.buffers[native.SLOTS.index("items")] = self.cand
        self.b.dispatch(launch)
        hdr = np.frombuffer(self.b.read(self.grip_out, 32), np.uint32)
        found, kept, status = int(hdr[2]), int(hdr[3]), int(hdr[4])
        nodes = np.frombuffer(self.b.read(self.grip_out, 4 * kept, 32), np.uint32) if kept else np.zeros(0, np.uint32)
        off_at = 32 + 4 * (limit + (limit & 1))
        offs = (np.frombuffer(self.b.read(self.grip_out, 24 * kept, off_at), np.float64).reshape(-1, 3)
                if kept else np.zeros((0, 3)))
        return nodes, offs, found, status


C_SIZE = 80


def scene_positions(seed, n_side=40, jitter=True):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij")
    pos = np.stack([i, j, k]).reshape(3, -1).astype(np.float64) * 0.01
    if jitter:
        pos += rng.normal(scale=0.002, size=pos.shape) * (rng.random(pos.shape[1]) < 0.5)
    mass = (rng.random(pos.shape[1]) < 0.9).astype(np.float32) * rng.uniform(0.5, 2.0, pos.shape[1]).astype(np.float32)
    return pos.astype(np.float32), mass


@pytest.mark.parametrize("jitter", [False, True])
def test_pick_equals_oracle(jitter):
    pos, mass = scene_positions(1, jitter=jitter)
    g = Grab(pos, mass)
    rng = np.random.default_rng(2)
    hits = 0
    for trial in range(60):
        target = pos[:, rng.integers(0, pos.shape[1])].astype(np.float64)
        origin = target + rng.normal(size=3) * 0.5
        d = target - origin + (rng.normal(size=3) * 0.01 if trial % 2 else 0)
        d /= np.linalg.norm(d)
        r = [0.012, 0.004, 0.05][trial % 3]
        want = O.pick_ray(pos, mass, origin, d, r)
        got = g.pick(origin, d, r, grid=[1, 7, 64][trial % 3])
        assert got[0] == want[0], trial
        if want[0] >= 0:
            hits += 1
            assert got[1] == want[1]
    assert hits > 30
    # a ray pointing away misses
    assert g.pick([0, 0, -1], [0, 0, -1], 0.1)[0] == -1


@pytest.mark.parametrize("radius,limit", [(0.0, 4096), (0.015, 4096), (0.03, 4096), (0.03, 5), (0.05, 4096),
                                          (0.12, 4096), (0.12, 300), (0.3, 4096)])
def test_grip_equals_oracle(radius, limit):
    # 0.12 and 0.3 m hold tens of thousands of candidates: the radix selection.
    pos, mass = scene_positions(3, jitter=radius != 0.015)
    mass[0] = 1.0
    g = Grab(pos, mass, limit=4096)
    n = pos.shape[1]
    begin, end = n // 8, n - n // 8
    rng = np.random.default_rng(4)
    for trial in range(4):
        anchor = int(rng.integers(begin, end))
        want = O.grip_radius(pos, mass, anchor, begin, end, radius, limit)
        got = g.grip(anchor, begin, end, radius, limit=limit)
        if mass[anchor] <= 0:
            assert got[2] == 0 and got[3] == 1 and len(got[0]) == 0
            continue
        assert got[3] == 0
        assert got[2] == want[2]
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])


def test_grip_rejects_bad_range_and_reuses_its_buffer():
    pos, mass = scene_positions(5, n_side=12)
    g = Grab(pos, mass, limit=64)
    n = pos.shape[1]
    assert g.grip(5, 10, n, 0.02, limit=64)[3] == 3   # anchor outside its body
    anchor = int(np.nonzero(mass > 0)[0][20])
    first = g.grip(anchor, 0, n, 0.02, limit=64)
    second = g.grip(anchor, 0, n, 0.02, limit=64)
    assert np.array_equal(first[0], second[0]) and first[2] == second[2]
