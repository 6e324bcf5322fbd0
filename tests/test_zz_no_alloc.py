"""vs_grip_noalloc (the grip with the library's static candidate array and the
question passed by value) against the grip of vs_dispatch with VS_K_GRIP
(withinRadius, grip.js:99-138), which takes its candidate queue from a caller
buffer: node ids, offsets, found counts and status must be identical, on two
calls in a row, the second with a smaller input."""
import numpy as np
import pytest

from vsim import native

pytestmark = pytest.mark.gpu


def lattice(seed, n_side):
    rng = np.random.default_rng(seed)
    i, j, k = np.meshgrid(*(np.arange(n_side),) * 3, indexing="ij")
    pos = np.stack([i, j, k]).reshape(3, -1).astype(np.float64) * 0.01
    pos += rng.normal(scale=0.002, size=pos.shape) * (rng.random(pos.shape[1]) < 0.5)
    mass = (rng.random(pos.shape[1]) < 0.9).astype(np.float32) * rng.uniform(0.5, 2.0, pos.shape[1]).astype(np.float32)
    return pos.astype(np.float32), mass


def read_out(b, out, limit):
    hdr = np.frombuffer(b.read(out, 32), np.uint32).copy()
    kept = int(hdr[3])
    nodes = np.frombuffer(b.read(out, 4 * kept, 32), np.uint32).copy() if kept else np.zeros(0, np.uint32)
    at = 32 + 4 * (limit + (limit & 1))
    offs = np.frombuffer(b.read(out, 24 * kept, at), np.float64).reshape(-1, 3).copy() if kept else np.zeros((0, 3))
    return hdr, nodes, offs


def both(b, pos, mass, anchor, begin, end, radius, limit, grid=32):
    n = pos.shape[1]
    hp = b.alloc(12 * n)
    b.write(hp, np.ascontiguousarray(pos, np.float32))
    hm = b.alloc(4 * n)
    b.write(hm, np.ascontiguousarray(mass, np.float32))
    out_bytes = 32 + 4 * (limit + (limit & 1)) + 24 * limit
    out_a, out_b = b.alloc(out_bytes), b.alloc(out_bytes)
    params = b.alloc(80)
    cand = b.alloc(16 * n)
    q = native.VsGrabParams()
    q.radius, q.anchor, q.begin, q.end = radius, anchor, begin, end
    b.write(params, q)
    launch = b.launch("sim-grip-radius", grid, plane=n, items=n, pair_cap=limit, nodePos=hp, nodeInvMass=hm,
                      params=params, out=out_a)
    launch.buffers[native.SLOTS.index("items")] = cand
    b.dispatch(launch)
    b.grip_noalloc(hp, n, hm, q, limit, grid, out_b)
    b.sync()
    got = read_out(b, out_b, limit), read_out(b, out_a, limit)
    for h in (hp, hm, out_a, out_b, params, cand):
        b.free(h)
    return got


def test_noalloc_grip_equals_the_dispatched_grip_twice():
    b = native.Backend(0)
    # First call: 64k nodes, a 0.12 m hand (tens of thousands of candidates,
    # the radix selection past 4096). Second: a smaller body and hand.
    cases = [(40, 0.12, 4096, 3), (12, 0.03, 64, 5)]
    for n_side, radius, limit, seed in cases:
        pos, mass = lattice(seed, n_side)
        n = pos.shape[1]
        begin, end = n // 8, n - n // 8
        anchor = int(np.nonzero(mass[begin:end] > 0)[0][n // 3]) + begin
        print(f"noalloc grip: {n} nodes, radius {radius}, limit {limit}, anchor {anchor}", flush=True)
        (h1, n1, o1), (h0, n0, o0) = both(b, pos, mass, anchor, begin, end, radius, limit)
        assert h0[4] == 0 and h1[4] == 0                     # status ok
        assert h1[2] == h0[2] and h1[3] == h0[3]             # found, kept
        assert h1[0] == 0 and h1[1] == 0                     # counters of the output header left clear
        assert h0[3] > 1
        assert np.array_equal(n1, n0)
        assert np.array_equal(o1, o0)
    b.close()


def test_noalloc_grip_rejects_a_body_beyond_its_array():
    b = native.Backend(0)
    pos, mass = lattice(1, 4)
    n = pos.shape[1]
    hp, hm, out = b.alloc(12 * n), b.alloc(4 * n), b.alloc(32 + 8 + 24)
    q = native.VsGrabParams()
    q.radius, q.anchor, q.begin, q.end = 0.01, 0, 0, (1 << 18) + 1
    with pytest.raises(native.VsimError, match="exceeds the static candidate array"):
        b.grip_noalloc(hp, n, hm, q, 1, 8, out)
    b.close()
