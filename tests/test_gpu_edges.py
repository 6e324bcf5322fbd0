"""The reference's edge cases through the HIP kernels, on small hand-built
scenes, against the f32 oracle: bucket keys and accepted contact pairs bit
for bit, positions within 1e-5 of the largest coordinate.

* coincident contact nodes take the +-y normal by index (collide.wgsl:72-80,
  contact.test.js:297-305);
* a pair whose inverse masses sum to zero is skipped (collide.wgsl:60-62);
* a fully collapsed cell (every corner on one point, F = 0) stays finite and
  still (elastic.test.js:370-382), in the colour-pair and one-colour passes;
* records outside the grid are clamped into its edge buckets
  (contacts.wgsl:84-91, contact.test.js:131-137);
* splats outside their lattice extrapolate from the nearest cell
  (sample.js:42-55);
* several overfull contact scenes share the world's overflow list in one
  batched, graph-replayed launch and still equal each scene alone;
* timing a pass leaves the simulation where it was, and contact passes are
  refused by vs_world_time_pass;
* a removal does not wait for the frames already queued, and the scene
  admitted into the freed ranges is written only after them.
"""
import numpy as np
import pytest

from oracle.oracle import OracleScene
from test_gpu_parity import RTOL, rel_err, world_for
from vsim import native, scenes

pytestmark = pytest.mark.gpu

H = 1.0 / 32


def two_blocks(static_a=(), static_b=(), gravity=(0.0, 0.0, 0.0), gap=0.0, y=0.25):
    """Blocks A (x in [-2h, 0]) and B (x in [gap, gap + 2h]) of 2^3 cells: at
    gap 0 B's x = 0 face nodes sit exactly on A's."""
    mat = dict(youngs_modulus=5e4, poisson_ratio=0.3, density=1000.0)
    A = scenes.Body(scenes.block_lattice((2, 2, 2), (-2 * H, y, 0.0), H, 1000.0), scenes.Material(**mat))
    B = scenes.Body(scenes.block_lattice((2, 2, 2), (gap, y, 0.0), H, 1000.0), scenes.Material(**mat))
    A.static_nodes = np.asarray(static_a, np.uint32)
    B.static_nodes = np.asarray(static_b, np.uint32)
    rate = max(scenes.wave_speed(b.material) / H for b in (A, B))
    return scenes.Scene([A, B], dt=float(np.float32(scenes.stable_step_duration(rate))), substeps=4, gravity=gravity,
                        tank_min=(-0.5, 0.0, -0.5), tank_max=(0.5, 1.0, 0.5), contact=1, name="pair")


def frames_against_oracle(sc, frames, cap=32, **kw):
    w = world_for(sc, **kw)
    w.log_pairs(cap)
    sid = w.add(sc)
    o = OracleScene(sc, "f32", pair_cap=cap)
    for f in range(frames):
        w.frame(sc.substeps, False, f > 0)
        o.frame(sc.substeps)
        pos, vel = w.read_nodes(sid)
        assert np.all(np.isfinite(pos))
        assert rel_err(pos, o.p) <= RTOL and rel_err(vel, o.v) <= RTOL
        if sc.contact_active:
            keys, counts, pairs = w.read_contacts(sid)
            assert np.array_equal(keys, o.keys)
            assert np.array_equal(counts, o.pair_count)
            for s in np.nonzero(counts)[0]:
                n = min(counts[s], cap)
                assert np.array_equal(pairs[s, :n], o.pairs[s, :n])
    return w, sid, o, pos


def face(i):
    """Body-local node ids of the x = i face of a 3^3-node block."""
    return [i + 3 * (j + 3 * k) for k in range(3) for j in range(3)]


def test_coincident_contact_nodes():
    sc = two_blocks()
    # The branch is taken: after the first substep's predict and elastic
    # sweep, face pairs are within 1e-9 m of each other (the collide test).
    o = OracleScene(sc, "f32")
    o.predict()
    for c in range(8):
        o.elastic(c)
    p = o.p.T.astype(np.float64)
    close = [np.linalg.norm(p[a] - p[27 + b]) <= 1e-9 for a, b in zip(face(2), face(0))]
    assert sum(close) >= 3
    w, sid, o, pos = frames_against_oracle(sc, 2)
    # coincident nodes are pushed apart along y (the normal (0, +-1, 0))
    d = pos[:, face(2)] - pos[:, [27 + n for n in face(0)]]
    assert np.max(np.abs(d[1])) > 1e-4


def test_zero_mass_pair_is_skipped():
    # A wholly static, B's touching face static: the face pairs sum to w = 0,
    # so no correction; B's other nodes and A stay exactly where they were.
    sc = two_blocks(static_a=range(27), static_b=face(0))
    w, sid, o, pos = frames_against_oracle(sc, 2)
    rest = sc.arrays()["pos"].T
    assert np.array_equal(pos[:, :27], rest[:, :27])
    assert np.array_equal(pos[:, [27 + n for n in face(0)]], rest[:, [27 + n for n in face(0)]])
    keys, counts, pairs = w.read_contacts(sid)
    assert counts.sum() > 0          # the pairs pass the reach test and are then skipped


@pytest.mark.parametrize("separate_colours", [0, 1])
def test_collapsed_cells_stay_finite(separate_colours):
    sc = scenes.jelly_block(3, h=H, pinned=False, tank_margin=1.0, name="collapsed")
    sc.gravity = (0.0, 0.0, 0.0)
    lat = sc.bodies[0].lattice
    lat.rest[:] = lat.rest[13].copy()                  # every node on one point: F = 0
    w, sid, o, pos = frames_against_oracle(sc, 2, separate_colours=separate_colours)
    assert np.array_equal(pos, np.tile(lat.rest[13][:, None], (1, lat.node_count)))


def test_records_outside_the_grid_are_clamped():
    # Bodies above the tank's open top (no y = max plane) and one starting
    # past x = min - edge: their records bin into the grid's edge buckets.
    sc = two_blocks(gravity=(0.0, -9.81, 0.0), gap=-0.25 * H, y=1.3)
    sc.bodies[0].lattice.rest[:, 0] -= 0.7
    sc.bodies[1].lattice.rest[:, 0] -= 0.7
    sc.substeps = 1                  # the keys read back are the first substep's, before finalize's walls
    origin, edge, dims = sc.grid()
    w, sid, o, pos = frames_against_oracle(sc, 1)
    keys = o.keys
    gx, gy = keys % dims[0], (keys // dims[0]) % dims[1]
    assert np.all(gy == dims[1] - 1) and np.any(gx == 0)


def test_splats_outside_the_lattice_extrapolate():
    sc = scenes.jelly_block(4, h=H, splats=64, seed=21, pinned=True, name="outside")
    b = sc.bodies[0]
    rng = np.random.default_rng(5)
    lo, hi = np.asarray(b.lattice.origin), np.asarray(b.lattice.origin) + 4 * H
    far = np.concatenate([lo - rng.uniform(0.05, 0.8, (32, 3)) * H, hi + rng.uniform(0.05, 0.8, (32, 3)) * H])
    b.splat_mean = far.astype(np.float32)
    w = world_for(sc)
    sid = w.add(sc)
    o = OracleScene(sc, "f32")
    for f in range(2):
        w.frame(sc.substeps, True, False)
        o.frame(sc.substeps, splats=True)
    mean, cov = w.read_splats(sid)
    assert np.all(np.isfinite(mean)) and np.all(np.isfinite(cov))
    assert rel_err(mean, o.out_mean) <= RTOL and rel_err(cov, o.out_cov) <= RTOL
    # extrapolated, not clamped: the means stay outside the deformed block
    pos, _ = w.read_nodes(sid)
    assert np.any(mean[0] < pos[0].min()) and np.any(mean[0] > pos[0].max())


def test_overfull_scenes_share_the_overflow_list():
    mix = []
    for count, side, bucket in ((6, 8, 0.08), (8, 8, 0.25), (4, 6, 0.05)):
        sc = scenes.contact_stack(count, side=side, seed=count)
        sc.bucket_size = bucket
        mix.append(sc)
    w = world_for(*mix)
    ids = [w.add(s) for s in mix]
    for f in range(3):
        w.frame(8, False, f > 0)
    together = [w.read_nodes(i)[0] for i in ids]
    for sc, got in zip(mix, together):
        alone = world_for(sc)
        a = alone.add(sc)
        o = OracleScene(sc, "f32")
        for f in range(3):
            alone.frame_scene(a, 8)
            o.frame(8)
        keys = o.keys
        assert int((np.bincount(keys) > 32).sum()) > 0
        assert np.array_equal(alone.read_nodes(a)[0], got)
        assert rel_err(got, o.p) <= RTOL
        alone.close()


def test_time_pass_leaves_the_simulation_and_refuses_contact():
    sc = scenes.c3_scene(bodies=8, seed=9)
    worlds = []
    for timed in (False, True):
        w = world_for(sc)
        sid = w.add(sc)
        w.frame(sc.substeps, False, False)
        if timed:
            ms, n = w.time_pass("elastic", 3)
            assert n > 0 and ms > 0
            for kind in ("bin", "sort", "collide", "settle"):
                with pytest.raises(native.VsimError, match="contact passes"):
                    w.time_pass(kind, 1)
        w.frame(sc.substeps, False, False)
        worlds.append(w.read_nodes(sid))
    assert np.array_equal(worlds[0][0], worlds[1][0]) and np.array_equal(worlds[0][1], worlds[1][1])


def test_removal_leaves_queued_frames_and_admission_waits_for_them():
    # Scene A's frames are queued and A is removed without waiting; B is
    # admitted into the ranges A held (the world's only scene) while A's
    # frames may still be running, then stepped. B must equal B stepped in a
    # world of its own, and A's queued frames must not see B's uploads.
    a = scenes.jelly_block(20, seed=31, name="a")
    b = scenes.jelly_block(20, seed=32, pinned=False, name="b")
    w = world_for(a)                 # room for one of the two: B takes A's ranges
    sa = w.add(a)
    for f in range(6):
        w.frame(a.substeps, False, True)
    w.remove(sa)
    sb = w.add(b)
    for f in range(3):
        w.frame(b.substeps, False, True)
    got = w.read_nodes(sb)
    w.close()
    alone = world_for(b)
    s = alone.add(b)
    for f in range(3):
        alone.frame(b.substeps, False, True)
    ref = alone.read_nodes(s)
    alone.close()
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
