"""GPU parity: the HIP path through the C ABI against the f32 oracle on the
same seeded inputs. Positions and velocities must agree to 1e-5 relative
(BASELINE tolerance); with -ffp-contract=off and IEEE divide/sqrt the kernels
reproduce the oracle's arithmetic, so the tests also report exact agreement."""
import numpy as np
import pytest

from vsim import native, scenes

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def world_for(*scs, **kw):
    nodes = sum(s.n_nodes for s in scs)
    cells = sum(s.n_cells for s in scs)
    splats = sum(s.n_splats for s in scs)
    surf = sum(s.n_surface_active for s in scs)
    buckets = sum(int(np.prod(s.grid()[2])) for s in scs if s.contact_active)
    driven = sum(len(s.drive.nodes) for s in scs if s.drive is not None)
    return native.World(max_scenes=max(4, len(scs)), max_bodies=max(4, sum(len(s.bodies) for s in scs)),
                        max_nodes=nodes + 16, max_cells=cells + 16, max_surface=surf + 16, max_buckets=buckets + 16,
                        max_driven=driven + 16, max_splats=splats + 16, **kw)


def rel_err(gpu, ref):
    scale = max(float(np.max(np.abs(ref))), 1e-30)
    return float(np.max(np.abs(gpu.astype(np.float64) - ref.astype(np.float64)))) / scale


def step_both(sc, frames, substeps=None, graph=False, plain=False, splats=False, **kw):
    from oracle.oracle import OracleScene
    w = world_for(sc, **kw)
    sid = w.add(sc)
    o = OracleScene(sc, "f32")
    n = substeps or sc.substeps
    for f in range(frames):
        if sc.drive is not None:
            t, v = sc.drive.frame_target(f, n * sc.dt)
            w.set_target(sid, t, v)
        if plain:
            w.frame_scene(sid, n, splats)
        else:
            w.frame(n, splats, graph)
        o.frame(n, splats)
    pos, vel = w.read_nodes(sid)
    return w, sid, o, pos, vel


@pytest.mark.parametrize("plain", [False, True])
def test_c1_one_substep(plain):
    sc = scenes.c1_scene(splats=0)
    w, sid, o, pos, vel = step_both(sc, 1, substeps=1, plain=plain)
    print(f"one substep: exact {np.mean(pos == o.p):.6f}, rel {rel_err(pos, o.p):.3e}")
    assert rel_err(pos, o.p) <= RTOL
    assert rel_err(vel, o.v) <= RTOL


def test_c1_ten_frames_batched_graph():
    sc = scenes.c1_scene(splats=0)
    w, sid, o, pos, vel = step_both(sc, 10, graph=True)
    print(f"10 frames: exact {np.mean(pos == o.p):.6f}, rel {rel_err(pos, o.p):.3e}")
    assert rel_err(pos, o.p) <= RTOL
    assert rel_err(vel, o.v) <= 1e-4


def test_contact_keys_and_pairs_bit_exact():
    from oracle.oracle import OracleScene
    sc = scenes.contact_stack(4)
    w = world_for(sc)
    w.log_pairs(64)
    sid = w.add(sc)
    o = OracleScene(sc, "f32", pair_cap=64)
    for f in range(3):
        w.frame(sc.substeps)
        o.frame(sc.substeps)
        keys, counts, pairs = w.read_contacts(sid)
        assert np.array_equal(keys, o.keys)
        assert np.array_equal(counts, o.pair_count)
        for s in np.nonzero(counts)[0]:
            n = min(counts[s], 64)
            assert np.array_equal(pairs[s, :n], o.pairs[s, :n])
        pos, vel = w.read_nodes(sid)
        assert rel_err(pos, o.p) <= RTOL
    print(f"pairs in last substep: {int(counts.sum())}")
    assert counts.sum() > 0


@pytest.mark.parametrize("count,side,bucket", [(4, 6, 0.05), (6, 8, 0.06), (8, 8, 0.08), (8, 8, 0.25)])
def test_overfull_buckets_bit_exact(count, side, bucket):
    """Buckets several voxels wide hold up to 86-199 records (0.25 m: up to
    1271): each overfull bucket keeps its 32 lowest record ids, ascending
    (sort.wgsl:27-37). On the GPU sort takes them from the bucket's members
    (arrival slots and bin's overflow list); buckets of more than 512 records
    go to the wave-wide rescan of the scene's records. Keys, accepted pairs
    and positions against the oracle over three frames."""
    from oracle.oracle import OracleScene
    sc = scenes.contact_stack(count, side=side)
    sc.bucket_size = bucket
    w = world_for(sc)
    w.log_pairs(128)
    sid = w.add(sc)
    o = OracleScene(sc, "f32", pair_cap=128)
    overfull = 0
    for f in range(3):
        w.frame(sc.substeps, False, f > 0)
        o.frame(sc.substeps)
        keys, counts, pairs = w.read_contacts(sid)
        assert np.array_equal(keys, o.keys)
        assert np.array_equal(counts, o.pair_count)
        for s in np.nonzero(counts)[0]:
            n = min(counts[s], 128)
            assert np.array_equal(pairs[s, :n], o.pairs[s, :n])
        overfull += int((np.bincount(keys) > 32).sum())
        pos, vel = w.read_nodes(sid)
        assert rel_err(pos, o.p) <= RTOL and rel_err(vel, o.v) <= RTOL
    print(f"overfull buckets at frame ends: {overfull}; exact {np.mean(pos == o.p):.6f}")
    assert overfull > 0


def test_contact_many_bodies_one_substep():
    from oracle.oracle import OracleScene
    sc = scenes.c3_scene(bodies=64)
    w, sid, o, pos, vel = step_both(sc, 1, substeps=1)
    assert rel_err(pos, o.p) <= RTOL
    assert rel_err(vel, o.v) <= RTOL


@pytest.mark.parametrize("law", [0, 1])
def test_drive_handle_frames(law):
    sc = scenes.jelly_block(12, pinned=False, tank_margin=0.3, name="drag")
    sc.tank_min = (sc.tank_min[0], 0.0, sc.tank_min[2])
    sc.drive = scenes.handle_drive(sc, seed=3, law=law)
    w, sid, o, pos, vel = step_both(sc, 6)
    print(f"drive law {law}: exact {np.mean(pos == o.p):.6f}, rel {rel_err(pos, o.p):.3e}")
    assert rel_err(pos, o.p) <= RTOL
    held = sc.drive.nodes
    if law == 0:
        # hard law: held nodes sit exactly on their goals after each frame
        goal = np.asarray(sc.drive.frame_target(5, sc.substeps * sc.dt)[0], np.float32)[:, None] + sc.drive.offsets.T
        assert np.max(np.abs(pos[:, held] - goal)) < 1e-5


def test_stretch_limit_sweeps():
    sc = scenes.jelly_block(8, pinned=True, name="stretch")
    sc.bodies[0].material.max_stretch = 1.02
    sc.drive = scenes.handle_drive(sc, seed=5, law=0, amplitude=0.3, frequency=2.0)
    w, sid, o, pos, vel = step_both(sc, 4)
    assert rel_err(pos, o.p) <= RTOL


def test_splats_after_frames():
    sc = scenes.c1_scene(splats=50_000)
    w, sid, o, pos, vel = step_both(sc, 3, splats=True)
    mean, cov = w.read_splats(sid)
    print(f"splats: mean exact {np.mean(mean == o.out_mean):.6f}, cov exact {np.mean(cov == o.out_cov):.6f}")
    assert rel_err(mean, o.out_mean) <= RTOL
    assert rel_err(cov, o.out_cov) <= RTOL


def rebuilt_cov(rot, axis):
    # R diag(axis)^2 R^T per splat, in f64, as 6 planes (xx xy xz yy yz zz).
    x, y, z, w = (r.astype(np.float64) for r in rot)
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)]),
                  np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)]),
                  np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)])])
    s2 = axis.astype(np.float64) ** 2
    c = np.einsum("ikn,kn,jkn->ijn", R, s2, R)
    return np.stack([c[0, 0], c[0, 1], c[0, 2], c[1, 1], c[1, 2], c[2, 2]])


@pytest.mark.parametrize("graph", [False, True])
def test_splat_engine_axes(graph):
    # Engine format after three frames: rotations and axis lengths against the
    # f32 oracle's Jacobi, and the covariance they describe against the one
    # the same launch wrote.
    sc = scenes.c1_scene(splats=50_000)
    w, sid, o, pos, vel = step_both(sc, 3, splats=False)
    w.frame(0, native.SPLAT_COV | native.SPLAT_AXES, graph)
    o.deform_splats(axes=True)
    mean, cov = w.read_splats(sid)
    m2, rot, axis = w.read_splat_axes(sid)
    print(f"axes: rotation exact {np.mean(rot == o.out_rot):.6f}, lengths exact {np.mean(axis == o.out_axis):.6f}")
    assert np.array_equal(m2, mean)
    assert rel_err(cov, o.out_cov) <= RTOL
    assert np.max(np.abs(rot.astype(np.float64) - o.out_rot)) <= RTOL
    assert rel_err(axis, o.out_axis) <= RTOL
    assert np.max(np.abs(np.sum(rot.astype(np.float64) ** 2, axis=0) - 1)) < 1e-6
    back = rebuilt_cov(rot, axis)
    big = np.max(np.abs(cov), axis=0)
    assert np.max(np.max(np.abs(back - cov), axis=0) / big) < 1e-5


def test_splat_axes_only_leaves_covariance_untouched():
    sc = scenes.jelly_block(6, splats=500, seed=3, name="ax")
    w = world_for(sc)
    sid = w.add(sc)
    w.frame(1, native.SPLAT_COV, False)
    _, cov1 = w.read_splats(sid)
    w.frame(0, native.SPLAT_AXES, False)
    _, cov2 = w.read_splats(sid)
    _, rot, axis = w.read_splat_axes(sid)
    assert np.array_equal(cov1, cov2)
    assert np.all(np.isfinite(rot)) and np.all(axis >= 0)


def test_batched_equals_alone_and_graph_equals_eager():
    mix = [scenes.jelly_block(s, splats=300, seed=s, name=f"m{s}") for s in (5, 9, 13)]
    mix[1].drive = scenes.handle_drive(mix[1], seed=2)
    w = world_for(*mix)
    ids = [w.add(s) for s in mix]
    for f in range(4):
        for sid, sc in zip(ids, mix):
            if sc.drive is not None:
                w.set_target(sid, *sc.drive.frame_target(f, sc.substeps * sc.dt))
        w.frame(10, splats=True, graph=(f % 2 == 0))
    together = [w.read_nodes(i)[0] for i in ids]
    for sc, got in zip(mix, together):
        alone = world_for(sc)
        a = alone.add(sc)
        for f in range(4):
            if sc.drive is not None:
                alone.set_target(a, *sc.drive.frame_target(f, sc.substeps * sc.dt))
            alone.frame_scene(a, 10, True)
        assert np.array_equal(alone.read_nodes(a)[0], got)


def test_autotune_leaves_state_and_results_unchanged():
    """vs_world_autotune times every launch-shape candidate from the same
    snapshot of the node planes and puts the snapshot back: the scenes are
    where they were, and frames run with the shape it picks equal frames run
    with the default shape (a dense block, a sparse lattice, a contact scene
    with a stretch limit, a driven block)."""
    def mixed():
        mix = [scenes.jelly_block(s, splats=200, seed=s, name=f"t{s}") for s in (6, 11)]
        mix[1].drive = scenes.handle_drive(mix[1], seed=3)
        mix.append(_sparse_scene(5))
        c = scenes.c3_scene(bodies=8, seed=9)
        for b in c.bodies:
            b.material.max_stretch = 1.05
        mix.append(c)
        return mix

    out = []
    for tune in (False, True):
        mix = mixed()
        w = world_for(*mix)
        ids = [w.add(s) for s in mix]
        w.frame(4, splats=True, graph=False)
        before = [w.read_nodes(i) for i in ids]
        if tune:
            report = w.autotune(4)
            print(report[:200])
            after = [w.read_nodes(i) for i in ids]
            for (p, v), (p2, v2) in zip(before, after):
                assert np.array_equal(p, p2) and np.array_equal(v, v2)
        for f in range(3):
            for sid, sc in zip(ids, mix):
                if sc.drive is not None:
                    w.set_target(sid, *sc.drive.frame_target(f, 4 * sc.dt))
            w.frame(4, splats=True, graph=(f > 0))
        out.append([w.read_nodes(i)[0] for i in ids] + [w.read_splats(i)[1] for i in ids])
        w.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def serve(requests, max_nodes, realtime=False):
    nodes = max_nodes
    cells = max(s.n_cells for s, _ in requests) * 4
    w = native.World(max_scenes=8, max_bodies=8, max_nodes=nodes, max_cells=cells, max_driven=4096,
                     max_splats=1024, persistent=512)
    srv = native.Server(w, substeps=10, graph=True, keep_results=True)
    ids = [srv.submit(sc, 0.0, frames) for sc, frames in requests]
    rep = srv.run(realtime=realtime)
    return srv, w, ids, rep


def test_server_results_are_the_same_alone_and_in_the_mix():
    from oracle.oracle import OracleScene
    reqs = []
    for i, side in enumerate([6, 10, 8, 12, 7, 9]):
        sc = scenes.jelly_block(side, seed=40 + i, name=f"r{i}")
        if i % 2:
            sc.drive = scenes.handle_drive(sc, seed=i)
        reqs.append((sc, 3 + i % 3))
    budget = 2 * max(sc.n_nodes for sc, _ in reqs) + 16
    srv, w, ids, rep = serve(reqs, budget)
    print(rep)
    assert rep["completed"] == len(reqs) and rep["max_active"] >= 2
    assert rep["iterations"] < sum(f for _, f in reqs)   # frames of different scenes shared launches
    for (sc, frames), rid in zip(reqs, ids):
        mixed = srv.result(rid)
        alone_srv, _, alone_ids, _ = serve([(sc, frames)], budget)
        assert np.array_equal(alone_srv.result(alone_ids[0]), mixed)
        if sc.drive is None:
            o = OracleScene(sc, "f32")
            for _ in range(frames):
                o.frame(10)
            assert rel_err(mixed, o.p) <= RTOL


def _sparse_scene(seed, shape=(9, 12, 10), fill=0.65, pinned=False):
    from tests.test_gpu_lattice import occupied_scene
    return occupied_scene(np.random.default_rng(seed).random(shape) < fill, pinned=pinned, name=f"sp{seed}")


@pytest.mark.parametrize("case", ["c1", "odd_block", "sparse", "sparse_pinned", "sparse_device", "hard_drive",
                                  "compliant_drive", "contact", "stretch"])
def test_merged_node_passes_equal_separate_and_oracle(case):
    # A substep's finalize and the next one's predict as one node pass: the
    # same frames, bit for bit, as the two passes, and the oracle's frames.
    from oracle.oracle import OracleScene
    device = False
    if case == "c1":
        sc = scenes.c1_scene(splats=0)
    elif case == "odd_block":
        sc = scenes.jelly_block((5, 6, 7), seed=3, name="odd")
    elif case == "sparse":
        sc = _sparse_scene(11)
    elif case == "sparse_pinned":
        sc = _sparse_scene(12, shape=(8, 9, 14), pinned=True)
    elif case == "sparse_device":
        sc, device = _sparse_scene(13, shape=(11, 7, 9)), True
    elif case in ("hard_drive", "compliant_drive"):
        sc = scenes.jelly_block(10, seed=5, pinned=False, tank_margin=0.3, name="hd")
        sc.drive = scenes.handle_drive(sc, seed=6, law=0 if case == "hard_drive" else 1)
    elif case == "contact":
        sc = scenes.contact_stack(count=3, side=4)
    else:
        sc = scenes.jelly_block(6, seed=8, name="st")
        sc.bodies[0].material.max_stretch = 1.05
    out = []
    for separate in (0, 1):
        w = world_for(sc, separate_node_passes=separate)
        sid = w.add(sc, device_lattice=device)
        for f in range(3):
            if sc.drive is not None:
                w.set_target(sid, *sc.drive.frame_target(f, sc.substeps * sc.dt))
            w.frame(sc.substeps, False, f == 1)
        out.append(w.read_nodes(sid))
        prof = w.profile(sc.substeps)
        # merged: one node pass a substep after the first, or none when every
        # body's last colour pair finalizes and predicts all its nodes
        fp = prof["finalize+predict"][1]
        assert fp == 0 if separate else fp in (0, sc.substeps - 1)
        # merged: one frame-start predict, or none when the first colour pair
        # predicts every node (odd-sided dense lattices without drive)
        assert prof["predict"][1] == sc.substeps if separate else prof["predict"][1] <= 1
        w.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    if sc.drive is None:
        o = OracleScene(sc, "f32")
        for f in range(3):
            o.frame(sc.substeps)
        print(f"{case}: exact vs oracle {np.mean(out[0][0] == o.p):.6f}")
        assert rel_err(out[0][0], o.p) <= RTOL
        assert rel_err(out[0][1], o.v) <= RTOL


def _pair_case(case):
    if case == "c1":
        return [scenes.c1_scene(splats=0)]
    if case == "odd_block":
        return [scenes.jelly_block((5, 6, 7), seed=3, name="odd")]
    if case == "thin":
        # one cell along x (a row of one lane), odd and even y/z counts
        return [scenes.jelly_block((1, 3, 2), seed=4, name="thin"), scenes.jelly_block((2, 1, 5), seed=5, name="thin2")]
    if case == "wide_rows":
        # rows of 36 and 66 lanes: the second does not fit a 64-lane
        # workgroup and runs per colour beside the first
        return [scenes.jelly_block((71, 3, 4), h=0.004, seed=6, name="w71"),
                scenes.jelly_block((130, 2, 3), h=0.002, seed=7, name="w130")]
    if case == "short_rows":
        # rows of two and three lanes, 128 and 85 of them to a 256-lane
        # workgroup
        return [scenes.jelly_block((2, 40, 40), h=0.01, seed=14, name="s2"),
                scenes.jelly_block((4, 30, 30), h=0.01, seed=15, name="s4")]
    if case == "even_rows":
        # even rows (a lane at each row's end holds its last node alone)
        # in the folded last pair and the predicting first pair, one block
        # pinned, one falling onto the tank floor
        return [scenes.jelly_block((16, 5, 6), seed=16, name="e16"),
                scenes.jelly_block((10, 7, 4), seed=17, pinned=False, tank_margin=0.01, name="e10")]
    if case == "contact":
        return [scenes.contact_stack(count=3, side=4)]
    if case == "stretch":
        sc = scenes.jelly_block((6, 5, 7), seed=8, name="st")
        sc.bodies[0].material.max_stretch = 1.05
        return [sc]
    if case == "hard_drive":
        sc = scenes.jelly_block(10, seed=5, pinned=False, tank_margin=0.3, name="hd")
        sc.drive = scenes.handle_drive(sc, seed=6, law=0)
        return [sc]
    if case == "mixed":
        # dense blocks (pairs) and a sparse lattice (per colour) in one batch
        return [scenes.jelly_block((9, 8, 7), seed=9, name="m1"), _sparse_scene(21),
                scenes.jelly_block(12, seed=10, name="m2")]
    if case == "mixed_fold":
        # scenes whose last colour pair finalizes and predicts its nodes
        # beside scenes with contact or tensile sweeps after the colours
        stack = scenes.contact_stack(count=2, side=5)
        stack.substeps = 10
        st = scenes.jelly_block((7, 4, 6), seed=12, name="st2")
        st.bodies[0].material.max_stretch = 1.02
        return [scenes.jelly_block((9, 8, 7), seed=11, name="f1"), stack, st,
                scenes.jelly_block((4, 9, 3), seed=13, pinned=False, name="f2")]
    raise ValueError(case)


@pytest.mark.parametrize("block", [64, 256])
@pytest.mark.parametrize("case", ["c1", "odd_block", "thin", "wide_rows", "short_rows", "even_rows", "contact", "stretch",
                                  "hard_drive", "mixed", "mixed_fold"])
def test_colour_pairs_equal_separate_colours_and_oracle(case, block):
    # Colours 2c and 2c+1 of a dense lattice solved by rows in one launch:
    # the same frames, bit for bit, as one launch per colour, batched, from a
    # graph and plain; and within tolerance of the oracle.
    from oracle.oracle import OracleScene
    scs = _pair_case(case)
    print(f"case {case} block {block}: dims {[b.lattice.cell_dims for s in scs for b in s.bodies]}", flush=True)
    results = {}
    for separate in (1, 0):
        for plain in (False, True):
            w = world_for(*scs, separate_colours=separate, block_cells=block)
            sids = [w.add(s) for s in scs]
            for f in range(2):
                for s, sid in zip(scs, sids):
                    if s.drive is not None:
                        w.set_target(sid, *s.drive.frame_target(f, s.substeps * s.dt))
                if plain:
                    for s, sid in zip(scs, sids):
                        w.frame_scene(sid, s.substeps, False)
                else:
                    w.frame(scs[0].substeps, False, f == 1)
            results[(separate, plain)] = [w.read_nodes(sid) for sid in sids]
            w.close()
    ref = results[(1, False)]
    for key, got in results.items():
        for (p, v), (rp, rv) in zip(got, ref):
            assert np.array_equal(p, rp) and np.array_equal(v, rv), key
    for s, (p, v) in zip(scs, results[(0, False)]):
        if s.drive is not None:
            continue
        o = OracleScene(s, "f32")
        for f in range(2):
            o.frame(s.substeps)
        print(f"{case}/{s.name}: exact vs oracle {np.mean(p == o.p):.6f}")
        assert rel_err(p, o.p) <= RTOL
        assert rel_err(v, o.v) <= RTOL


def _sphere_scene(drive):
    sc = scenes.c2_scene(side=12, splats=2000, seed=7, sphere=True)
    if not drive:
        sc.drive = None
    return sc


@pytest.mark.parametrize("drive", [False, True])
@pytest.mark.parametrize("schedule", ["pairs", "separate_colours", "separate_node_passes", "plain"])
def test_sphere_collider_frames(drive, schedule):
    """A block draped over the sphere collider (its cap two cells into the
    block's underside), with and without the drag handle, through every
    schedule: the folded node passes of the colour pairs, one colour per
    launch, separate finalize and predict, and the plain per-scene path."""
    sc = _sphere_scene(drive)
    kw = {"separate_colours": int(schedule == "separate_colours"),
          "separate_node_passes": int(schedule == "separate_node_passes")}
    w, sid, o, pos, vel = step_both(sc, 4, graph=schedule == "pairs", plain=schedule == "plain", splats=True, **kw)
    print(f"sphere {schedule} drive={drive}: exact {np.mean(pos == o.p):.6f}, rel {rel_err(pos, o.p):.3e}")
    assert rel_err(pos, o.p) <= RTOL and rel_err(vel, o.v) <= RTOL
    mean, cov = w.read_splats(sid)
    assert rel_err(mean, o.out_mean) <= RTOL and rel_err(cov, o.out_cov) <= RTOL
    c = np.asarray(sc.sphere[:3], np.float64)[:, None]
    assert np.all(np.linalg.norm(pos.astype(np.float64) - c, axis=0) >= sc.sphere[3] * (1 - 1e-5))


def test_poisson_stream_results_equal_each_request_served_alone():
    """A shortened serving stream (Poisson arrivals at 50/s, seeded drag
    handles, 4k-60k nodes, 20 frames each) served in real time: requests
    that arrive while others run are admitted into the running batch, and
    each result is bit-identical to the same request served alone."""
    reqs = scenes.c5_requests(16, max_nodes=60_000, frames=20)
    budget = 160_000
    cells = int(budget * 1.05)
    w = native.World(max_scenes=16, max_bodies=16, max_nodes=budget, max_cells=cells, max_driven=16 * 27,
                     max_splats=1024, persistent=512)
    srv = native.Server(w, substeps=10, graph=True, keep_results=True)
    ids = [srv.submit(sc, t, frames) for t, sc, frames in reqs]
    rep = srv.run(realtime=True)
    print(rep)
    assert rep["completed"] == len(reqs) and rep["max_active"] >= 2
    for (t, sc, frames), rid in zip(reqs, ids):
        alone_srv, _, alone_ids, _ = serve([(sc, frames)], sc.n_nodes + 16)
        assert np.array_equal(alone_srv.result(alone_ids[0]), srv.result(rid)), sc.name
        alone_srv.close()
    srv.close()
    w.close()
