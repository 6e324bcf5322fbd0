"""Parity at the workloads' full sizes.

* 256 scenes of lattice side 8..48 (9.8 M nodes, 38 M splats): the batched,
  graph-replayed frame equals the plain per-scene frame bit for bit, nodes
  and splats (a size-independent property of the schedule: colour pairs,
  folded node passes and node tiles against one scene's slice of the same
  tables).
* 64^3 block with 2 M splats and a 27-node drag handle, and 512 bodies of 8^3
  in contact: one frame against the f32 oracle, positions within 1e-5 of the
  largest coordinate (the figures printed are the exact-match fraction), and
  the contact grid's bucket keys and accepted pairs bit for bit.
"""
import numpy as np
import pytest

from vsim import native, scenes

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def world_for(scs, **kw):
    nodes = sum(s.n_nodes for s in scs)
    cells = sum(s.n_cells for s in scs)
    splats = sum(s.n_splats for s in scs)
    surf = sum(s.n_surface_active for s in scs)
    buckets = sum(int(np.prod(s.grid()[2])) for s in scs if s.contact_active)
    driven = sum(len(s.drive.nodes) for s in scs if s.drive is not None)
    return native.World(max_scenes=len(scs) + 1, max_bodies=sum(len(s.bodies) for s in scs) + 1,
                        max_nodes=nodes + 16, max_cells=cells + 16, max_surface=surf + 16, max_buckets=buckets + 16,
                        max_driven=driven + 16, max_splats=splats + 16, **kw)


def rel_err(gpu, ref):
    scale = max(float(np.max(np.abs(ref))), 1e-30)
    return float(np.max(np.abs(gpu.astype(np.float64) - ref.astype(np.float64)))) / scale


def test_256_scenes_batched_equals_plain():
    scs = scenes.c4_scenes(256)
    print(f"256 scenes: {sum(s.n_nodes for s in scs)} nodes, {sum(s.n_splats for s in scs)} splats", flush=True)
    out = []
    for batched in (True, False):
        w = world_for(scs)
        ids = [w.add(s) for s in scs]
        for f in range(2):
            if batched:
                w.frame(scs[0].substeps, splats=True, graph=True)
            else:
                for sid in ids:
                    w.frame_scene(sid, scs[0].substeps, True)
        out.append([(w.read_nodes(sid), w.read_splats(sid)) for sid in ids])
        w.close()
    for ((p, v), (m, c)), ((p2, v2), (m2, c2)) in zip(*out):
        assert np.array_equal(p, p2) and np.array_equal(v, v2)
        assert np.array_equal(m, m2) and np.array_equal(c, c2)


@pytest.mark.parametrize("case", ["block64_drag", "bodies512_contact"])
def test_full_size_frame_against_oracle(case):
    from oracle.oracle import OracleScene
    sc = scenes.c2_scene() if case == "block64_drag" else scenes.c3_scene()
    print(f"{case}: {sc.n_nodes} nodes, {sc.n_splats} splats, {sc.n_surface_active} contact records, "
          f"{sc.substeps} substeps", flush=True)
    cap = 8 if sc.contact_active else 0
    w = world_for([sc])
    if cap:
        w.log_pairs(cap)
    sid = w.add(sc)
    o = OracleScene(sc, "f32", pair_cap=cap)
    if sc.drive is not None:
        t, v = sc.drive.frame_target(0, sc.substeps * sc.dt)
        w.set_target(sid, t, v)
        o.set_target(t, v)
    w.frame(sc.substeps, splats=sc.n_splats > 0, graph=True)
    o.frame(sc.substeps, splats=sc.n_splats > 0)
    pos, vel = w.read_nodes(sid)
    print(f"{case}: exact {np.mean(pos == o.p):.6f}, rel {rel_err(pos, o.p):.2e}")
    assert rel_err(pos, o.p) <= RTOL and rel_err(vel, o.v) <= RTOL
    if sc.n_splats:
        mean, cov = w.read_splats(sid)
        assert rel_err(mean, o.out_mean) <= RTOL and rel_err(cov, o.out_cov) <= RTOL
    if cap:
        # the last substep's bucket keys and accepted pairs (up to `cap` logged)
        keys, counts, pairs = w.read_contacts(sid)
        assert np.array_equal(keys, o.keys)
        assert np.array_equal(counts, o.pair_count)
        for s in np.nonzero(counts)[0]:
            k = min(counts[s], cap)
            assert np.array_equal(pairs[s, :k], o.pairs[s, :k])
        print(f"{case}: {int(counts.sum())} pairs in the last substep")
    w.close()
