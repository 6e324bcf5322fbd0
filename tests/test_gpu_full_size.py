"""Parity at the workloads' full sizes.

* 256 scenes of lattice side 8..48 (9.8 M nodes, 38 M splats): the batched,
  graph-replayed frame equals the plain per-scene frame bit for bit, nodes
  and splats (a size-independent property of the schedule: colour pairs,
  folded node passes and node tiles against one scene's slice of the same
  tables).
* 64^3 block with 2 M splats and a 27-node drag handle, and 512 bodies of 8^3
  in contact (buckets of 1 and of 2 voxels): one frame against the f32
  oracle, positions within 1e-5 of the largest coordinate (the figures
  printed are the exact-match fraction), and the contact grid's bucket keys
  and accepted pairs bit for bit.
* the 512 bodies piled up after 80 frames: one more frame against the oracle
  from the GPU's state, with about a hundred overfull buckets.
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


FULL_SIZE = {
    "block64_drag": scenes.c2_scene,
    "bodies512_contact": scenes.c3_scene,
    # the configuration's hash grid: buckets of 2 voxels, up to 64 records
    # at worst against 32 slots (broadphase.js:82-85)
    "bodies512_contact_2voxel": lambda: scenes.c3_scene(bucket_voxels=2.0),
}


def check_contacts(w, sid, o, cap, case):
    """The last substep's bucket keys and accepted pairs (up to `cap`
    logged) against the oracle's."""
    keys, counts, pairs = w.read_contacts(sid)
    assert np.array_equal(keys, o.keys)
    assert np.array_equal(counts, o.pair_count)
    for s in np.nonzero(counts)[0]:
        k = min(counts[s], cap)
        assert np.array_equal(pairs[s, :k], o.pairs[s, :k])
    overfull = int((np.bincount(keys) > 32).sum())
    print(f"{case}: {int(counts.sum())} pairs in the last substep, {overfull} overfull buckets")
    return overfull


@pytest.mark.parametrize("case", sorted(FULL_SIZE))
def test_full_size_frame_against_oracle(case):
    from oracle.oracle import OracleScene
    sc = FULL_SIZE[case]()
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
        check_contacts(w, sid, o, cap, case)
    w.close()


@pytest.mark.parametrize("bucket_voxels", [1.0, 2.0])
def test_piled_contact_frame_against_oracle(bucket_voxels):
    """The 512 bodies after 80 frames on the GPU, fallen into a pile whose
    crowded buckets hold far more than 32 records: the GPU's state is loaded
    into the oracle and one more frame is compared (positions, velocities,
    the last substep's keys and pairs). The overfull buckets keep their 32
    lowest record ids (sort.wgsl:27-37, bodyContact.js:87-107)."""
    from oracle.oracle import OracleScene
    sc = scenes.c3_scene(bucket_voxels=bucket_voxels)
    cap = 8
    w = world_for([sc])
    w.log_pairs(cap)
    sid = w.add(sc)
    for f in range(80):
        w.frame(sc.substeps, False, True)
    pos, vel = w.read_nodes(sid)
    print(f"piled ({bucket_voxels:g}-voxel buckets): state after 80 frames read back", flush=True)
    o = OracleScene(sc, "f32", pair_cap=cap)
    o.p[:] = pos
    o.q[:] = pos
    o.v[:] = vel
    w.frame(sc.substeps, False, True)
    o.frame(sc.substeps)
    pos, vel = w.read_nodes(sid)
    print(f"piled: exact {np.mean(pos == o.p):.6f}, rel {rel_err(pos, o.p):.2e}")
    assert rel_err(pos, o.p) <= RTOL and rel_err(vel, o.v) <= RTOL
    assert check_contacts(w, sid, o, cap, f"piled {bucket_voxels:g}-voxel") > 0
    w.close()
