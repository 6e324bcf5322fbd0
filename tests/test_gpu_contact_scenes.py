"""Several contact scenes in one world. The world's grid, its overflow lists
and its surface records are shared by every scene (each scene a range of
them), so a batched bin/sort/collide/settle launch works on all of them at
once. Three scenes, one with overfull buckets and one of 27 bodies in a
packed tank, stepped as graph-replayed batched frames and as per-scene
frames: bucket keys, accepted pairs (as each scene's own record numbers) and
positions against the f32 oracle of each scene alone."""
import numpy as np
import pytest

from vsim import scenes

from test_gpu_parity import RTOL, rel_err, world_for

pytestmark = pytest.mark.gpu

CAP = 128


def _scenes():
    a = scenes.contact_stack(4)
    b = scenes.contact_stack(6, side=6, seed=4)
    b.bucket_size = 0.06  # buckets of 6 voxels: overfull
    c = scenes.c3_scene(bodies=27, box=0.3, bucket_voxels=2, seed=5)
    return a, b, c


@pytest.mark.parametrize("plain", [False, True])
def test_contact_scenes_batched_and_plain_equal_oracle(plain):
    from oracle.oracle import OracleScene
    scs = _scenes()
    w = world_for(*scs)
    w.log_pairs(CAP)
    ids = [w.add(s) for s in scs]
    os_ = [OracleScene(s, "f32", pair_cap=CAP) for s in scs]
    n = 8
    pairs_seen = [0, 0, 0]
    overfull = 0
    for f in range(2):
        if plain:
            for sid in ids:
                w.frame_scene(sid, n, False)
        else:
            w.frame(n, False, True)
        for o in os_:
            o.frame(n)
        for k, (sid, o) in enumerate(zip(ids, os_)):
            keys, counts, pairs = w.read_contacts(sid)
            pos, vel = w.read_nodes(sid)
            assert np.array_equal(keys, o.keys)
            assert np.array_equal(counts, o.pair_count)
            for s in np.nonzero(counts)[0]:
                m = min(counts[s], CAP)
                assert np.array_equal(pairs[s, :m], o.pairs[s, :m])
            assert rel_err(pos, o.p) <= RTOL and rel_err(vel, o.v) <= RTOL
            pairs_seen[k] += int(counts.sum())
            overfull += int((np.bincount(keys) > 32).sum())
    print(f"pairs per scene {pairs_seen}, overfull buckets {overfull}")
    assert all(p > 0 for p in pairs_seen) and overfull > 0
    w.close()
