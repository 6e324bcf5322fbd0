"""Batched frames as two chains of launches side by side (`branches=2`: the
scenes split by work into two groups, each group's launches on a stream of
its own, captured as two branches of the frame's graph) against one chain,
bit for bit, and against the f32 oracle: blocks of several sizes with
splats, a dragged block, tensile sweeps, and two contact scenes (which share
the grid's overflow lists and so go to the same chain); graph-replayed and
eager frames; and the same world served by the continuous-batching server."""
import numpy as np
import pytest

from vsim import native, scenes

from test_gpu_parity import RTOL, rel_err, world_for

pytestmark = pytest.mark.gpu


def _mix():
    out = [scenes.jelly_block(s, splats=300, seed=40 + s, name=f"b{s}") for s in (5, 9, 13, 17, 7)]
    drag = scenes.jelly_block(10, seed=5, pinned=False, tank_margin=0.3, name="drag")
    drag.drive = scenes.handle_drive(drag, seed=6, law=0)
    st = scenes.jelly_block((6, 5, 7), seed=8, name="st")
    st.bodies[0].material.max_stretch = 1.05
    stack = scenes.contact_stack(4)
    over = scenes.contact_stack(6, side=6, seed=4)
    over.bucket_size = 0.06
    return out + [drag, st, stack, over]


def _frames(scs, branches, graph, frames=3):
    w = world_for(*scs, branches=branches)
    sids = [w.add(s) for s in scs]
    n = 8
    for f in range(frames):
        for s, sid in zip(scs, sids):
            if s.drive is not None:
                w.set_target(sid, *s.drive.frame_target(f, n * s.dt))
        w.frame(n, True, graph)
    out = [(w.read_nodes(sid), w.read_splats(sid)) for sid in sids]
    w.close()
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_two_chains_equal_one_chain_and_oracle(graph):
    from oracle.oracle import OracleScene
    scs = _mix()
    one = _frames(scs, 1, graph)
    two = _frames(scs, 2, graph)
    for s, ((p1, v1), (m1, c1)), ((p2, v2), (m2, c2)) in zip(scs, one, two):
        assert np.array_equal(p1, p2) and np.array_equal(v1, v2), s.name
        assert np.array_equal(m1, m2) and np.array_equal(c1, c2), s.name
    for s, ((p, v), _) in zip(scs, one):
        o = OracleScene(s, "f32")
        for f in range(3):
            if s.drive is not None:
                o.set_target(*s.drive.frame_target(f, 8 * s.dt))
            o.frame(8, True)
        assert rel_err(p, o.p) <= RTOL and rel_err(v, o.v) <= RTOL, s.name


def test_served_with_two_chains_equals_one_chain():
    reqs = [(scenes.jelly_block(side, seed=60 + side, name=f"r{side}"), 4 + side % 3) for side in (6, 11, 8, 14, 5, 9)]
    nodes = 3 * max(s.n_nodes for s, _ in reqs)
    cells = 3 * max(s.n_cells for s, _ in reqs)
    results = []
    for branches in (1, 2):
        w = native.World(max_scenes=8, max_bodies=8, max_nodes=nodes, max_cells=cells, max_driven=64,
                         max_splats=1024, branches=branches)
        srv = native.Server(w, substeps=8, graph=True, keep_results=True)
        ids = [srv.submit(sc, 0.0, frames) for sc, frames in reqs]
        rep = srv.run(realtime=False)
        assert rep["completed"] == len(reqs) and rep["max_active"] >= 2
        results.append([srv.result(rid) for rid in ids])
        srv.close()
        w.close()
    for a, b in zip(results[0], results[1]):
        assert np.array_equal(a, b)
