"""The eight colours of a small dense body in one launch, the body in LDS
(k_elastic_bodies), against the same bodies through the colour-pair launches
(`colour_launches=1`) bit for bit, and against the f32 oracle: contact
bodies of 8^3 cells, a dragged 10^3 block (11^3 nodes, the largest such
body, 125 cells a colour), an 11^3
block beside it that is too large and keeps the colour pairs, tensile-limit
sweeps, and a folded scene (no contact, no drive) whose bodies keep the
colour pairs, all in one world; batched from a graph and per scene. Then a
contact scene whose sweep carries the settle and the bin, over frames of an
odd substep count."""
import numpy as np
import pytest

from vsim import scenes

from test_gpu_parity import RTOL, rel_err, world_for

pytestmark = pytest.mark.gpu


def _mix():
    bodies = scenes.c3_scene(bodies=27, box=0.3, seed=5)
    drag = scenes.jelly_block(10, seed=5, pinned=False, tank_margin=0.3, name="drag10")
    drag.drive = scenes.handle_drive(drag, seed=6, law=0)
    big = scenes.jelly_block(11, seed=7, pinned=False, tank_margin=0.3, name="drag11")
    big.drive = scenes.handle_drive(big, seed=8, law=1)
    st = scenes.jelly_block((6, 5, 7), seed=8, name="st")
    st.bodies[0].material.max_stretch = 1.05
    fold = scenes.jelly_block(9, seed=9, name="fold")
    return [bodies, drag, big, st, fold]


def _run(scs, colour_launches, plain, frames, n=8):
    w = world_for(*scs, colour_launches=colour_launches)
    sids = [w.add(s) for s in scs]
    for f in range(frames):
        for s, sid in zip(scs, sids):
            if s.drive is not None:
                w.set_target(sid, *s.drive.frame_target(f, n * s.dt))
        if plain:
            for sid in sids:
                w.frame_scene(sid, n, False)
        else:
            w.frame(n, False, True)
    out = [w.read_nodes(sid) for sid in sids]
    w.close()
    return out


@pytest.mark.parametrize("plain", [False, True])
def test_body_sweep_equals_colour_launches_and_oracle(plain):
    from oracle.oracle import OracleScene
    scs = _mix()
    frames = 2
    swept = _run(scs, 0, plain, frames)
    pairs = _run(scs, 1, plain, frames)
    for s, (p, v), (rp, rv) in zip(scs, swept, pairs):
        assert np.array_equal(p, rp) and np.array_equal(v, rv), s.name
    for s, (p, v) in zip(scs, swept):
        o = OracleScene(s, "f32")
        for f in range(frames):
            if s.drive is not None:
                o.set_target(*s.drive.frame_target(f, 8 * s.dt))
            o.frame(8)
        print(f"{s.name}: exact vs oracle {np.mean(p == o.p):.6f}")
        assert rel_err(p, o.p) <= RTOL and rel_err(v, o.v) <= RTOL


@pytest.mark.parametrize("plain", [False, True])
def test_settle_and_bin_on_the_sweep_over_odd_substep_counts(plain):
    """A contact scene whose bodies all take the body sweep, beside a scene
    that keeps the colour pairs: the sweep carries the previous substep's
    settle and this substep's bin, and substep s bins into counter set s & 1.
    Five substeps a frame, so frames start on either set. Positions,
    velocities and the last substep's bucket keys bit for bit the colour-pair
    launches' (separate bin and settle launches), and the f32 oracle's; the
    oracle logs contact pairs, so the bodies did meet."""
    from oracle.oracle import OracleScene
    scs = [scenes.c3_scene(bodies=27, box=0.3, seed=11), scenes.jelly_block(9, seed=12, name="fold")]
    frames, n = 4, 5

    def run(colour_launches):
        w = world_for(*scs, colour_launches=colour_launches)
        sids = [w.add(s) for s in scs]
        for f in range(frames):
            if plain:
                for sid in sids:
                    w.frame_scene(sid, n, False)
            else:
                w.frame(n, False, True)
        out = [w.read_nodes(sid) for sid in sids]
        keys = w.read_contacts(sids[0])[0]
        w.close()
        return out, keys

    swept, swept_keys = run(0)
    pairs, pair_keys = run(1)
    assert np.array_equal(swept_keys, pair_keys)
    for s, (p, v), (rp, rv) in zip(scs, swept, pairs):
        assert np.array_equal(p, rp) and np.array_equal(v, rv), s.name
    met = 0
    for s, (p, v) in zip(scs, swept):
        o = OracleScene(s, "f32", pair_cap=4 if s.contact_active else 0)
        for f in range(frames):
            o.frame(n)
            if s.contact_active:
                met = max(met, int(np.sum(o.pair_count)))
        print(f"{s.name}: exact vs oracle {np.mean(p == o.p):.6f}")
        assert rel_err(p, o.p) <= RTOL and rel_err(v, o.v) <= RTOL
    assert met > 0
