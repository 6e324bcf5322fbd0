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
