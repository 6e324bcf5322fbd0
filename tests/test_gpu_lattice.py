"""Lattices built on the device from occupancy grids (buildLattice and
surfaceNodes, lattice.js:128-239) against the host builder the lattice tests
pin: rest positions, inverse masses (static and hard-held nodes zeroed),
colour-grouped cells, colour counts and surface records must be identical, and
so must the simulation that runs on them."""
import numpy as np
import pytest

from vsim import native, scenes

pytestmark = pytest.mark.gpu


def world_for(*scs):
    nodes = sum(s.n_nodes for s in scs)
    cells = sum(s.n_cells for s in scs)
    surf = sum(s.n_surface_active for s in scs)
    buckets = sum(int(np.prod(s.grid()[2])) for s in scs if s.contact_active)
    driven = sum(len(s.drive.nodes) for s in scs if s.drive is not None)
    splats = sum(s.n_splats for s in scs)
    return native.World(max_scenes=8, max_bodies=16, max_nodes=nodes + 16, max_cells=cells + 16,
                        max_surface=surf + 16, max_buckets=buckets + 16, max_driven=driven + 16,
                        max_splats=splats + 16)


def occupied_scene(occ, h=0.012, origin=(0.1, -0.05, 0.2), pinned=True, contact=0, name="occ", splats=0, seed=0):
    mat = scenes.Material(**scenes.C1_MATERIAL)
    lat = scenes.build_lattice(occ, origin, h, mat.density)
    static = None
    if pinned:
        nx, ny, nz = lat.node_dims
        static = np.nonzero(np.tile(np.repeat(np.arange(ny), nx), nz) == 0)[0].astype(np.uint32)
    body = scenes.Body(lat, mat, static)
    if splats:
        body.splat_mean, body.splat_quat, body.splat_scale = scenes.seed_splats(np.random.default_rng(seed), lat, splats)
    cz, cy, cx = occ.shape
    tmin = tuple(origin[a] - 0.1 for a in range(3))
    tmax = tuple(origin[a] + (cx, cy, cz)[a] * h + 0.1 for a in range(3))
    return scenes.Scene([body], dt=float(np.float32(1 / 600)), tank_min=tmin, tank_max=tmax, contact=contact, name=name)


def sphere_shell(n, thick=3):
    c = (n - 1) / 2
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    return (r <= c) & (r >= c - thick)


CASES = {
    "cube1": lambda: occupied_scene(np.ones((1, 1, 1), bool)),
    "block_2x3x5": lambda: occupied_scene(np.ones((5, 3, 2), bool)),
    "block_7x8x9": lambda: occupied_scene(np.ones((9, 8, 7), bool)),
    "random_60pct": lambda: occupied_scene(np.random.default_rng(5).random((13, 10, 11)) < 0.6),
    "random_sparse_big": lambda: occupied_scene(np.random.default_rng(6).random((33, 17, 20)) < 0.3, pinned=False),
    "sphere_shell": lambda: occupied_scene(sphere_shell(24), pinned=False, contact=2, splats=2000, seed=3),
}


def check_same_lattice(w, sid_host, sid_dev, sc):
    arrays = sc.arrays()
    off = 0
    for b, body in enumerate(sc.bodies):
        lat = body.lattice
        rest_h, w_h, cells_h, surf_h, info_h = w.read_lattice(sid_host, b)
        rest_d, w_d, cells_d, surf_d, info_d = w.read_lattice(sid_dev, b)
        assert info_d == info_h
        assert info_d["color_count"] == [int(v) for v in lat.color_counts]
        assert np.array_equal(rest_d, rest_h) and np.array_equal(rest_d, lat.rest.T)
        assert np.array_equal(w_d, w_h)
        assert np.array_equal(w_d, arrays["w"][off:off + lat.node_count])
        assert np.array_equal(cells_d, cells_h) and np.array_equal(cells_d, lat.cells)
        assert np.array_equal(surf_d, surf_h)
        if sc.contact:
            assert np.array_equal(surf_d, scenes.surface_nodes(lat))
        off += lat.node_count


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_lattice_equals_host_lattice(case):
    sc = CASES[case]()
    w = world_for(sc, sc)
    a = w.add(sc)
    b = w.add(sc, device_lattice=True)
    print(f"{case}: {sc.n_nodes} nodes, {sc.n_cells} cells")
    check_same_lattice(w, a, b, sc)
    for f in range(3):
        w.frame(10, splats=bool(sc.n_splats), graph=f == 1)
    pa, va = w.read_nodes(a)
    pb, vb = w.read_nodes(b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    if sc.n_splats:
        assert all(np.array_equal(x, y) for x, y in zip(w.read_splats(a), w.read_splats(b)))


def test_device_lattice_zeroes_hard_held_nodes_and_contact_records():
    # Two bodies, contact on, a hard drive on body 0: the held nodes lose their
    # mass and both bodies' surfaces become records, body by body.
    s0 = scenes.jelly_block(6, seed=1, name="a")
    s1 = scenes.jelly_block((5, 4, 7), seed=2, pinned=False, origin=(0.2, 0.0, 0.0), name="b")
    sc = scenes.Scene(s0.bodies + s1.bodies, dt=s0.dt, tank_min=(-0.2, -0.2, -0.2), tank_max=(0.5, 0.5, 0.5),
                      contact=2, name="pair")
    sc.drive = scenes.handle_drive(sc, seed=7, law=0)
    w = world_for(sc, sc)
    a = w.add(sc)
    b = w.add(sc, device_lattice=True)
    check_same_lattice(w, a, b, sc)
    held = sc.drive.nodes
    assert np.all(w.read_lattice(b, 0)[1][held] == 0)
    for f in range(2):
        for sid in (a, b):
            w.set_target(sid, *sc.drive.frame_target(f, sc.substeps * sc.dt))
        w.frame(10)
    assert np.array_equal(w.read_nodes(a)[0], w.read_nodes(b)[0])
    ka = w.read_contacts(a)
    kb = w.read_contacts(b)
    assert np.array_equal(ka[0], kb[0])


def test_server_admits_device_built_scenes():
    reqs = [(occupied_scene(np.random.default_rng(40 + i).random((7 + i, 6, 8)) < 0.7, name=f"q{i}"), 4)
            for i in range(4)]
    nodes = sum(s.n_nodes for s, _ in reqs)
    results = []
    for device in (False, True):
        w = native.World(max_scenes=8, max_bodies=8, max_nodes=nodes, max_cells=nodes, max_driven=64,
                         max_splats=64, persistent=256)
        srv = native.Server(w, substeps=10, graph=True, keep_results=True)
        ids = [srv.submit(sc, 0.0, f, device_lattice=device) for sc, f in reqs]
        srv.run(realtime=False)
        results.append([srv.result(i) for i in ids])
        srv.close()
        w.close()
    for x, y in zip(*results):
        assert np.array_equal(x, y)
