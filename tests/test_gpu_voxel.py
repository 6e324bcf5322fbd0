"""The device voxelisers (voxelizeCloud, voxelizeMesh, voxelizeSdf) through the
C ABI against the oracle's restatement (oracle/voxel.c), which
tests/test_oracle_voxel.py pins on the reference's voxeliser tests. Both
evaluate the JS's f64 expressions in the same order, so the grids must be
identical: dims, origin, cell edge and every byte. The last test runs a
capture end to end: PLY bytes, decoded cloud, device grid, device lattice,
frames, against the host-built lattice."""
import math

import numpy as np
import pytest

from _geometry import box_mesh, cloud_at, icosphere, sphere_shell
from oracle import oracle as O
from vsim import ingest, native, scenes

pytestmark = pytest.mark.gpu


def same_grid(dev, ref):
    assert dev.dims == ref.dims
    assert dev.origin == ref.origin and dev.h == ref.h
    assert np.array_equal(dev.cells, ref.occupancy), int((dev.cells != ref.occupancy).sum())


def random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 3] < 0] *= -1
    return q.astype(np.float32)


# ---------------------------------------------------------------- shapes
@pytest.mark.parametrize("kind,size,h,offset,padding", [
    ("sphere", [0.1], 0.2 / 32, (0, 0, 0), 2),
    ("sphere", [0.037], 0.0041, (0.3, -0.21, 1.7), 0),
    ("box", [0.1, 0.05, 0.075], 0.025, (0, 0, 0), 2),
    ("box", [0.1, 0.05, 0.075], 0.025, (0.3, -0.2, 0.05), 3),
    ("box", [0.21, 0.013, 0.4], 0.0117, (-1, 2, 0.5), 1),
    ("torus", [0.08, 0.03], 0.22 / 28, (0, 0, 0), 2),
    ("torus", [0.3, 0.11], 0.006, (0.01, 0.02, -0.03), 2),
    ("torus", [0.3, 0.11], 0.0071, [(0.01, 0.02, -0.03), (0.3, -0.7, 0.11), (-0.05, 0.0, 0.9)], 2),
    ("box", [0.2, 0.1, 0.3], 0.013, [(0.1, 0.2, 0.3), (0.7, -0.1, 0.05)], 2),
])
def test_shape_equals_oracle(kind, size, h, offset, padding):
    dev = ingest.voxelize_shape(kind, size, h, offset=offset, padding=padding)
    ref = O.voxelize_shape(kind, size, h, offset=offset, padding=padding)
    print(kind, dev.dims, dev.count())
    same_grid(dev, ref)
    assert dev.count() > 0


# ---------------------------------------------------------------- meshes
def transformed(p, seed):
    rng = np.random.default_rng(seed)
    q = random_rotations(rng, 1)[0].astype(np.float64)
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    return (p.astype(np.float64) @ R.T + rng.normal(size=3)).astype(np.float32)


def soup(seed, n=400):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-0.3, 0.3, size=(3 * n, 3)).astype(np.float32)
    return p, np.arange(3 * n, dtype=np.uint32)


MESHES = {
    "box": lambda: (*box_mesh(0.5), 1.0 / 20, {}),
    "box_open_hole": lambda: (box_mesh(0.5)[0], np.concatenate([box_mesh(0.5)[1][:12], box_mesh(0.5)[1][18:]]),
                              1.0 / 16, {}),
    "box_no_close": lambda: (*box_mesh(0.5), 1.0 / 12, {"close": False}),
    "icosphere_coarse": lambda: (*icosphere(0.1, 3), 0.2 / 10, {}),
    "icosphere_fine": lambda: (*icosphere(0.1, 3), 0.2 / 40, {}),
    "icosphere_turned": lambda: (transformed(icosphere(0.2, 4)[0], 3), icosphere(0.2, 4)[1], 0.0093, {"padding": 3}),
    "soup": lambda: (*soup(1), 0.011, {}),
    "soup_bounds": lambda: (*soup(2, 50), 0.02, {"bounds": ([-0.5, -0.4, -0.35], [0.3, 0.45, 0.3])}),
    "big_sphere": lambda: (*icosphere(0.1, 5), 0.2 / 150, {}),
}


@pytest.mark.parametrize("case", sorted(MESHES))
def test_mesh_equals_oracle(case):
    p, ix, h, kw = MESHES[case]()
    dev = ingest.voxelize_mesh(p, ix, h, **kw)
    ref = O.voxelize_mesh(p, ix, h, **kw)
    print(case, len(ix) // 3, "triangles", dev.dims, dev.count())
    same_grid(dev, ref)


def test_mesh_sphere_volume():
    p, ix = icosphere(0.1, 4)
    exact = 4 / 3 * math.pi * 0.1 ** 3
    assert 0.9 < ingest.voxelize_mesh(p, ix, 0.2 / 60).volume() / exact < 1.15


# ---------------------------------------------------------------- clouds
SHELL = sphere_shell(0.05, 6000)


def anisotropic_cloud(seed, n, spread=0.01):
    rng = np.random.default_rng(seed)
    pos = rng.normal(scale=0.1, size=(n, 3)).astype(np.float32)
    rot = random_rotations(rng, n)
    scale = np.exp(rng.normal(np.log(spread), 0.6, size=(n, 3))).astype(np.float32)
    return pos, rot, scale


CLOUDS = {
    "shell_fill": lambda: (*cloud_at(SHELL, 0.0008), 0.005, {}),
    "shell_hollow": lambda: (*cloud_at(SHELL, 0.0008), 0.005, {"fill": False}),
    "shell_fine": lambda: (*cloud_at(SHELL, 0.0008), 0.0032, {}),
    "open_capture": lambda: (*cloud_at(sphere_shell(0.05, 6000, 0.75), 0.0008), 0.005, {}),
    "seal0": lambda: (*cloud_at(SHELL, 0.0008), 0.004, {"seal": 0}),
    "seal3_sigma3": lambda: (*cloud_at(SHELL, 0.0005), 0.004, {"seal": 3, "sigma": 3.0}),
    "turned_blobs": lambda: (*anisotropic_cloud(5, 3000), 0.008, {}),
    "turned_blobs_hollow": lambda: (*anisotropic_cloud(6, 3000, 0.004), 0.006, {"fill": False, "padding": 4}),
    "one_needle": lambda: (np.zeros((1, 3), np.float32),
                           np.array([[0, 0, math.sqrt(0.5), math.sqrt(0.5)]], np.float32),
                           np.array([[0.05, 0.004, 0.004]], np.float32), 0.0125, {"fill": False}),
    "large_capture": lambda: (*anisotropic_cloud(7, 200_000, 0.003), 0.004, {}),
    # one cell thick on y, no padding: the stamp box reaches index -1
    "thin_grid": lambda: (*flat_cloud(), 0.012, {"padding": 0}),
    "thin_grid_hollow": lambda: (*flat_cloud(), 0.02, {"padding": 0, "fill": False}),
    # one floater covering most of the grid among small Gaussians
    "floater": lambda: (*with_floater(), 0.004, {}),
}


def flat_cloud():
    rng = np.random.default_rng(3)
    p = (rng.normal(size=(20, 3)) * [0.01, 0.001, 0.03]).astype(np.float32)
    return p, np.tile(np.array([0, 0, 0, 1], np.float32), (20, 1)), np.full((20, 3), 0.002, np.float32)


def with_floater():
    p, r, s = anisotropic_cloud(8, 2000, 0.003)
    s[17] = [0.08, 0.05, 0.1]
    return p, r, s


@pytest.mark.parametrize("case", sorted(CLOUDS))
def test_cloud_equals_oracle(case):
    p, r, s, h, kw = CLOUDS[case]()
    dev = ingest.voxelize_cloud(p, r, s, h, **kw)
    ref = O.voxelize_cloud(p, r, s, h, **kw)
    print(case, len(p), "Gaussians", dev.dims, dev.count())
    same_grid(dev, ref)


def test_cloud_centres_are_solid():
    p, r, s = cloud_at(SHELL, 0.0008)
    for h in (0.013, 0.0065, 0.0032):
        g = ingest.voxelize_cloud(p, r, s, h)
        idx = np.floor((p.astype(np.float64) - g.origin) / g.h).astype(np.int64)
        assert g.array()[idx[:, 2], idx[:, 1], idx[:, 0]].all()


def test_refusals_carry_the_reference_messages():
    with pytest.raises(ValueError, match="a cloud with no Gaussians has no volume"):
        ingest.voxelize_cloud(np.zeros((0, 3)), np.zeros((0, 4)), np.zeros((0, 3)), 0.01)
    with pytest.raises(ValueError, match="a cloud with no extent has no volume"):
        ingest.voxelize_cloud(*cloud_at(np.zeros((1, 3), np.float32), 0.0), 0.01)
    p, r, s = cloud_at(SHELL[:40].copy(), 0.001)
    p[3, 0] = np.nan
    with pytest.raises(ValueError, match="no extent"):
        ingest.voxelize_cloud(p, r, s, 0.01)
    for bad in (0.0, -1.0, float("nan"), float("inf")):
        with pytest.raises(ValueError, match="elementSize must be a positive length in metres, got"):
            ingest.voxelize_shape("sphere", [0.1], bad)
    with pytest.raises(ValueError, match="more cells"):
        ingest.voxelize_shape("box", [100, 100, 100], 0.001)
    with pytest.raises(ValueError, match="finite"):
        ingest.voxelize_mesh(box_mesh(0.5)[0], np.array([0, 1, 9], np.uint32), 0.1)


# ---------------------------------------------------------------- end to end
def test_capture_to_frames_on_the_device():
    # PLY bytes -> cloudFromPly -> device grid -> device lattice -> 3 frames,
    # against the same body with its lattice built on the host.
    from test_ply import build_ply
    rng = np.random.default_rng(11)
    pts = sphere_shell(0.05, 3000)
    n = len(pts)
    q = rng.normal(size=(n, 4)).astype(np.float32)
    cols = {"x": pts[:, 0], "y": pts[:, 1], "z": pts[:, 2], "opacity": rng.normal(size=n),
            "f_dc_0": rng.normal(size=n), "f_dc_1": rng.normal(size=n), "f_dc_2": rng.normal(size=n),
            "scale_0": np.full(n, math.log(0.002)), "scale_1": np.full(n, math.log(0.002)),
            "scale_2": np.full(n, math.log(0.001)), "rot_0": q[:, 0], "rot_1": q[:, 1], "rot_2": q[:, 2],
            "rot_3": q[:, 3]}
    cloud = ingest.cloud_from_ply(build_ply([(k, "float", v) for k, v in cols.items()]))
    grid = ingest.voxelize_cloud(cloud.positions, cloud.rotations, cloud.scales, 0.008)
    ref = O.voxelize_cloud(cloud.positions, cloud.rotations, cloud.scales, 0.008)
    same_grid(grid, ref)
    mat = scenes.Material(**scenes.C1_MATERIAL)
    body = grid.body(mat, splats=(cloud.positions, cloud.rotations, cloud.scales))
    lo = [grid.origin[a] - 0.05 for a in range(3)]
    hi = [grid.origin[a] + grid.dims[a] * grid.h + 0.05 for a in range(3)]
    sc = scenes.Scene([body], dt=float(np.float32(1 / 600)), tank_min=tuple(lo), tank_max=tuple(hi), name="capture")
    w = native.World(max_scenes=2, max_bodies=2, max_nodes=2 * sc.n_nodes + 16, max_cells=2 * sc.n_cells + 16,
                     max_surface=64, max_buckets=64, max_driven=16, max_splats=2 * n + 16)
    a = w.add(sc)
    b = w.add(sc, device_lattice=True)
    for f in range(3):
        w.frame(10, splats=True, graph=f == 1)
    pa, va = w.read_nodes(a)
    pb, vb = w.read_nodes(b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert all(np.array_equal(x, y) for x, y in zip(w.read_splats(a), w.read_splats(b)))
    active = body.lattice.inv_mass > 0
    assert np.isfinite(pa).all() and (pa[1][active] < body.lattice.rest[active, 1]).mean() > 0.9   # it fell
