"""The oracle's voxelisers (gridForBounds, voxelizeSdf, voxelizeMesh,
morphologicalClose: voxelize.js; voxelizeCloud: cloudVolume.js), pinned by
the known answers and properties of
packages/visionary-sim/test/unit/voxelize.test.js and cloudVolume.test.js on
the same synthetic inputs, plus closed forms."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from _geometry import box_mesh, cloud_at, icosphere, sphere_shell


# ------------------------------------------------------------ gridForBounds
def test_grid_for_bounds_known_answers():
    assert O.grid_for_bounds([0, 0, 0], [0.4, 0.2, 0.1], 0.02).dims == (24, 14, 9)
    assert O.grid_for_bounds([0, 0, 0], [0.41, 0.2, 0.1], 0.02).dims[0] == 25
    ext = float(np.float32(0.15))
    assert ext > 0.15
    assert O.grid_for_bounds([0, 0, 0], [ext] * 3, 0.0125).dims == (16, 16, 16)   # float32 dust
    g = O.grid_for_bounds([-1, -1, -1], [1, 1, 1], 0.25, padding=2)
    assert abs(g.origin[0] - (-1 - 2 * g.h)) < 1e-12 and g.occupancy.sum() == 0
    for bad in (0.0, -1.0, float("nan")):
        with pytest.raises(ValueError, match="elementSize"):
            O.grid_for_bounds([0, 0, 0], [0.1] * 3, bad)


# --------------------------------------------------------------- voxelizeSdf
def test_sdf_sphere_volume():
    r = 0.1
    g = O.voxelize_shape("sphere", [r], 2 * r / 32)
    ratio = g.volume() / (4 / 3 * math.pi * r ** 3)
    assert 1.0 < ratio < 1.16


def test_sdf_torus_keeps_its_hole():
    g = O.voxelize_shape("torus", [0.08, 0.03], 0.22 / 28)
    centre = [math.floor((0 - g.origin[a]) / g.h) for a in range(3)]
    assert g.occupancy[centre[0] + g.dims[0] * (centre[1] + g.dims[1] * centre[2])] == 0
    assert g.occupancy.sum() > 0


def test_sdf_box_is_exact_and_translates():
    # Centres land on the faces, so an axis-aligned box fills exactly its
    # cells plus the half-cell rim the containment rule claims.
    g = O.voxelize_shape("box", [0.1, 0.05, 0.075], 0.025)
    assert g.occupancy.sum() == (8 + 1) * (4 + 1) * (6 + 1)
    t = O.voxelize_shape("box", [0.1, 0.05, 0.075], 0.025, offset=(0.3, -0.2, 0.05))
    assert np.array_equal(t.occupancy, g.occupancy) and t.dims == g.dims
    assert abs(t.origin[0] - g.origin[0] - 0.3) < 1e-12


def test_close_seals_a_pit_without_eating_the_body():
    g = O.voxelize_shape("box", [0.1, 0.1, 0.1], 0.2 / 12)
    before = int(g.occupancy.sum())
    pit = g.dims[0] // 2 + g.dims[0] * (g.dims[1] // 2 + g.dims[1] * (g.dims[2] // 2))
    g.occupancy[pit] = 0
    O.close_grid(g)
    assert g.occupancy[pit] == 1 and int(g.occupancy.sum()) == before


# -------------------------------------------------------------- voxelizeMesh
def test_mesh_closed_box_volume():
    p, ix = box_mesh(0.5)
    g = O.voxelize_mesh(p, ix, 1.0 / 20)
    assert 0.95 < g.volume() / 1.0 < 1.35


def test_mesh_survives_an_open_hole():
    p, ix = box_mesh(0.5)
    holed = np.concatenate([ix[:12], ix[18:]])   # drop the -y face
    g = O.voxelize_mesh(p, holed, 1.0 / 16)
    assert g.volume() / 1.0 > 0.7


def test_mesh_obj_box_fill_ratio():
    p, ix = box_mesh(0.5)
    g = O.voxelize_mesh(p, ix, 1.0 / 12)
    ratio = g.occupancy.sum() / g.occupancy.size
    assert 0.1 < ratio < 0.75


def test_mesh_sphere_volume_converges():
    p, ix = icosphere(0.1, 3)
    exact = 4 / 3 * math.pi * 0.1 ** 3
    coarse = O.voxelize_mesh(p, ix, 0.2 / 10).volume() / exact
    fine = O.voxelize_mesh(p, ix, 0.2 / 40).volume() / exact
    assert abs(fine - 1) < abs(coarse - 1) and 0.9 < fine < 1.2


# -------------------------------------------------------------- voxelizeCloud
RADIUS = 0.05
SPREAD = 0.0008
SHELL = sphere_shell(RADIUS, 6000)


def test_cloud_centres_are_solid_at_every_size():
    p, r, s = cloud_at(SHELL, SPREAD)
    for h in (0.013, 0.0086, 0.0065, 0.0043, 0.0032):
        g = O.voxelize_cloud(p, r, s, h)
        assert sum(g.at(x) == 0 for x in p) == 0, h


def test_cloud_shell_becomes_solid():
    p, r, s = cloud_at(SHELL, SPREAD)
    exact = 4 / 3 * math.pi * RADIUS ** 3
    shell = O.voxelize_cloud(p, r, s, 0.005, fill=False)
    solid = O.voxelize_cloud(p, r, s, 0.005)
    assert shell.volume() / exact < 0.7
    assert 1.15 < solid.volume() / exact < 1.45


def test_cloud_volume_converges():
    p, r, s = cloud_at(SHELL, SPREAD)
    exact = 4 / 3 * math.pi * RADIUS ** 3
    err = lambda h: abs(O.voxelize_cloud(p, r, s, h).volume() - exact) / exact
    assert err(0.0037) < err(0.01) * 0.7


def test_open_capture_stays_hollow():
    p, r, s = cloud_at(sphere_shell(RADIUS, 6000, 0.75), SPREAD)
    shell = O.voxelize_cloud(p, r, s, 0.005, fill=False).occupancy.sum()
    filled = O.voxelize_cloud(p, r, s, 0.005).occupancy.sum()
    assert filled < shell * 1.5


def test_cloud_padding_ring_clear():
    g = O.voxelize_cloud(*cloud_at(SHELL, SPREAD), 0.0065)
    a = g.array()
    faces = [a[0], a[-1], a[:, 0], a[:, -1], a[:, :, 0], a[:, :, -1]]
    assert sum(int(f.sum()) for f in faces) == 0


def test_cloud_stamps_along_its_own_axes():
    p = np.zeros((1, 3), np.float32)
    s = np.array([[0.05, 0.004, 0.004]], np.float32)
    r = np.array([[0, 0, math.sqrt(0.5), math.sqrt(0.5)]], np.float32)
    g = O.voxelize_cloud(p, r, s, 0.0125, fill=False)
    assert g.at([0, 0.08, 0]) == 1 and g.at([0.08, 0, 0]) == 0


def test_cloud_refuses_nothing():
    with pytest.raises(ValueError, match="no Gaussians"):
        O.voxelize_cloud(np.zeros((0, 3)), np.zeros((0, 4)), np.zeros((0, 3)), 0.01)
    with pytest.raises(ValueError, match="no extent"):
        O.voxelize_cloud(*cloud_at(np.zeros((1, 3), np.float32), 0.0), 0.01)


def test_fill_enclosed_matches_scipy_labelling():
    # The flood is the complement of the background component that touches
    # the faces; scipy's connected-component labelling (6-connectivity) gives
    # the same set independently.
    from scipy import ndimage
    rng = np.random.default_rng(4)
    for trial in range(6):
        occ = (rng.random((14, 11, 17)) < [0.25, 0.45, 0.6][trial % 3]).astype(np.uint8)
        g = O.Grid((17, 11, 14), (0, 0, 0), 0.1, occ.reshape(-1).copy())
        grid = O.OGrid()
        grid.dims[:] = g.dims
        O.load("f64").o_fill_enclosed(grid, g.occupancy.ctypes.data)
        lab, _ = ndimage.label(occ == 0)
        border = set(np.unique(np.concatenate([lab[0].ravel(), lab[-1].ravel(), lab[:, 0].ravel(),
                                                lab[:, -1].ravel(), lab[:, :, 0].ravel(), lab[:, :, -1].ravel()])))
        outside = np.isin(lab, list(border - {0}))
        want = (~outside).astype(np.uint8)
        assert np.array_equal(g.occupancy.reshape(occ.shape), want)


def test_refusals_of_unusable_inputs():
    # Math.min/max carry a NaN Gaussian into the bounds: the cloud has no extent.
    p, r, s = cloud_at(SHELL[:50].copy(), SPREAD)
    p[7, 1] = np.nan
    with pytest.raises(ValueError, match="no extent"):
        O.voxelize_cloud(p, r, s, 0.01)
    # bounds beyond what a grid can hold
    with pytest.raises(ValueError, match="more cells"):
        O.grid_for_bounds([0, 0, 0], [1e6, 1e6, 1e6], 0.001)
    with pytest.raises(ValueError, match="more cells"):
        O.grid_for_bounds([0, 0, 0], [np.inf, 1, 1], 0.1)
    q, ix = box_mesh(0.5)
    q = q.copy()
    q[3, 2] = np.inf
    with pytest.raises(ValueError, match="finite"):
        O.voxelize_mesh(q, ix, 0.1)
    with pytest.raises(ValueError, match="finite"):
        O.voxelize_mesh(box_mesh(0.5)[0], np.array([0, 1, 8], np.uint32), 0.1)


def test_mesh_bounds_argument_lays_out_the_grid():
    p, ix = box_mesh(0.4333)
    own = O.voxelize_mesh(p, ix, 0.1)
    lo, hi = p.min(0).astype(np.float64), p.max(0).astype(np.float64)
    same = O.voxelize_mesh(p, ix, 0.1, bounds=(lo, hi))
    assert same.dims == own.dims and np.array_equal(same.occupancy, own.occupancy)
    wide = O.voxelize_mesh(p, ix, 0.1, bounds=(lo - [0.5, 0, 0], hi))
    assert wide.dims == (own.dims[0] + 5, own.dims[1], own.dims[2])
    assert abs(wide.origin[0] - (own.origin[0] - 0.5)) < 1e-12
    # Five columns further along x, the same body up to the rays that graze
    # a face diagonal (their sample points round differently off a new origin).
    diff = int((wide.array()[:, :, 5:] != own.array()).sum())
    assert not wide.array()[:, :, :5].any() and diff <= 0.01 * own.occupancy.sum()


def test_translation_chain():
    # Offsets exact in binary: two translations are one by their sum.
    one = O.voxelize_shape("torus", [0.25, 0.125], 0.0625, offset=(0.75, -0.5, 0.25))
    two = O.voxelize_shape("torus", [0.25, 0.125], 0.0625, offset=[(0.5, -0.25, 0.0), (0.25, -0.25, 0.25)])
    assert one.dims == two.dims and one.origin == two.origin and np.array_equal(one.occupancy, two.occupancy)
    # A chain is applied outermost first to the query point: with offsets
    # whose sum rounds, the chain's origin is the nested sum's.
    chain = [(0.1, 0.2, 0.3), (0.7, -0.1, 0.05)]
    g = O.voxelize_shape("sphere", [0.1], 0.01, offset=chain)
    assert abs(g.origin[0] - ((-0.1 + 0.1) + 0.7 - 0.02 - 0.005)) < 1e-15
    assert g.occupancy.sum() > 0
