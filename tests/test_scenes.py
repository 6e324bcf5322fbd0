"""Host-side scene assembly against the reference's lattice and material
rules (lattice.test.js, kuhn.test.js, compliance.test.js, timestep.js)."""
import math

import numpy as np
import pytest

from vsim import scenes


def corners(lat, linear):
    cx, cy, cz = lat.cell_dims
    nx, ny, _ = lat.node_dims
    i, j, k = linear % cx, (linear // cx) % cy, linear // (cx * cy)
    return {(i + (c & 1)) + nx * ((j + ((c >> 1) & 1)) + ny * (k + ((c >> 2) & 1))) for c in range(8)}


def test_same_colour_cells_share_no_node():
    rng = np.random.default_rng(2)
    occ = rng.random((5, 6, 7)) < 0.7
    lat = scenes.build_lattice(occ, (0, 0, 0), 0.01, 1000)
    start = 0
    for c in range(8):
        seen = set()
        for lin in lat.cells[start:start + lat.color_counts[c]]:
            nodes = corners(lat, int(lin))
            assert not (nodes & seen)
            seen |= nodes
        start += int(lat.color_counts[c])
    assert start == int(occ.sum())


def test_cells_are_colour_grouped_in_scan_order():
    lat = scenes.block_lattice((4, 3, 2), (0, 0, 0), 0.01, 1000)
    start = 0
    for c in range(8):
        seg = lat.cells[start:start + lat.color_counts[c]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        for lin in seg:
            i, j, k = lin % 4, (lin // 4) % 3, lin // 12
            assert scenes.cell_color(i, j, k) == c
        start += int(lat.color_counts[c])


def test_mass_is_conserved_and_lumped_on_corners():
    rng = np.random.default_rng(4)
    occ = rng.random((6, 5, 4)) < 0.6
    lat = scenes.build_lattice(occ, (0, 0, 0), 0.02, 1200)
    m = np.where(lat.inv_mass > 0, 1 / np.where(lat.inv_mass > 0, lat.inv_mass.astype(np.float64), 1), 0)
    assert m.sum() == pytest.approx(1200 * 0.02 ** 3 * occ.sum(), rel=1e-6)
    assert lat.mass == pytest.approx(1200 * 0.02 ** 3 * occ.sum(), rel=1e-12)


def test_surface_nodes_closed_forms():
    one = scenes.build_lattice(np.ones((1, 1, 1), bool), (0, 0, 0), 0.1, 1000)
    assert len(scenes.surface_nodes(one)) == 8
    for s in (2, 5, 16):
        lat = scenes.block_lattice((s, s, s), (0, 0, 0), 0.01, 1000)
        assert len(scenes.surface_nodes(lat)) == (s + 1) ** 3 - (s - 1) ** 3
    c1 = scenes.c1_scene(splats=0)
    assert len(scenes.surface_nodes(c1.bodies[0].lattice)) == 1538


def test_rest_positions_and_counts_of_the_16_block():
    sc = scenes.c1_scene(splats=0)
    lat = sc.bodies[0].lattice
    assert lat.node_count == 4913 and len(lat.cells) == 4096
    assert list(lat.color_counts) == [512] * 8
    n = 3 + 17 * (5 + 17 * 7)
    assert tuple(lat.rest[n]) == (np.float32(3 * 0.016), np.float32(5 * 0.016), np.float32(7 * 0.016))
    assert len(sc.bodies[0].static_nodes) == 17 * 17


def test_compliance_constants_of_the_16_block():
    sc = scenes.c1_scene(splats=0)
    dev, hyd, gamma = sc.bodies[0].constants()
    assert dev == pytest.approx(38.0859375, rel=1e-12)
    assert hyd == pytest.approx(25.390625, rel=1e-12)
    assert gamma == pytest.approx(1 + 0.4 / 0.6, rel=1e-12)


def test_stable_step_and_wave_rate():
    m = scenes.Material(**scenes.C1_MATERIAL)
    rate = scenes.wave_speed(m) / 0.016
    assert rate == pytest.approx(725, rel=2e-3)
    assert scenes.stable_step_duration(rate) == pytest.approx(0.75 / rate)
    assert scenes.stable_step_duration(0) == 1 / 240
    with pytest.raises(ValueError):
        scenes.stable_step_duration(float("nan"))


def test_poisson_clamp_and_floor_on_youngs_modulus():
    assert scenes.lame_parameters(scenes.Material(poisson_ratio=0.1)) == \
        scenes.lame_parameters(scenes.Material(poisson_ratio=0.25))
    assert scenes.lame_parameters(scenes.Material(youngs_modulus=0.0)) == \
        scenes.lame_parameters(scenes.Material(youngs_modulus=1.0))


def test_broadphase_grid_layout():
    sc = scenes.c3_scene(bodies=8)
    origin, edge, dims = sc.grid()
    assert edge == pytest.approx(0.01, rel=1e-6)
    assert dims == [math.ceil(1.0 / (2 * float(np.float32(0.005)))) + 2] * 3
    assert origin[0] == pytest.approx(-edge, rel=1e-6)


def test_seeded_workloads_are_reproducible():
    a, b = scenes.c4_scenes(4), scenes.c4_scenes(4)
    for x, y in zip(a, b):
        assert x.n_nodes == y.n_nodes
        assert np.array_equal(x.bodies[0].splat_mean, y.bodies[0].splat_mean)
    r = scenes.c5_requests(5)
    assert all(t2 > t1 for (t1, _, _), (t2, _, _) in zip(r, r[1:]))
