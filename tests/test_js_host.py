"""The JavaScript host (visionary-sim_amd/js) over the Node-API addon.

CPU: the host's clock, lattice and backend validation (tests/js/host.test.mjs,
addon faked). GPU: the Simulator runs a scene on the HIP backend and its
lattice, positions and splats are compared with the Python host and the f32
oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from vsim import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")


def run_node(*args, timeout=300):
    return subprocess.run([NODE, *args], cwd=ROOT, capture_output=True, text=True, timeout=timeout)


@pytest.mark.skipif(NODE is None, reason="node is not installed")
def test_js_host_cpu_suite():
    r = run_node("tests/js/run.mjs", "tests/js/host.test.mjs")
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0


def spec_for(sc, frames, tmp, splats=True, grab=None):
    b = sc.bodies[0]
    lat = b.lattice
    spec = dict(cells=list(lat.cell_dims), origin=list(lat.origin), h=lat.h,
                material=dict(youngsModulus=b.material.youngs_modulus, poissonRatio=b.material.poisson_ratio,
                              density=b.material.density, friction=b.material.friction, damping=b.material.damping,
                              relaxation=b.material.relaxation, maxStretch=b.material.max_stretch),
                pinned=b.static_nodes is not None, tank=dict(min=list(sc.tank_min), max=list(sc.tank_max)),
                dt=sc.dt, substeps=sc.substeps, frames=frames, splats=None, drive=None, grab=grab)
    if splats and b.n_splats:
        d = os.path.join(tmp, "splats")
        os.makedirs(d, exist_ok=True)
        b.splat_mean.astype(np.float32).tofile(os.path.join(d, "means.bin"))
        b.splat_quat.astype(np.float32).tofile(os.path.join(d, "quats.bin"))
        b.splat_scale.astype(np.float32).tofile(os.path.join(d, "scales.bin"))
        spec["splats"] = d
    if sc.drive is not None:
        fdt = sc.substeps * sc.dt
        spec["drive"] = dict(law=sc.drive.law, nodes=[int(n) for n in sc.drive.nodes],
                             offsets=sc.drive.offsets.astype(float).tolist(),
                             targets=[list(map(float, sc.drive.target(f * fdt))) for f in range(frames + 1)])
    path = os.path.join(tmp, "spec.json")
    with open(path, "w") as f:
        json.dump(spec, f)
    return path


def load(out, name, dtype=np.float32):
    return np.fromfile(os.path.join(out, name), dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["block", "drag"])
def test_js_simulator_matches_oracle(tmp_path, case):
    from oracle.oracle import OracleScene
    if case == "block":
        sc = scenes.jelly_block(16, splats=20_000, seed=21, name="js")
    else:
        sc = scenes.jelly_block(12, splats=0, seed=22, pinned=False, tank_margin=0.3, name="jsdrag")
        sc.tank_min = (sc.tank_min[0], 0.0, sc.tank_min[2])
        sc.drive = scenes.handle_drive(sc, seed=4)
    frames = 5
    lat0 = sc.bodies[0].lattice
    centre = [lat0.origin[a] + lat0.cell_dims[a] * lat0.h / 2 + (0.003 if a != 1 else 0) for a in range(3)]
    grab = dict(origin=[centre[0], centre[1] + 1.0, centre[2]], direction=[0.0, -1.0, 0.0]) if case == "block" else None
    spec = spec_for(sc, frames, str(tmp_path), grab=grab)
    out = str(tmp_path / "out")
    r = run_node("tests/js/sim_gpu.mjs", spec, out)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0
    lat = sc.bodies[0].lattice
    # The JS host builds exactly the Python host's lattice.
    assert np.array_equal(load(out, "rest.bin").reshape(-1, 3), lat.rest)
    assert np.array_equal(load(out, "cells.bin", np.uint32), lat.cells)
    assert np.array_equal(load(out, "surface.bin", np.uint32), scenes.surface_nodes(lat))
    arrays = sc.arrays()
    assert np.array_equal(load(out, "w.bin"), arrays["w"])
    o = OracleScene(sc, "f32")
    for _ in range(frames):
        o.frame(sc.substeps)
        o.deform_splats(axes=True)
    pos = load(out, "pos.bin").reshape(3, -1)
    vel = load(out, "vel.bin").reshape(3, -1)
    scale = np.max(np.abs(o.p))
    err = np.max(np.abs(pos - o.p)) / scale
    print(f"{case}: exact {np.mean(pos == o.p):.6f} rel {err:.3e}")
    assert err <= 1e-5
    assert np.max(np.abs(vel - o.v)) / max(np.max(np.abs(o.v)), 1e-6) <= 1e-4
    if sc.bodies[0].n_splats:
        mean = load(out, "splat_means.bin").reshape(3, -1)
        cov = load(out, "splat_cov.bin").reshape(6, -1)
        assert np.max(np.abs(mean - o.out_mean)) / np.max(np.abs(o.out_mean)) <= 1e-5
        assert np.max(np.abs(cov - o.out_cov)) / np.max(np.abs(o.out_cov)) <= 1e-5
        rot = load(out, "splat_rot.bin").reshape(4, -1)
        axis = load(out, "splat_axis.bin").reshape(3, -1)
        print(f"engine axes: rotation exact {np.mean(rot == o.out_rot):.6f}, lengths exact "
              f"{np.mean(axis == o.out_axis):.6f}")
        assert np.max(np.abs(rot - o.out_rot)) <= 1e-5
    if grab is not None:
        # The device pick and grip against the oracle's restatement on the
        # same final positions (grab mass: material inverse mass, statics 0).
        from oracle import oracle as O
        got = json.load(open(os.path.join(out, "meta.json")))["grab"]
        body = sc.bodies[0]
        gm = body.lattice.inv_mass.copy()
        gm[body.static_nodes] = 0
        anchor, _ = O.pick_ray(pos, gm, grab["origin"], grab["direction"], body.lattice.h * 1.2)
        assert anchor >= 0
        nodes, offs, found = O.grip_radius(pos, gm, anchor, 0, pos.shape[1], 0.03, 4096)
        print(f"grab: anchor {anchor}, held {len(nodes)} of {found}")
        assert got["held"] == [int(n) for n in nodes]
        assert np.array_equal(np.asarray(got["offsets"], np.float32), offs.astype(np.float32))
        assert got["hit"]["position"] == [float(v) for v in pos[:, anchor]]
        assert got["hit"]["held"] == len(nodes)
        assert np.max(np.abs(axis - o.out_axis)) / np.max(np.abs(o.out_axis)) <= 1e-5


@pytest.mark.gpu
def test_js_capture_voxelised_on_the_device(tmp_path):
    # A PLY capture through the JS host: the device grid and the lattice the
    # JS host builds on it equal the oracle's grid and the Python lattice, and
    # the frames stepped on it equal the f32 oracle's.
    import math
    from _geometry import sphere_shell
    from oracle import oracle as O
    from oracle.oracle import OracleScene
    from test_ply import build_ply
    from vsim import ingest
    rng = np.random.default_rng(12)
    pts = sphere_shell(0.05, 2500)
    n = len(pts)
    cols = {"x": pts[:, 0], "y": pts[:, 1], "z": pts[:, 2], "opacity": rng.normal(size=n),
            "f_dc_0": rng.normal(size=n), "f_dc_1": rng.normal(size=n), "f_dc_2": rng.normal(size=n),
            "scale_0": np.full(n, math.log(0.003)), "scale_1": np.full(n, math.log(0.002)),
            "scale_2": np.full(n, math.log(0.001)), "rot_0": rng.normal(size=n), "rot_1": rng.normal(size=n),
            "rot_2": rng.normal(size=n), "rot_3": rng.normal(size=n)}
    ply = tmp_path / "capture.ply"
    data = build_ply([(k, "float", v) for k, v in cols.items()])
    ply.write_bytes(data)
    cloud = ingest.cloud_from_ply(data)
    h = 0.009
    grid = O.voxelize_cloud(cloud.positions, cloud.rotations, cloud.scales, h)
    mat = scenes.Material(**scenes.C1_MATERIAL)
    body = scenes.Body(scenes.build_lattice(grid.array(), grid.origin, grid.h, mat.density), mat)
    lo = [grid.origin[a] - 0.05 for a in range(3)]
    hi = [grid.origin[a] + grid.dims[a] * grid.h + 0.05 for a in range(3)]
    sc = scenes.Scene([body], dt=float(np.float32(1 / 600)), tank_min=tuple(lo), tank_max=tuple(hi), name="jscap")
    frames = 3
    spec = dict(ply=str(ply), elementSize=h, tank=dict(min=lo, max=hi), dt=sc.dt, substeps=sc.substeps, frames=frames,
                material=dict(youngsModulus=mat.youngs_modulus, poissonRatio=mat.poisson_ratio, density=mat.density,
                              friction=mat.friction, damping=mat.damping, relaxation=mat.relaxation,
                              maxStretch=mat.max_stretch))
    path = tmp_path / "spec.json"
    path.write_text(json.dumps(spec))
    out = str(tmp_path / "out")
    r = run_node("tests/js/capture_gpu.mjs", str(path), out)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0
    meta = json.load(open(os.path.join(out, "meta.json")))
    assert tuple(meta["cellDims"]) == grid.dims and tuple(meta["origin"]) == grid.origin and meta["h"] == grid.h
    lat = body.lattice
    assert np.array_equal(load(out, "cells.bin", np.uint32), lat.cells)
    assert np.array_equal(load(out, "rest.bin").reshape(-1, 3), lat.rest)
    assert np.array_equal(load(out, "w.bin"), sc.arrays()["w"])
    o = OracleScene(sc, "f32")
    for _ in range(frames):
        o.frame(sc.substeps)
    pos = load(out, "pos.bin").reshape(3, -1)
    err = np.max(np.abs(pos - o.p)) / np.max(np.abs(o.p))
    print(f"capture: {lat.node_count} nodes, exact {np.mean(pos == o.p):.6f} rel {err:.3e}")
    assert err <= 1e-5
