"""Benchmark: node-substeps/s of the visionary-sim substep on one MI355X.

Default workload: 256 independent scenes (jelly blocks, lattice side ~ U{8..48},
10k-500k Gaussians each, 10 substeps per 1/60 s frame), stepped batched -- one
launch per pass across every scene, the frame replayed from a hipGraph -- and,
beside it, the plain per-scene launch path on the same scenes. A step is one
frame of every scene: 10 substeps (predict, the 8 colours -- in pairs, by rows
of cells, one launch per pair -- finalize) and the splat deformation. Inputs
are synthetic and resident on the device before the timed region.

The autotuner (vs_world_autotune) picks the launch shapes and schedule first
(--no-autotune keeps the defaults).

Prints ONE JSON line. `value` is the batched whole-job node-substeps/s;
`plain`, `speedup`, frame latency p50/p99, the per-pass profile, the roofline
of the dominant kernel (the colour-pair launches, timed inside a frame, with
the back-to-back figure beside it; its bound is the VALU when the committed
counter summary's VALUBusy for the launch exceeds its HBM fraction), each
pass's achieved GB/s against HBM peak, the single-scene workloads (16^3
block, 64^3 drag block with its sphere collider, 512 contact bodies at
2-voxel buckets, falling and piled), the serving loop at
50 requests/s and saturated, and the CPU baseline (the f32 oracle, this
repository's C restatement of the reference's WGSL, on a bounded sample,
all cores) ride along.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "visionary-sim_amd", "py"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
VALU_ISSUE_PER_S = 256 * 2.4e9   # wave instructions per second: one per CU per cycle
# Algorithmic bytes (SoA f32) per item of each pass, per launch:
ELASTIC_BYTES_PER_CELL = 8 * (12 + 12 + 4) + 8 * 12 + 4   # gather p, prev, w; scatter p; cell id
PREDICT_BYTES_PER_NODE = 64
FINALIZE_BYTES_PER_NODE = 52
SPLAT_BYTES = 76


def _profile_rows(name, kernel):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return []
    return [line.split("|") for line in open(path) if line.startswith(f"| {kernel}<")]


def _most_dispatched(name, kernel, column):
    """`column` of the row of `kernel` with the most dispatches in a committed
    profile summary: the plain colour-pair launch of the frame (the first and
    last pairs, which also run node passes, are fewer). None when the summary
    is absent."""
    rows = [r for r in _profile_rows(name, kernel) if len(r) > column]
    if not rows:
        return None
    return float(max(rows, key=lambda r: int(r[3]))[column])


def profile_source(name):
    """Where a committed counter figure comes from: the summary file and the
    source state it names on its 'Measured at' line."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    stamp = next((l.strip() for l in open(path) if l.startswith("Measured at")), "Measured at: unstated")
    return f"profiles/{name} ({stamp}); rocprofv3 counters, not measured in this run"


def elastic_traffic_bytes(kernel):
    """HBM bytes per plain colour-pair launch of `kernel` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary, scaled by the copy calibration."""
    v = _most_dispatched("batched256_hbm_traffic.md", kernel, 6)
    return v * 1e6 if v is not None else None


def elastic_valu_instructions(kernel):
    """VALU wave instructions per plain colour-pair launch of `kernel`
    (rocprofv3 SQ_INSTS_VALU summary)."""
    v = _most_dispatched("batched256_valu.md", kernel, 4)
    return v * 1e6 if v is not None else None


def elastic_valu_busy(kernel):
    """(VALUBusy, VALUUtilization) of the plain colour-pair launch of
    `kernel` as fractions, from the same summary: the share of cycles its
    VALUs issue and the share of lanes active per VALU instruction."""
    busy = _most_dispatched("batched256_valu.md", kernel, 7)
    lanes = _most_dispatched("batched256_valu.md", kernel, 8)
    return (busy / 100 if busy is not None else None), (lanes / 100 if lanes is not None else None)


def pair_class_nodes(scs, cls):
    """Nodes a colour-pair launch of row class `cls` covers: cell rows y =
    py + 2a, z = pz + 2b of every dense body, two node rows each way."""
    py, pz = cls & 1, cls >> 1
    total = 0
    for sc in scs:
        for b in sc.bodies:
            cx, cy, cz = b.lattice.cell_dims
            total += (cx + 1) * 2 * ((cy + 1 - py) // 2) * 2 * ((cz + 1 - pz) // 2)
    return total


def pair_launch_bytes(scs):
    """Algorithmic bytes of one colour-pair launch averaged over the four row
    classes: 40 B (p, prev, w read; p written) per covered node."""
    return 40.0 * sum(pair_class_nodes(scs, c) for c in range(4)) / 4


def plain_pair_bytes(scs, substeps, merged=True):
    """Average algorithmic bytes of the colour-pair launches of a frame that
    run no node pass (the profile's "pair" kind): with merged node passes,
    every class-3 launch finalizes and the frame's first class-0 launch
    predicts, so those are left out."""
    total, count = 0.0, 0
    for s in range(substeps):
        for cls in range(4):
            if merged and (cls == 3 or (s == 0 and cls == 0)):
                continue
            total += 40.0 * pair_class_nodes(scs, cls)
            count += 1
    return total / count if count else 0.0


def leftover_nodes(scs, pairs=True, merged=True):
    """Nodes the node pass moves inside a frame when the last colour pair
    finalizes and predicts its own: every node of a scene with contact or
    tensile sweeps, the rows class 3 leaves (y = 0, z = 0 and the last ones on
    odd sides) of the dense bodies of other scenes."""
    total = 0
    for sc in scs:
        fold = not sc.contact_active and all(b.material.max_stretch <= 1.0 for b in sc.bodies)
        for b in sc.bodies:
            cx, cy, cz = b.lattice.cell_dims
            dense = len(b.lattice.cells) == cx * cy * cz
            n = b.lattice.node_count
            if fold and dense and pairs and merged:
                n -= (cx + 1) * 2 * (cy // 2) * 2 * (cz // 2)
            total += n
    return total


def pass_rates(prof, scs, substeps, splats, pair_bytes, pairs, merged):
    """Achieved algorithmic GB/s of each pass of one instrumented batched
    frame (HIP events around each launch, so small launches read low), against
    HBM peak. Contact figures are lower bounds: the bytes per examined slot
    are left out."""
    nodes = sum(s.n_nodes for s in scs)
    records = sum(s.n_surface_active for s in scs)
    per_launch = {
        "predict": nodes * PREDICT_BYTES_PER_NODE,
        "finalize": leftover_nodes(scs, pairs, merged) * FINALIZE_BYTES_PER_NODE,
        "finalize+predict": leftover_nodes(scs, pairs, merged) * 52,
        "elastic": pair_bytes,
        "splat": splats * SPLAT_BYTES,
        "bin": records * 40,
        "sort": records * 4,
        "collide": records * 44,
        "settle": records * 44,
    }
    out = {}
    for k, (ms, n) in prof.items():
        if not n or k not in per_launch:
            continue
        us = ms / n * 1e3
        gbs = per_launch[k] / (us * 1e-6) / 1e9
        out[k] = {"launches": n, "avg_us": us, "bytes_per_launch": per_launch[k], "gb_s": gbs,
                  "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    return out


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sync(world):
    world.sync()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def make_world(scs, **kw):
    from vsim import native
    nodes = sum(s.n_nodes for s in scs)
    cells = sum(s.n_cells for s in scs)
    splats = sum(s.n_splats for s in scs)
    driven = sum(len(s.drive.nodes) for s in scs if s.drive is not None)
    surf = sum(s.n_surface_active for s in scs)
    buckets = sum(int(np.prod(s.grid()[2])) for s in scs if s.contact_active)
    return native.World(max_scenes=len(scs) + 1, max_bodies=sum(len(s.bodies) for s in scs) + 1,
                        max_nodes=nodes + 1024, max_cells=cells + 1024, max_surface=surf + 1024,
                        max_buckets=buckets + 1024, max_driven=driven + 64, max_splats=splats + 1024, **kw)


def set_targets(world, ids, scs, frame):
    for sid, sc in zip(ids, scs):
        if sc.drive is not None:
            t, v = sc.drive.frame_target(frame, sc.substeps * sc.dt)
            world.set_target(sid, t, v)


_CPU_SCENES = []


def _cpu_worker(args):
    """One process of the all-cores baseline: whole frames of every
    `stride`-th scene from `start` until `budget_s` is spent; returns
    node-substeps done."""
    from oracle.oracle import OracleScene
    start, stride, budget_s = args
    t0, done = time.perf_counter(), 0
    for i in range(start, len(_CPU_SCENES), stride):
        sc = _CPU_SCENES[i]
        o = OracleScene(sc, "f32")
        for f in range(2):
            o.frame(sc.substeps, splats=True)
            done += sc.n_nodes * sc.substeps
        if time.perf_counter() - t0 > budget_s:
            break
    return done


def cpu_baseline(scs, budget_s, cores):
    """The f32 oracle on whole frames of the workload's scenes: one core
    until `budget_s` is spent, then `cores` processes (forked before the GPU
    is touched, one scene at a time each) for about as long. Reported value:
    all cores."""
    from oracle.oracle import OracleScene
    done_ns, t0, frames, scenes_used = 0, time.perf_counter(), 0, 0
    elapsed = 0.0
    for sc in scs:
        o = OracleScene(sc, "f32")
        scenes_used += 1
        for f in range(2):
            a = time.perf_counter()
            o.frame(sc.substeps, splats=True)
            elapsed += time.perf_counter() - a
            done_ns += sc.n_nodes * sc.substeps
            frames += 1
        if time.perf_counter() - t0 > budget_s:
            break
    one = done_ns / elapsed
    out = dict(value=one, unit="node-substeps/s", cores=1, kind="port",
               sample=f"{scenes_used} scenes x 2 frames ({frames} scene-frames, {done_ns} node-substeps, "
                      f"{elapsed:.1f} s) of the same workload, splats included", one_core=one)
    if cores > 1:
        import multiprocessing as mp
        _CPU_SCENES[:] = list(scs)
        a = time.perf_counter()
        with mp.get_context("fork").Pool(cores) as pool:
            done = sum(pool.map(_cpu_worker, [(k, cores, budget_s) for k in range(cores)]))
        wall = time.perf_counter() - a
        _CPU_SCENES[:] = []
        out.update(value=done / wall, cores=cores,
                   sample=f"{cores} processes, each whole frames (2 a scene) of every {cores}th scene of the same "
                          f"workload until {budget_s:.0f} s: {done} node-substeps in {wall:.1f} s wall, splats "
                          f"included; one core alone: {one:.4g} node-substeps/s ({out['sample']})")
    return out


def serve_leg(count, budget_nodes, realtime=True):
    """Scenes with log-uniform 4k-500k nodes, each running 120 frames with a
    seeded drag handle, admitted while a world of `budget_nodes` node
    capacity holds them; continuous batching. `realtime`: the requests
    arrive as a Poisson process (50/s), so the figure is mostly the arrival
    rate; otherwise all are waiting at the start, and the server runs at
    capacity (admission only by the budget)."""
    from vsim import native, scenes
    reqs = scenes.c5_requests(count)
    cells = int(budget_nodes * 1.05)
    world = native.World(max_scenes=256, max_bodies=256, max_nodes=budget_nodes, max_cells=cells,
                         max_driven=256 * 27, max_splats=1024, persistent=2048, branches=2)
    srv = native.Server(world, substeps=10, graph=True)
    for t, sc, frames in reqs:
        srv.submit(sc, t, frames)
    rep = srv.run(realtime=realtime)
    lat = srv.latencies()
    srv.close()
    world.close()
    return {"requests": count, "arrival_rate_per_s": 50 if realtime else None, "frames_per_request": 120,
            "budget_nodes": budget_nodes, "wall_s": rep["wall_s"],
            "node_substeps_per_s": rep["node_substeps"] / rep["wall_s"],
            "frame_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                 "max": float(lat.max())},
            "queue_s": {"mean": rep["mean_queue_s"], "max": rep["max_queue_s"]},
            "max_active": rep["max_active"], "iterations": rep["iterations"], "completed": rep["completed"]}


def capture_cloud(n, seed=1):
    """A synthetic capture: n Gaussians on a noisy ellipsoid shell (0.5 x 0.36 x
    0.6 m), random orientations, log-normal scales around 2 mm."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = (d * (1 + rng.normal(scale=0.02, size=(n, 1))) * [0.25, 0.18, 0.3]).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 3] < 0] *= -1
    s = np.exp(rng.normal(math.log(0.002), 0.5, size=(n, 3))).astype(np.float32)
    return pos, q.astype(np.float32), s


def torus_mesh(major, minor, around, ring):
    """A closed torus surface of 2 * around * ring triangles."""
    u, v = np.meshgrid(np.arange(around) / around * 2 * np.pi, np.arange(ring) / ring * 2 * np.pi, indexing="ij")
    r = major + minor * np.cos(u)
    p = np.stack([r * np.cos(v), minor * np.sin(u), r * np.sin(v)], -1).reshape(-1, 3).astype(np.float32)
    i, j = np.meshgrid(np.arange(around), np.arange(ring), indexing="ij")
    a, b = i * ring + j, ((i + 1) % around) * ring + j
    c, d = ((i + 1) % around) * ring + (j + 1) % ring, i * ring + (j + 1) % ring
    tri = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return p, tri.astype(np.uint32).reshape(-1)


def ingest_leg(cpu):
    """The device voxelisers on capture-sized inputs (best of 3 calls, host
    arrays in and out, PCIe included), and the oracle's single-core time on
    the same algorithm and inputs."""
    from vsim import ingest

    def best(fn, reps=3):
        fn()
        t = []
        for _ in range(reps):
            a = time.perf_counter()
            g = fn()
            t.append(time.perf_counter() - a)
        return min(t), g

    out = {}
    p, q, s = capture_cloud(1_000_000)
    t, g = best(lambda: ingest.voxelize_cloud(p, q, s, 0.004))
    out["cloud"] = {"gaussians": len(p), "element_size": 0.004, "grid": list(g.dims), "occupied": g.count(),
                    "gpu_ms": t * 1e3, "gaussians_per_s": len(p) / t}
    mp, mix = torus_mesh(0.3, 0.1, 256, 512)
    t, g = best(lambda: ingest.voxelize_mesh(mp, mix, 0.002))
    out["mesh"] = {"triangles": len(mix) // 3, "element_size": 0.002, "grid": list(g.dims), "occupied": g.count(),
                   "gpu_ms": t * 1e3, "triangles_per_s": len(mix) // 3 / t}
    t, g = best(lambda: ingest.voxelize_shape("torus", [0.3, 0.1], 0.002))
    out["shape"] = {"kind": "torus", "element_size": 0.002, "grid": list(g.dims), "gpu_ms": t * 1e3,
                    "cells_per_s": int(np.prod(g.dims)) / t}
    if cpu:
        from oracle import oracle as O
        k = len(p)
        a = time.perf_counter()
        O.voxelize_cloud(p, q, s, 0.004)
        tc = time.perf_counter() - a
        a = time.perf_counter()
        O.voxelize_mesh(mp, mix, 0.002)
        tm = time.perf_counter() - a
        a = time.perf_counter()
        O.voxelize_shape("torus", [0.3, 0.1], 0.002)
        ts = time.perf_counter() - a
        out["cpu_oracle"] = {"cores": 1, "kind": "port",
                             "cloud": {"gaussians": k, "ms": tc * 1e3, "gaussians_per_s": k / tc},
                             "mesh": {"ms": tm * 1e3, "triangles_per_s": len(mix) // 3 / tm},
                             "shape": {"ms": ts * 1e3}}
    return out


def workload_scenes(name):
    """The seeded scenes and config of a named workload."""
    from vsim import scenes
    if name == "batched256":
        return scenes.c4_scenes(256), {
            "workload": "256 scenes, lattice side U{8..48}, 10k-500k Gaussians each, E 1e5 nu 0.3 rho 1000, "
                        "h 0.016 m, bottom layer pinned, 10 substeps per 1/60 s frame", "scenes": 256}
    if name == "c1":
        return [scenes.c1_scene()], {"workload": "16^3 block, 50k Gaussians, bottom layer pinned, 10 substeps per "
                                                 "frame", "scenes": 1}
    if name == "c2":
        return [scenes.c2_scene()], {"workload": "64^3 block, 2M Gaussians, ground plane and sphere collider, 27-node "
                                                 "sinusoid handle, 10 substeps per frame", "scenes": 1}
    voxels = 1.0 if name == "c3_1voxel" else 2.0
    scs = [scenes.c3_scene(bucket_voxels=voxels)]
    piled = ", fallen into a pile (timed after 80 frames)" if name == "c3_pile" else ""
    return scs, {"workload": f"512 bodies of 8^3 cells (h 0.01 m) in a 1 m tank, node-node and node-wall contact, "
                             f"hash grid of {voxels:g}-voxel buckets, E log-uniform 1e4-1e6, {scs[0].substeps} "
                             f"substeps per frame{piled}", "scenes": 1}


PRE_ROLL = {"c3_pile": 80}  # untimed frames before a workload is timed


def timed_frames(world, ids, scs, substeps, steps, warmup, splats=True):
    """Graph-replayed batched frames: warmup, then `steps` frames between a
    synchronisation on each side. Returns wall seconds and per-frame ms from
    HIP events recorded between frames on the world's stream."""
    for f in range(warmup):
        set_targets(world, ids, scs, f)
        world.frame(substeps, splats=splats, graph=True)
    sync(world)
    marks = [world.record()]
    t0 = time.perf_counter()
    for f in range(steps):
        set_targets(world, ids, scs, warmup + f)
        world.frame(substeps, splats=splats, graph=True)
        marks.append(world.record())
    sync(world)
    wall = time.perf_counter() - t0
    return wall, np.array([world.elapsed_ms(marks[i], marks[i + 1]) for i in range(steps)])


def plain_frames(world, ids, scs, substeps, steps, splats=True):
    """The plain path: per-scene launches, one scene after another."""
    for sid in ids:
        world.frame_scene(sid, substeps, splats=splats)
    sync(world)
    t1 = time.perf_counter()
    for f in range(steps):
        set_targets(world, ids, scs, f)
        for sid in ids:
            world.frame_scene(sid, substeps, splats=splats)
    sync(world)
    return time.perf_counter() - t1


def workload_leg(name, steps, warmup, plain_steps, autotune):
    """One single-scene workload on the GPU: frame time, node-substeps/s,
    frame latency p50/p99 (graph-replayed frames), the plain per-scene
    launches beside it, and the per-pass profile of one instrumented frame."""
    scs, config = workload_scenes(name)
    nodes = sum(s.n_nodes for s in scs)
    substeps = scs[0].substeps
    splats = sum(s.n_splats for s in scs) > 0
    world = make_world(scs)
    ids = [world.add(sc) for sc in scs]
    for f in range(PRE_ROLL.get(name, 0)):
        world.frame(substeps, splats=False, graph=True)
        if f % 8 == 7:
            sync(world)  # a few frames in flight at a time
    sync(world)
    pick = json.loads(world.autotune(substeps))["best"] if autotune else None
    wall, frame_ms = timed_frames(world, ids, scs, substeps, steps, warmup, splats)
    prof = world.profile(substeps, splats=splats)
    plain_wall = plain_frames(world, ids, scs, substeps, plain_steps, splats) if plain_steps else None
    world.close()
    out = {"config": config, "nodes": nodes, "substeps_per_frame": substeps, "steps": steps,
           "ms_per_frame": wall * 1e3 / steps, "node_substeps_per_s": nodes * substeps * steps / wall,
           "frame_latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99))},
           "plain_node_substeps_per_s": nodes * substeps * plain_steps / plain_wall if plain_wall else None,
           "profile_ms": {k: {"total": v[0], "launches": v[1]} for k, v in prof.items() if v[1]},
           "autotune": pick}
    log(f"[bench] {name}: {out['ms_per_frame']:.3f} ms/frame, {out['node_substeps_per_s']:.4e} node-substeps/s, "
        f"p50 {out['frame_latency_ms']['p50']:.3f} p99 {out['frame_latency_ms']['p99']:.3f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batched256", choices=["batched256", "c1", "c2", "c3", "c3_1voxel", "c3_pile"])
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of the one-core CPU baseline, and again of the all-cores one")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of the autotuner's pick")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the single-scene workloads (16^3, 64^3 drag, 512 contact bodies) run beside the "
                         "256-scene one")
    ap.add_argument("--separate-node-passes", action="store_true",
                    help="run each substep's finalize and the next one's predict as two node passes (same "
                         "results; the default merges them inside a frame)")
    ap.add_argument("--separate-colours", action="store_true",
                    help="one launch per colour for every body (same results; the default solves the colours "
                         "of dense lattices in pairs, by rows of cells, one launch per pair)")
    ap.add_argument("--colour-launches", action="store_true",
                    help="small bodies of contact, stretch and drive scenes through the colour-pair launches too "
                         "(same results; the default runs each such body's eight colours in one launch, in LDS)")
    ap.add_argument("--branches", type=int, default=0,
                    help="1 or 2: run batched frames as that many chains of launches side by side and skip the "
                         "autotuner (same results; by default the autotuner picks)")
    ap.add_argument("--serve-requests", type=int, default=120,
                    help="scene requests of the serving legs (4k-500k nodes, 120 frames); 0 skips them")
    ap.add_argument("--serve-budget-nodes", type=float, default=6e6,
                    help="node capacity of the serving world (its HBM budget for node state)")
    ap.add_argument("--no-ingest", action="store_true", help="skip the voxeliser leg")
    ap.add_argument("--ingest-only", action="store_true",
                    help="run only the voxeliser leg (GPU side) and print its JSON")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a device-to-device copy of 1 GiB (5 times), the reference kernel "
                         "for scaling FETCH_SIZE/WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = (1 << 30) // 4
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)
        return
    if args.ingest_only:
        print(json.dumps(ingest_leg(cpu=False)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); scenes shard across GPUs with no "
              "communication, so the whole-job figure is N x the one-GPU figure by construction.", file=sys.stderr)
        sys.exit(2)

    t_gen = time.perf_counter()
    scs, config = workload_scenes(args.workload)
    nodes = sum(s.n_nodes for s in scs)
    cells = sum(s.n_cells for s in scs)
    splats = sum(s.n_splats for s in scs)
    substeps = scs[0].substeps
    log(f"[bench] {len(scs)} scenes: {nodes} nodes, {cells} cells, {splats} splats (generated in "
        f"{time.perf_counter() - t_gen:.1f} s)")

    # The CPU baseline runs first, so its worker processes fork before this
    # process touches the GPU.
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scs, args.cpu_budget, min(16, os.cpu_count() or 1))
        log(f"[bench] cpu baseline: {cpu['value']:.4e} node-substeps/s on {cpu['cores']} cores ({cpu['sample']})")

    merged = not args.separate_node_passes
    world = make_world(scs, separate_node_passes=int(args.separate_node_passes),
                       separate_colours=int(args.separate_colours), colour_launches=int(args.colour_launches),
                       branches=args.branches)
    ids = [world.add(sc) for sc in scs]
    for f in range(PRE_ROLL.get(args.workload, 0)):
        world.frame(substeps, splats=False, graph=True)
        if f % 8 == 7:
            sync(world)  # a few frames in flight at a time
    sync(world)
    log("[bench] scenes resident")
    tune = None
    # The autotuner may pick either schedule; an explicit schedule flag keeps it.
    if not (args.no_autotune or args.separate_colours or args.separate_node_passes or args.branches):
        tune = json.loads(world.autotune(substeps))
        for c in tune["candidates"]:
            log(f"[bench] autotune {c}")
        log(f"[bench] autotune best {tune['best']}")

    # ---- batched, graph-replayed frames (the timed region)
    wall, frame_ms = timed_frames(world, ids, scs, substeps, args.steps, args.warmup, splats > 0)
    value = nodes * substeps * args.steps / wall
    log(f"[bench] batched: {wall * 1e3 / args.steps:.3f} ms/frame, {value:.4e} node-substeps/s")

    # ---- per-pass profile of one batched frame, HIP events around each
    # launch on the world's stream (the stream the kernels run on)
    prof = world.profile(substeps, splats=splats > 0)
    log(f"[bench] profile: " + ", ".join(f"{k} {v[0]:.3f} ms/{v[1]}" for k, v in prof.items() if v[1]))
    # ---- the colour passes alone: rounds of a substep's colour launches
    # back to back, HIP events only at the two ends (state restored after)
    el_ms, el_n = world.time_pass("elastic", substeps)
    burst_us = el_ms / el_n * 1e3
    per_sweep = el_n // substeps
    pairs = per_sweep == 4
    el_kernel = "k_elastic_rows" if pairs else "k_elastic"
    colour_model_bytes = cells * (ELASTIC_BYTES_PER_CELL - 4 if pairs else ELASTIC_BYTES_PER_CELL) / per_sweep
    if pairs and prof["pair"][1]:
        # The dominant kernel as the frame runs it: the colour-pair launches
        # with no node pass on board, timed inside the frame.
        kernel_bytes = plain_pair_bytes(scs, substeps, merged)
        kernel_us = prof["pair"][0] / prof["pair"][1] * 1e3
        burst_bytes = pair_launch_bytes(scs)
    else:
        kernel_bytes = burst_bytes = colour_model_bytes
        kernel_us = prof["elastic"][0] / prof["elastic"][1] * 1e3
    achieved = kernel_bytes / (kernel_us * 1e-6) / 1e9
    burst_gbs = burst_bytes / (burst_us * 1e-6) / 1e9
    log(f"[bench] colour pair in frame: {kernel_us:.2f} us avg, {achieved:.0f} GB/s algorithmic; back to back "
        f"{burst_us:.2f} us, {burst_gbs:.0f} GB/s")
    counted = args.workload == "batched256" and not args.separate_colours
    valu = elastic_valu_instructions(el_kernel) if counted else None
    valu_frac = valu / (kernel_us * 1e-6 * VALU_ISSUE_PER_S) if valu else None
    valu_busy, valu_lanes = elastic_valu_busy(el_kernel) if counted else (None, None)
    hbm_frac = achieved / HBM_PEAK_GBS
    # ---- the splat pass alone, covariance output and with the engine format
    # (rotation + axis lengths by four Jacobi sweeps) as well
    from vsim import native
    splat_us = {}
    if splats:
        for label, mode in (("cov", native.SPLAT_COV), ("cov+axes", native.SPLAT_COV | native.SPLAT_AXES)):
            world.frame(0, mode, False)
            ms, n = world.time_pass("splat", 3)
            splat_us[label] = ms / n * 1e3 if n else None
        log(f"[bench] splat pass: " + ", ".join(f"{k} {v:.1f} us" for k, v in splat_us.items() if v))

    # ---- plain per-scene launches (the reference's dispatch shape)
    plain = None
    if not args.no_plain and args.plain_steps > 0:
        plain_wall = plain_frames(world, ids, scs, substeps, args.plain_steps, splats > 0)
        plain = nodes * substeps * args.plain_steps / plain_wall
        log(f"[bench] plain: {plain_wall * 1e3 / args.plain_steps:.3f} ms/frame, {plain:.4e} node-substeps/s")
    world.close()

    legs = None
    if args.workload == "batched256" and not args.no_legs:
        legs = {name: workload_leg(name, 20, 3, 3, not args.no_autotune) for name in ("c1", "c2", "c3", "c3_pile")}

    serving = None
    if args.serve_requests > 0:
        serving = {"poisson_50_per_s": serve_leg(args.serve_requests, int(args.serve_budget_nodes), realtime=True),
                   "saturated": serve_leg(args.serve_requests, int(args.serve_budget_nodes), realtime=False)}
        log(f"[bench] serving: {serving}")

    ingest_out = None
    if not args.no_ingest:
        ingest_out = ingest_leg(cpu=not args.no_cpu)
        log(f"[bench] ingest: {ingest_out}")

    # ---- the whole frame as the timed loop runs it (one or two chains of
    # launches side by side): the colour launches' and the splat pass's
    # algorithmic bytes over the frame time; node passes and contact are left
    # out, so this is a lower bound of the frame's rate.
    frame_bytes = substeps * per_sweep * (burst_bytes if pairs else colour_model_bytes) + splats * SPLAT_BYTES
    frame_s = wall / args.steps
    frame_roof = {"bytes": frame_bytes, "ms": frame_s * 1e3, "achieved": frame_bytes / frame_s / 1e9,
                  "frac": frame_bytes / frame_s / 1e9 / HBM_PEAK_GBS,
                  "branches": tune["best"].get("branches") if tune else (args.branches or None),
                  "note": "colour launches and splat pass only, over the timed frame"}
    roofline = {
        "bound": "valu" if (valu_busy or valu_frac or 0) > hbm_frac else "hbm",
        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_frac,
        "traffic": elastic_traffic_bytes(el_kernel) if counted else None,
        "kernel": el_kernel + (" (colour pair, no node pass on board)" if pairs else " (colour pass)"),
        "bytes_per_launch": kernel_bytes,
        "avg_launch_us": kernel_us,
        "timing": "HIP events around each launch of one batched frame on the world's stream",
        "back_to_back": {"avg_launch_us": burst_us, "bytes_per_launch": burst_bytes, "achieved": burst_gbs,
                         "frac": burst_gbs / HBM_PEAK_GBS},
        "colour_model_bytes_per_launch": colour_model_bytes,
        "valu_issue_frac": valu_frac,
        "valu_issue_note": "wave instructions per CU per cycle at 2.4 GHz, this run's launch time; gfx950 "
                           "dual-issues VALU instructions, so 1.0 is not a ceiling",
        "valu_busy_frac": valu_busy,
        "valu_lanes_active_frac": valu_lanes,
        "valu_wave_instructions_per_launch": valu,
        "counters_source": profile_source("batched256_valu.md") if counted else None,
        "frame": frame_roof,
    }
    out = {
        "metric": "node-substeps/s, batched vs plain, at \u22641e-5 rel error; p50/p99 frame latency",
        "value": value,
        "unit": "node-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded lattices and Gaussians, resident in HBM)",
        "config": config,
        "plain": plain,
        "speedup_vs_plain": value / plain if plain else None,
        "frame_latency_ms": {"p50": float(np.percentile(frame_ms, 50)), "p99": float(np.percentile(frame_ms, 99))},
        "nodes": nodes, "cells": cells, "splats": splats, "substeps_per_frame": substeps,
        "profile_ms": {k: {"total": v[0], "launches": v[1]} for k, v in prof.items() if v[1]},
        "passes": pass_rates(prof, scs, substeps, splats, burst_bytes, pairs, merged),
        "roofline": roofline,
        "splat_pass_us": splat_us,
        "workloads": legs,
        "serving": serving,
        "ingest": ingest_out,
        "cpu_baseline": cpu,
    }
    if tune:
        out["autotune"] = tune["best"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
