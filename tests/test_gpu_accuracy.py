"""Error of the GPU's f32 frames against the f64 oracle: positions, the
constraint residuals C_D = |F| and C_H = det F - gamma of every Kuhn
tetrahedron (project.wgsl:75-110), and the splat covariances
(covariance.wgsl:18-42). The GPU is held to the f32 oracle bit for bit by the
parity tests; these state how far that f32 arithmetic sits from f64 after
whole frames. The bounds are generous; the printed figures are the report."""
import numpy as np
import pytest

from vsim import native, scenes

pytestmark = pytest.mark.gpu

# Kuhn tetrahedra (kuhn.js:28-35), corner c = (c & 1, (c >> 1) & 1, (c >> 2) & 1).
TETS = [(0, 1, 3, 7), (0, 5, 1, 7), (0, 3, 2, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 6, 4, 7)]


def corner(c):
    return np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], float)


def residuals(pos, body):
    """(C_D, C_H) of every tetrahedron of a dense body, in f64, from node
    positions `pos` (3, N) of the body: F = Ds Dm^-1 / h."""
    lat = body.lattice
    cx, cy, cz = lat.cell_dims
    nx, ny = cx + 1, cy + 1
    h = lat.h
    gamma = body.constants()[2]
    k, j, i = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
    base = (i + nx * (j + ny * k)).reshape(-1)
    p = pos.astype(np.float64)
    out = []
    for tet in TETS:
        dm = np.stack([corner(c) - corner(tet[0]) for c in tet[1:]], axis=1)
        inv = np.linalg.inv(dm) / h
        ids = [base + int(corner(c)[0]) + nx * (int(corner(c)[1]) + ny * int(corner(c)[2])) for c in tet]
        ds = np.stack([p[:, ids[a]] - p[:, ids[0]] for a in (1, 2, 3)], axis=2)  # (3, cells, 3)
        F = np.einsum("rca,ab->crb", ds, inv)
        out.append((np.sqrt((F ** 2).sum(axis=(1, 2))), np.linalg.det(F) - gamma))
    cd = np.stack([o[0] for o in out])
    ch = np.stack([o[1] for o in out])
    return cd, ch


def run_gpu(sc, frames):
    nodes, cells, splats = sc.n_nodes, sc.n_cells, sc.n_splats
    driven = len(sc.drive.nodes) if sc.drive is not None else 0
    w = native.World(max_scenes=2, max_bodies=2, max_nodes=nodes + 16, max_cells=cells + 16,
                     max_splats=splats + 16, max_driven=driven + 16)
    sid = w.add(sc)
    for f in range(frames):
        if sc.drive is not None:
            w.set_target(sid, *sc.drive.frame_target(f, sc.substeps * sc.dt))
        w.frame(sc.substeps, True, f > 0)
    pos, _ = w.read_nodes(sid)
    mean, cov = w.read_splats(sid)
    w.close()
    return pos, mean, cov


def run_oracle(sc, frames, precision):
    from oracle.oracle import OracleScene
    o = OracleScene(sc, precision)
    for f in range(frames):
        if sc.drive is not None:
            o.set_target(*sc.drive.frame_target(f, sc.substeps * sc.dt))
        o.frame(sc.substeps, splats=True)
    return o.p.copy(), o.out_mean.copy(), o.out_cov.copy()


def rel(a, ref):
    d = np.abs(a.astype(np.float64) - ref.astype(np.float64))
    scale = float(np.max(np.abs(ref)))
    return float(d.max()) / scale, float(np.sqrt((d ** 2).mean())) / scale


@pytest.mark.parametrize("case", ["block16", "drag24"])
def test_error_against_f64(case):
    if case == "block16":
        sc = scenes.c1_scene(splats=20_000)
    else:
        sc = scenes.jelly_block(24, splats=20_000, seed=21, pinned=False, tank_margin=0.3, name="drag24")
        sc.drive = scenes.handle_drive(sc, seed=22, law=0)
    frames = 10
    print(f"case {case}: {sc.n_nodes} nodes, {frames} frames of {sc.substeps} substeps", flush=True)
    pos, mean, cov = run_gpu(sc, frames)
    p64, m64, c64 = run_oracle(sc, frames, "f64")
    p32, _, _ = run_oracle(sc, frames, "f32")
    assert np.array_equal(pos, p32.astype(np.float32))  # the parity bar: the f32 oracle, bit for bit
    ep = rel(pos, p64)
    cd, ch = residuals(pos, sc.bodies[0])
    cd64, ch64 = residuals(p64, sc.bodies[0])
    ed = float(np.max(np.abs(cd - cd64))) / float(np.mean(np.abs(cd64)))
    eh = float(np.max(np.abs(ch - ch64))) / float(np.mean(np.abs(ch64)))
    em = rel(mean, m64)
    ec = rel(cov, c64)
    print(f"{case} vs f64: positions max {ep[0]:.2e} rms {ep[1]:.2e}; C_D max {ed:.2e}, C_H max {eh:.2e} "
          f"(of their mean magnitude); splat means max {em[0]:.2e}; covariances max {ec[0]:.2e} rms {ec[1]:.2e}")
    assert ep[0] < 1e-4 and em[0] < 1e-4
    assert ed < 1e-3 and eh < 1e-2
    assert ec[0] < 1e-3
