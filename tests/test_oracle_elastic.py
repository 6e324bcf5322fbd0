"""Pins the oracle's elastic projection (project.wgsl:75-153 + xpbdCouple.wgsl)
against the known answers of the reference's elastic.test.js, which the
closed forms of the stable Neo-Hookean block fix: rest is a fixed point,
momentum is conserved, an inert corner does not move, Kelvin-Voigt damping
ignores rigid motion and removes energy, and the collapse barrier sits where
collapseScale says."""
import math

import numpy as np
import pytest

from vsim import scenes

H = 0.03125             # elastic.test.js:186 (h and 1/h exact in f32)
DT = 1.0 / (60 * 20)    # elastic.test.js:189
RATIOS = [0.25, 0.3, 0.35, 0.45, 0.49]
KUHN_TETS = [[0, 1, 3, 7], [0, 5, 1, 7], [0, 3, 2, 7], [0, 2, 6, 7], [0, 4, 5, 7], [0, 6, 4, 7]]
CORNERS = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], float)


def rest_inverse(t, h=H):
    """Inverse rest shape of Kuhn tet t, column-major, from numpy's inverse."""
    i = KUHN_TETS[t]
    Dm = np.stack([(CORNERS[i[k]] - CORNERS[i[0]]) * h for k in (1, 2, 3)], axis=1)
    return np.linalg.inv(Dm).T.reshape(9)  # column-major = row-major of the transpose


def rest_tet(nu, t=0):
    m = scenes.Material(poisson_ratio=nu)
    vol = scenes.tet_rest_volume(H)
    dev, hyd, gamma = scenes.element_compliance(m, vol)
    node_mass = m.density * vol / 4
    mu, lam = scenes.lame_parameters(m)
    p = CORNERS[KUHN_TETS[t]] * H
    return dict(p=p, B=rest_inverse(t), w=np.full(4, 1 / node_mass), c=(dev / DT ** 2, hyd / DT ** 2, gamma),
                mu=mu, lam=lam, vol=vol)


def project(o, s, p, prev=None, relax=0.0, w=None):
    return o.project_tet(p, p if prev is None else prev, s["w"] if w is None else w, s["B"], *s["c"], relax)


def moved(a, b):
    return float(np.max(np.linalg.norm(np.asarray(b) - np.asarray(a), axis=1)))


def volume(p):
    return np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1)) / 6


def run(o, s, p, substeps, relax=0.0):
    pos = np.array(p, float)
    vel = np.zeros_like(pos)
    for _ in range(substeps):
        prev = pos.copy()
        pos = pos + vel * DT
        pos = project(o, s, pos, prev, relax)
        vel = (pos - prev) / DT
    return pos


def constraints(s, p):
    F = np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1) @ s["B"].reshape(3, 3).T
    return np.linalg.norm(F), np.linalg.det(F) - s["c"][2]


def stored(s, p):
    def at(q):
        cd, ch = constraints(s, q)
        return 0.5 * s["mu"] * s["vol"] * cd * cd + 0.5 * s["lam"] * s["vol"] * ch * ch
    return at(p) - at(s["p"])


def test_rest_is_a_fixed_point_at_every_ratio(oracle):
    for nu in RATIOS:
        s = rest_tet(nu)
        assert moved(s["p"], project(oracle, s, s["p"])) < 1e-9, nu


def test_every_kuhn_tet_holds_rest(oracle):
    for t in range(6):
        s = rest_tet(0.25, t)
        assert moved(s["p"], project(oracle, s, s["p"])) < 1e-9, t


def test_f32_rest_fixed_point(oracle):
    # h = 1/32 keeps B exact in f32, so the f32 build is at rest to f32 noise.
    for nu in RATIOS:
        s = rest_tet(nu)
        out = oracle.project_tet(s["p"], s["p"], s["w"], s["B"], *s["c"], 0.0, precision="f32")
        assert moved(s["p"], out) < 1e-7, nu


def test_free_tet_keeps_its_volume_for_a_second(oracle):
    s = rest_tet(0.25)
    assert volume(run(oracle, s, s["p"], 1200)) / volume(s["p"]) == pytest.approx(1, abs=5e-7)


def test_squashed_tet_rings_about_rest_volume(oracle):
    s = rest_tet(0.25)
    ratio = volume(run(oracle, s, s["p"] * 0.9, 1200)) / volume(s["p"])
    assert 0.6 < ratio < 1.4


def test_momentum_is_conserved(oracle):
    s = rest_tet(0.3)
    sq = s["p"] * np.array([1, 0.7, 1])
    out = project(oracle, s, sq)
    assert np.all(np.abs((out - sq).sum(axis=0)) < 1e-12)


def test_collapsed_tet_stays_finite(oracle):
    s = rest_tet(0.35)
    out = project(oracle, s, np.zeros((4, 3)))
    assert np.all(np.isfinite(out))


def test_inert_corner_does_not_move_and_changes_the_others(oracle):
    s = rest_tet(0.3)
    sq = s["p"] * np.array([1, 0.85, 1])
    driven = s["w"].copy()
    driven[0] = 0
    out = project(oracle, s, sq, w=driven)
    assert moved(sq[:1], out[:1]) == 0
    free = project(oracle, s, sq)
    for i in range(1, 4):
        assert moved(free[i:i + 1], out[i:i + 1]) > 1e-9


RELAX = 0.04 / DT


def test_kelvin_voigt_is_a_no_op_at_zero_relaxation(oracle):
    s = rest_tet(0.35)
    sq = s["p"] * np.array([1, 0.8, 1])
    bare = project(oracle, s, sq)
    damped = project(oracle, s, sq, prev=s["p"], relax=0.0)
    assert moved(bare, damped) == 0


def test_kelvin_voigt_keeps_rest_and_ignores_translation(oracle):
    for nu in RATIOS:
        s = rest_tet(nu)
        assert moved(s["p"], project(oracle, s, s["p"], relax=RELAX)) < 1e-9
    s = rest_tet(0.35)
    shifted = s["p"] + np.array([0.004, -0.011, 0.002])
    assert moved(shifted, project(oracle, s, shifted, prev=s["p"], relax=RELAX)) < 1e-9


def test_kelvin_voigt_ignores_spin_to_second_order(oracle):
    s = rest_tet(0.35)
    a = 0.002
    R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
    spun = s["p"] @ R.T
    sheared = s["p"] + np.outer(s["p"][:, 1] * a, [1, 0, 0])
    shear = moved(sheared, project(oracle, s, sheared, prev=s["p"], relax=RELAX))
    assert moved(spun, project(oracle, s, spun, prev=s["p"], relax=RELAX)) < shear / 100


def test_kelvin_voigt_lowers_the_peak_deformation(oracle):
    s = rest_tet(0.35)

    def peak(relax):
        pos = s["p"].copy()
        vel = np.stack([s["p"][:, 1] * 2, np.zeros(4), np.zeros(4)], axis=1)
        worst = 0.0
        for _ in range(200):
            prev = pos.copy()
            pos = pos + vel * DT
            pos = project(oracle, s, pos, prev, relax)
            vel = (pos - prev) / DT
            worst = max(worst, stored(s, pos))
        return worst

    assert peak(0.012 / DT) < peak(0) * 0.98


def collapse_scale(gamma):
    """Root of s^3 + s^2 + s = gamma - 1 by bisection (compliance.js:78-90)."""
    lo, hi = 0.0, 4.0
    for _ in range(40):
        m = (lo + hi) / 2
        if m ** 3 + m ** 2 + m < gamma - 1:
            lo = m
        else:
            hi = m
    return (lo + hi) / 2


def test_collapse_scale_known_values():
    assert collapse_scale(4) == pytest.approx(1, abs=1e-6)
    assert collapse_scale(2) == pytest.approx(0.5437, abs=5e-4)
    assert collapse_scale(1) == pytest.approx(0, abs=1e-6)
    assert collapse_scale(5) > 1


def test_barrier_is_where_collapse_scale_says(oracle):
    # elastic.test.js:539-567: bisect for the largest uniform squash a free
    # tetrahedron survives and compare with the closed-form barrier.
    for nu in (0.25, 0.3, 0.35):
        s = rest_tet(nu)
        rest = volume(s["p"])

        def survives(f):
            r = volume(run(oracle, s, s["p"] * f, 3000)) / rest
            return 0.5 < r < 2

        ok, lost = 1.0, 0.05
        for _ in range(18):
            mid = (ok + lost) / 2
            if survives(mid):
                ok = mid
            else:
                lost = mid
        assert abs(ok - collapse_scale(s["c"][2])) < 0.5 * 10 ** -2, nu


def test_compliance_constants_reduce_to_lame():
    m = scenes.Material(poisson_ratio=0.3)
    mu, lam = scenes.lame_parameters(m)
    dev, hyd, _ = scenes.element_compliance(m, scenes.tet_rest_volume(H))
    assert (dev * mu) / (hyd * lam) == pytest.approx(1, abs=1e-12)
    assert dev * mu == pytest.approx(1 / scenes.tet_rest_volume(H), rel=1e-12)


def test_lattice_block_at_rest_without_gravity_stays_put(oracle):
    # The cell path (gather, 6 tets in order, scatter) on a free 3^3 block:
    # no gravity, no damping -> nothing moves beyond f32 rounding of B.
    sc = scenes.jelly_block(3, h=H, pinned=False, tank_margin=1.0)
    sc.gravity = (0.0, 0.0, 0.0)
    o = oracle.OracleScene(sc, "f64")
    before = o.p.copy()
    o.substeps(20)
    assert np.max(np.abs(o.p - before)) < 1e-12


def test_lattice_block_conserves_momentum(oracle):
    # A sheared free block: one colour sweep moves nodes but not their
    # mass-weighted centre (equal masses in a colour's cells).
    sc = scenes.jelly_block(4, h=H, pinned=False, tank_margin=1.0)
    sc.gravity = (0.0, 0.0, 0.0)
    o = oracle.OracleScene(sc, "f64")
    o.p[0] += 0.2 * (o.p[1] - o.p[1].min())
    o.q[:] = o.p
    mass = np.where(o.w > 0, 1 / np.where(o.w > 0, o.w, 1), 0)
    before = (o.p * mass).sum(axis=1)
    for c in range(8):
        o.elastic(c)
    after = (o.p * mass).sum(axis=1)
    assert np.all(np.abs(after - before) < 1e-12)
    assert np.max(np.abs(o.p - o.q)) > 1e-4


# ---- the fused evaluation order the GPU uses, pinned against the literal
# WGSL formulas (project.wgsl:75-153) above.

def kuhn(o, s, p, t, prev=None, relax=0.0, w=None, literal=False, precision="f64"):
    return o.project_kuhn_tet(p, p if prev is None else prev, s["w"] if w is None else w, t, H, *s["c"], relax,
                              literal=literal, precision=precision)


@pytest.mark.parametrize("literal", [True, False])
def test_both_forms_hold_rest_for_every_tet(oracle, literal):
    for t in range(6):
        for nu in RATIOS:
            s = rest_tet(nu, t)
            assert moved(s["p"], kuhn(oracle, s, s["p"], t, literal=literal)) < 1e-9


def test_fused_form_equals_the_literal_formulas_in_f64(oracle):
    rng = np.random.default_rng(12)
    worst = 0.0
    for trial in range(600):
        t = trial % 6
        s = rest_tet(RATIOS[trial % len(RATIOS)], t)
        p = s["p"] + rng.normal(scale=0.2 * H, size=(4, 3))
        prev = p - rng.normal(scale=0.05 * H, size=(4, 3))
        w = s["w"] * rng.uniform(0.2, 2.0, 4)
        if trial % 7 == 0:
            w[trial % 4] = 0.0
        relax = [0.0, 0.012 / DT][trial % 2]
        a = kuhn(oracle, s, p, t, prev, relax, w, literal=True)
        b = kuhn(oracle, s, p, t, prev, relax, w, literal=False)
        worst = max(worst, np.max(np.abs(a - b)) / np.max(np.abs(a - p)))
    assert worst < 1e-9


def test_fused_form_conserves_momentum_and_leaves_inert_corners(oracle):
    s = rest_tet(0.3, 2)
    sq = s["p"] * np.array([1, 0.7, 1])
    out = kuhn(oracle, s, sq, 2)
    assert np.all(np.abs((out - sq).sum(axis=0)) < 1e-12)
    w = s["w"].copy()
    w[1] = 0
    assert moved(sq[1:2], kuhn(oracle, s, sq, 2, w=w)[1:2]) == 0


def test_fused_substep_is_within_the_parity_tolerance_of_the_literal_one(oracle):
    # One substep of the 16^3 block: the GPU's f32 order against the literal
    # WGSL formulas in f32 (<= 1e-5 relative), and both against f64.
    sc = scenes.c1_scene(splats=0)
    fused = oracle.OracleScene(sc, "f32", literal=False)
    lit = oracle.OracleScene(sc, "f32", literal=True)
    ref = oracle.OracleScene(sc, "f64", literal=True)
    for o in (fused, lit, ref):
        o.frame(10)
    scale = np.max(np.abs(ref.p))
    assert np.max(np.abs(fused.p - lit.p)) / scale <= 1e-5
    e_f = np.max(np.abs(fused.p - ref.p)) / scale
    e_l = np.max(np.abs(lit.p - ref.p)) / scale
    print(f"f32 fused vs f64: {e_f:.3e}; f32 literal vs f64: {e_l:.3e}")
    assert e_f < 1e-5 and e_l < 1e-5


def test_fused_form_finds_the_same_collapse_barrier(oracle):
    # The barrier bisection of elastic.test.js:539-567, run on the fused form.
    s = rest_tet(0.3, 0)
    rest = volume(s["p"])

    def run_fused(p, substeps):
        pos = np.array(p, float)
        vel = np.zeros_like(pos)
        for _ in range(substeps):
            prev = pos.copy()
            pos = kuhn(oracle, s, pos + vel * DT, 0, prev)
            vel = (pos - prev) / DT
        return pos

    ok, lost = 1.0, 0.05
    for _ in range(14):
        mid = (ok + lost) / 2
        r = volume(run_fused(s["p"] * mid, 3000)) / rest
        if 0.5 < r < 2:
            ok = mid
        else:
            lost = mid
    assert abs(ok - collapse_scale(s["c"][2])) < 0.5 * 10 ** -2


def _stretch_block():
    sc = scenes.jelly_block(8, pinned=True, name="stretch")
    sc.bodies[0].material.max_stretch = 1.02
    sc.drive = scenes.handle_drive(sc, seed=5, law=0, amplitude=0.3, frequency=2.0)
    return sc


def _drag_block():
    sc = scenes.jelly_block(12, pinned=False, tank_margin=0.3, name="drag")
    sc.tank_min = (sc.tank_min[0], 0.0, sc.tank_min[2])
    sc.drive = scenes.handle_drive(sc, seed=3, law=0)
    return sc


WIDE_CASES = {
    "contact_stack": (lambda: scenes.contact_stack(), 3),
    "hard_drive": (_drag_block, 3),
    "stretch": (_stretch_block, 3),
    "bodies64_contact": (lambda: scenes.c3_scene(bodies=64), 2),
    # 64 bodies packed into a 0.4 m tank: hundreds of contacts a substep,
    # where the sum of a record's corrections row by row departs most from
    # collide.wgsl's one-after-another sum
    "bodies64_packed": (lambda: scenes.c3_scene(bodies=64, box=0.4), 1),
}


@pytest.mark.parametrize("case", sorted(WIDE_CASES))
def test_fused_frames_track_the_literal_formulas_where_the_branches_fire(oracle, case):
    # Scenes where contact, the hard drive, the tensile sweeps and fast
    # collisions exercise what a resting block does not. The GPU's fused f32
    # order against the literal WGSL formulas in f32, relative to the largest
    # coordinate: within 1e-5 after one substep and after whole frames; and
    # against f64, one substep within 1e-5 for both forms, while over whole
    # frames (where contact between fast bodies lets rounding grow) the fused
    # form stays as close to f64 as the literal one.
    make, frames = WIDE_CASES[case]
    sc = make()
    fused = oracle.OracleScene(sc, "f32", literal=False)
    lit = oracle.OracleScene(sc, "f32", literal=True)
    ref = oracle.OracleScene(sc, "f64", literal=True)
    start = np.asarray(sc.arrays()["pos"]).T
    for o in (fused, lit, ref):
        if sc.drive is not None:
            o.set_target(*sc.drive.frame_target(0, sc.substeps * sc.dt))
        o.substep()
    scale = np.max(np.abs(ref.p))
    one = (np.max(np.abs(fused.p - lit.p)) / scale, np.max(np.abs(fused.p - ref.p)) / scale,
           np.max(np.abs(lit.p - ref.p)) / scale)
    fused = oracle.OracleScene(sc, "f32", literal=False)
    lit = oracle.OracleScene(sc, "f32", literal=True)
    ref = oracle.OracleScene(sc, "f64", literal=True)
    for _ in range(frames):
        for o in (fused, lit, ref):
            o.frame()
    scale = np.max(np.abs(ref.p))
    d = np.max(np.abs(fused.p - lit.p)) / scale
    e_f = np.max(np.abs(fused.p - ref.p)) / scale
    e_l = np.max(np.abs(lit.p - ref.p)) / scale
    moved_by = np.max(np.abs(ref.p - start)) / scale
    print(f"{case}: one substep: fused vs literal {one[0]:.3e}, vs f64 {one[1]:.3e} / {one[2]:.3e}; "
          f"{frames} frames: fused vs literal {d:.3e}, vs f64 {e_f:.3e} / {e_l:.3e}; motion {moved_by:.3e}")
    assert moved_by > 1e-3
    assert max(one) <= 1e-5
    assert d <= 1e-5
    assert e_f <= 1.5 * e_l + 1e-6


@pytest.mark.parametrize("shape", ["collapsed", "flat", "inverted", "needle"])
def test_both_forms_agree_on_degenerate_tets(oracle, shape):
    # The branches a resting block never takes: F = 0 (every corner on one
    # point: C_D below its 1e-9 floor, zero cofactors), a flat tet (det F = 0),
    # an inverted one (det F < 0) and a needle (three corners on one point).
    # Literal and fused agree to 1e-9 of the motion in f64, the result is
    # finite, and F = 0 moves nothing.
    for t in range(6):
        s = rest_tet(0.3, t)
        p = s["p"].copy()
        if shape == "collapsed":
            p[:] = p[0]
        elif shape == "flat":
            p[:, 2] = p[0, 2]
        elif shape == "inverted":
            p[3] = 2 * p[0] - p[3]
        else:
            p[1:] = p[1]
        prev = p - 0.01 * H
        for relax in (0.0, 0.012 / DT):
            a = kuhn(oracle, s, p, t, prev, relax, literal=True)
            b = kuhn(oracle, s, p, t, prev, relax, literal=False)
            assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
            motion = np.max(np.abs(a - p))
            assert np.max(np.abs(a - b)) <= 1e-9 * max(motion, H), (shape, t, relax)
            if shape == "collapsed":
                assert np.array_equal(a, p) and np.array_equal(b, p)
            a32 = kuhn(oracle, s, p, t, prev, relax, literal=True, precision="f32")
            b32 = kuhn(oracle, s, p, t, prev, relax, literal=False, precision="f32")
            assert np.max(np.abs(a32 - b32)) <= 1e-5 * max(np.max(np.abs(a)), H), (shape, t, relax)
