"""Host-side scene assembly: materials, lattices, bodies, the five workloads.

Everything here runs once per scene on the CPU and fixes the order of every GPU
input, exactly as the reference's host JS does:

* material constants: ``lameParameters`` (body/material.js:101-108),
  ``elementCompliance``/``gammaOf`` (sim/constraints/compliance.js:30-51),
  ``waveSpeed`` (material.js:137-140), ``stableStepDuration``
  (sim/solver/timestep.js:38-45);
* the lattice: ``buildLattice`` (sim/discretize/lattice/lattice.js:166-239) --
  dense node grid, f32 rest positions, corner-lumped masses accumulated in f32,
  cells grouped by colour in z, y, x scan order; ``surfaceNodes`` (:128-158);
* contact records body by body (sim/constraints/bodyContact.js:72-85) with the
  half-cell radius of voxelLattice.js:17.

Synthetic inputs are seeded (numpy PCG64); a scene is a pure function of its
parameters and seed.
"""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import native

COLOR_COUNT = 8
CONTACT_RADIUS_CELLS = 0.5       # voxelLattice.js:17
POISSON_MIN, POISSON_MAX = 0.25, 0.49   # material.js:66
TARGET_CFL = 0.75                # timestep.js:17
MAX_STEP_DURATION = 1.0 / 240.0  # timestep.js:23
MAX_STEPS_PER_FRAME = 24         # timestep.js:30
RESPONSE_TIME = 0.05             # kinematicDrive.js:65
MAX_DRIVEN = 4096                # kinematicDrive.js:49


@dataclass
class Material:
    """Material (body/material.js:7-27); defaults are DEFAULT_MATERIAL (:69-77)."""
    youngs_modulus: float = 1.2e5
    poisson_ratio: float = 0.40
    density: float = 1000.0
    friction: float = 0.4
    damping: float = 0.6
    relaxation: float = 0.012
    max_stretch: float = 0.0


def lame_parameters(m):
    nu = min(max(m.poisson_ratio, POISSON_MIN), POISSON_MAX)
    e = max(m.youngs_modulus, 1.0)
    return e / (2 * (1 + nu)), (e * nu) / ((1 + nu) * (1 - 2 * nu))


def tet_rest_volume(h):
    return (1.0 / 6.0) * h ** 3


def element_compliance(m, rest_volume):
    mu, lam = lame_parameters(m)
    return 1.0 / (mu * rest_volume), 1.0 / (lam * rest_volume), 1.0 + mu / lam


def wave_speed(m):
    mu, lam = lame_parameters(m)
    return math.sqrt((lam + 2 * mu) / m.density)


def stable_step_duration(wave_rate):
    if not math.isfinite(wave_rate) or wave_rate < 0:
        raise ValueError(f"waveRate must be a finite non-negative number, got {wave_rate}")
    return MAX_STEP_DURATION if wave_rate == 0 else min(MAX_STEP_DURATION, TARGET_CFL / wave_rate)


def worst_case_bucket_occupancy(bucket_size, spacing):
    """Bound on one body's nodes in one bucket (broadphase.js:82-85): cubes of
    edge spacing/sqrt(3) hold at most one node each."""
    return math.ceil(math.sqrt(3) * bucket_size / spacing) ** 3


def cell_color(i, j, k):
    return (i & 1) | ((j & 1) << 1) | ((k & 1) << 2)


@dataclass
class Lattice:
    origin: tuple
    h: float
    cell_dims: tuple
    node_dims: tuple
    rest: np.ndarray          # (N, 3) float32
    inv_mass: np.ndarray      # (N,) float32, 0 where inactive
    cells: np.ndarray         # uint32 linear cell ids grouped by colour
    color_counts: np.ndarray  # (8,)
    mass: float
    volume: float
    occupancy: np.ndarray = None  # (cz, cy, cx) bool, the grid it was built from
    density: float = None

    @property
    def node_count(self):
        return int(np.prod(self.node_dims))


def build_lattice(occupancy, origin, h, density):
    """buildLattice (lattice.js:166-239) over an occupancy array of shape
    (cz, cy, cx), nonzero = occupied."""
    occ = np.ascontiguousarray(occupancy, dtype=bool)
    cz, cy, cx = occ.shape
    nx, ny, nz = cx + 1, cy + 1, cz + 1
    # node index i + nx * (j + ny * k): build in (k, j, i) order.
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rest = np.empty((nz * ny * nx, 3), np.float32)
    rest[:, 0] = (origin[0] + i3.reshape(-1) * h)
    rest[:, 1] = (origin[1] + j3.reshape(-1) * h)
    rest[:, 2] = (origin[2] + k3.reshape(-1) * h)
    # Incident occupied cells per node; masses accumulate in a Float32Array,
    # one f64 add of the corner mass rounded to f32 per incident cell.
    count = np.zeros((nz, ny, nx), np.int32)
    o = occ.astype(np.int32)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                count[dz:dz + cz, dy:dy + cy, dx:dx + cx] += o
    corner_mass = density * h * h * h / 8
    table = np.zeros(9, np.float32)
    for k in range(1, 9):
        table[k] = np.float32(float(table[k - 1]) + corner_mass)
    mass = table[count.reshape(-1)]
    inv = np.zeros_like(mass)
    nzm = mass > 0
    inv[nzm] = (1.0 / mass[nzm].astype(np.float64)).astype(np.float32)
    # Cells by colour, each colour in z, y, x scan order.
    kc, jc, ic = np.nonzero(occ)  # already in (k, j, i) lexicographic = scan order
    lin = (ic + cx * (jc + cy * kc)).astype(np.uint32)
    col = (ic & 1) | ((jc & 1) << 1) | ((kc & 1) << 2)
    cells = np.concatenate([lin[col == c] for c in range(COLOR_COUNT)]).astype(np.uint32)
    counts = np.array([(col == c).sum() for c in range(COLOR_COUNT)], np.uint32)
    occupied = int(occ.sum())
    return Lattice(tuple(float(v) for v in origin), float(h), (cx, cy, cz), (nx, ny, nz), rest, inv, cells, counts,
                   occupied * density * h ** 3, occupied * h ** 3, occ, float(density))


def surface_nodes(lat):
    """surfaceNodes (lattice.js:128-158): nodes with 1..7 of 8 incident cells occupied, ascending."""
    cx, cy, cz = lat.cell_dims
    occ = np.zeros(cx * cy * cz, np.int32)
    occ[lat.cells] = 1
    occ = occ.reshape(cz, cy, cx)
    nx, ny, nz = lat.node_dims
    filled = np.zeros((nz, ny, nx), np.int32)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                filled[dz:dz + cz, dy:dy + cy, dx:dx + cx] += occ
    flat = filled.reshape(-1)
    return np.nonzero((flat > 0) & (flat < 8))[0].astype(np.uint32)


def block_lattice(cells_xyz, origin, h, density):
    cx, cy, cz = cells_xyz
    return build_lattice(np.ones((cz, cy, cx), bool), origin, h, density)


@dataclass
class Body:
    lattice: Lattice
    material: Material
    static_nodes: np.ndarray = None   # body-local node ids prescribed static
    velocity: tuple = (0.0, 0.0, 0.0)
    splat_mean: np.ndarray = None     # (G, 3) float32
    splat_quat: np.ndarray = None     # (G, 4) float32 xyzw
    splat_scale: np.ndarray = None    # (G, 3) float32

    def gpu_inv_mass(self, prescribed=()):
        w = self.lattice.inv_mass.copy()
        if self.static_nodes is not None and len(self.static_nodes):
            w[self.static_nodes] = 0
        if len(prescribed):
            w[np.asarray(prescribed, np.int64)] = 0
        return w

    @property
    def n_splats(self):
        return 0 if self.splat_mean is None else len(self.splat_mean)

    def constants(self):
        """(alpha_dev, alpha_hyd, gamma) in f64, as elasticCells.js:228 computes them."""
        return element_compliance(self.material, tet_rest_volume(self.lattice.h))


@dataclass
class Drive:
    """A held node set (kinematicDrive.js:155-202) following a sinusoid."""
    nodes: np.ndarray            # scene-global node ids
    offsets: np.ndarray          # (n, 3) float32, node rest minus anchor
    anchor: np.ndarray           # (3,) float64 grab target at t = 0
    direction: np.ndarray        # (3,) unit vector
    amplitude: float = 0.1
    frequency: float = 1.0
    phase: float = 0.0
    law: int = 0                 # 0 hard, 1 compliant
    response_time: float = RESPONSE_TIME

    def target(self, t):
        """Grab target at time t; it starts on the anchor (the grab's own rule)."""
        s = math.sin(2 * math.pi * self.frequency * t + self.phase) - math.sin(self.phase)
        return self.anchor + self.amplitude * s * self.direction

    def frame_target(self, frame, frame_dt):
        """(position, velocity) handed to the drive for frame `frame`: the
        target at the frame's end and its displacement over the frame
        (simulator.js:309-315)."""
        now = self.target((frame + 1) * frame_dt)
        before = self.target(frame * frame_dt)
        return now, (now - before) / frame_dt


@dataclass
class Scene:
    bodies: list
    dt: float
    substeps: int = 10
    gravity: tuple = (0.0, -9.81, 0.0)
    tank_min: tuple = (-10.0, -10.0, -10.0)
    tank_max: tuple = (10.0, 10.0, 10.0)
    contact: int = 0              # 0 off, 1 on (with >= 2 bodies), 2 forced
    bucket_size: float = 0.0      # 0: 2 x the widest contact radius (bodyContact.js:87)
    drive: Drive = None
    name: str = "scene"
    sphere: tuple = None          # sphere collider (cx, cy, cz, radius); None: none
    _surface: list = field(default=None, repr=False)

    # -------- derived layout
    @property
    def n_nodes(self):
        return sum(b.lattice.node_count for b in self.bodies)

    @property
    def n_cells(self):
        return sum(len(b.lattice.cells) for b in self.bodies)

    @property
    def n_splats(self):
        return sum(b.n_splats for b in self.bodies)

    def node_offsets(self):
        offs, at = [], 0
        for b in self.bodies:
            offs.append(at)
            at += b.lattice.node_count
        return offs

    def surfaces(self):
        if self._surface is None:
            self._surface = [surface_nodes(b.lattice) for b in self.bodies]
        return self._surface

    @property
    def contact_active(self):
        return self.contact == 2 or (self.contact == 1 and len(self.bodies) >= 2)

    @property
    def n_surface_active(self):
        return sum(len(s) for s in self.surfaces()) if self.contact else 0

    def grid_bucket_size(self):
        if self.bucket_size > 0:
            return float(np.float32(self.bucket_size))
        # radii live in a Float32Array (voxelLattice.js:72); the edge is twice
        # the widest, formed in f64 (bodyContact.js:74-87).
        widest = max(float(np.float32(b.lattice.h * CONTACT_RADIUS_CELLS)) for b in self.bodies)
        return 2 * widest

    def prescribed(self):
        return set() if self.drive is None or self.drive.law != 0 else set(int(n) for n in self.drive.nodes)

    # -------- flattened arrays (scene-local indices)
    def arrays(self):
        """Every array the oracle and the GPU consume, scene-local."""
        offs = self.node_offsets()
        pres = self.prescribed()
        pos, w = [], []
        for b, off in zip(self.bodies, offs):
            local = [n - off for n in pres if off <= n < off + b.lattice.node_count]
            pos.append(b.lattice.rest)
            w.append(b.gpu_inv_mass(local))
        pos = np.concatenate(pos).astype(np.float32)
        vel = np.concatenate([np.tile(np.asarray(b.velocity, np.float32), (b.lattice.node_count, 1))
                              for b in self.bodies]).astype(np.float32)
        cells, color_start, color_count = [], [], []
        at = 0
        for c in range(COLOR_COUNT):
            color_start.append(at)
            for bi, b in enumerate(self.bodies):
                lat = b.lattice
                s = int(lat.color_counts[:c].sum())
                seg = lat.cells[s:s + int(lat.color_counts[c])]
                cells.append(np.stack([np.full(len(seg), bi, np.uint32), seg], axis=1))
                at += len(seg)
            color_count.append(at - color_start[-1])
        cells = np.concatenate(cells).astype(np.uint32) if cells else np.zeros((0, 2), np.uint32)
        s_node, s_body = [], []
        for bi, (surf, off) in enumerate(zip(self.surfaces(), offs)):
            s_node.append(surf.astype(np.int64) + off)
            s_body.append(np.full(len(surf), bi))
        s_node = np.concatenate(s_node).astype(np.uint32)
        s_body = np.concatenate(s_body).astype(np.uint32)
        s_rad = np.concatenate([np.full(len(s), np.float32(b.lattice.h * CONTACT_RADIUS_CELLS), np.float32)
                                for s, b in zip(self.surfaces(), self.bodies)])
        return dict(pos=pos, vel=vel, w=np.concatenate(w).astype(np.float32), cells=cells,
                    color_start=np.array(color_start), color_count=np.array(color_count),
                    s_node=s_node, s_body=s_body, s_radius=s_rad, node_offsets=np.array(offs))

    def grid(self):
        """broadphaseGrid (broadphase.js:46-58) over the tank, in f64; the
        uniforms then hold origin and edge as f32 (contacts.wgsl:1-8)."""
        edge = self.grid_bucket_size()
        if not edge > 0:
            edge = max(self.tank_max[a] - self.tank_min[a] for a in range(3))
        dims = [max(1, math.ceil((self.tank_max[a] - self.tank_min[a]) / edge)) + 2 for a in range(3)]
        origin = [float(np.float32(self.tank_min[a] - edge)) for a in range(3)]
        return origin, float(np.float32(edge)), dims

    # -------- C ABI descriptor
    def to_desc(self, device_lattice=False):
        """The scene as a vs_scene_desc. With `device_lattice`, every body goes
        as its occupancy grid and the device builds the lattice."""
        keep = []
        offs = self.node_offsets()
        pres = self.prescribed()
        bds = (native.VsBodyDesc * len(self.bodies))()
        for i, (b, off) in enumerate(zip(self.bodies, offs)):
            lat = b.lattice
            local = [n - off for n in pres if off <= n < off + lat.node_count]
            rest = np.ascontiguousarray(lat.rest, np.float32)
            w = np.ascontiguousarray(b.gpu_inv_mass(local), np.float32)
            cells = np.ascontiguousarray(lat.cells, np.uint32)
            surf = np.ascontiguousarray(self.surfaces()[i], np.uint32)
            keep += [rest, w, cells, surf]
            dev, hyd, gamma = b.constants()
            d = bds[i]
            d.cell_dims[:] = lat.cell_dims
            d.n_cells = len(cells)
            d.cells = native.ptr(cells, C.c_uint32)
            d.color_count[:] = [int(v) for v in lat.color_counts]
            d.origin[:] = lat.origin
            d.h = lat.h
            d.rest = native.ptr(rest)
            d.inv_mass = native.ptr(w)
            d.velocity[:] = b.velocity
            d.alpha_dev, d.alpha_hyd, d.gamma = dev, hyd, gamma
            d.relaxation = b.material.relaxation
            d.max_stretch = b.material.max_stretch
            d.friction = b.material.friction
            d.damping = b.material.damping
            d.n_surface = len(surf)
            d.surface = native.ptr(surf, C.c_uint32)
            d.contact_radius = lat.h * CONTACT_RADIUS_CELLS
            if device_lattice:
                occ = np.ascontiguousarray(lat.occupancy, np.uint8).reshape(-1)
                st = np.ascontiguousarray(b.static_nodes if b.static_nodes is not None else [], np.uint32)
                keep += [occ, st]
                d.occupancy = occ.ctypes.data_as(C.POINTER(C.c_uint8))
                d.origin64[:] = lat.origin
                d.h64 = lat.h
                d.density = lat.density
                d.n_static = len(st)
                d.static_nodes = native.ptr(st, C.c_uint32)
            if b.n_splats:
                mean = np.ascontiguousarray(b.splat_mean, np.float32)
                quat = np.ascontiguousarray(b.splat_quat, np.float32)
                scale = np.ascontiguousarray(b.splat_scale, np.float32)
                keep += [mean, quat, scale]
                d.n_splats = len(mean)
                d.splat_mean, d.splat_quat, d.splat_scale = native.ptr(mean), native.ptr(quat), native.ptr(scale)
        desc = native.VsSceneDesc()
        desc.n_bodies = len(self.bodies)
        desc.bodies = C.cast(bds, C.POINTER(native.VsBodyDesc))
        desc.dt = self.dt
        desc.gravity[:] = self.gravity
        desc.tank_min[:] = self.tank_min
        desc.tank_max[:] = self.tank_max
        desc.contact = self.contact
        origin, edge, dims = self.grid()
        desc.bucket_size = edge
        desc.grid_origin[:] = origin
        desc.grid_dims[:] = dims
        if self.drive is not None:
            dn = np.ascontiguousarray(self.drive.nodes, np.uint32)
            do = np.ascontiguousarray(self.drive.offsets, np.float32)
            keep += [dn, do]
            desc.n_driven = len(dn)
            desc.driven = native.ptr(dn, C.c_uint32)
            desc.driven_offset = native.ptr(do)
            desc.drive_law = self.drive.law
            desc.response_time = self.drive.response_time
        if self.sphere is not None:
            desc.sphere[:] = [float(v) for v in self.sphere]
        keep.append(bds)
        return desc, keep


# ------------------------------------------------------------------ splats
def seed_splats(rng, lat, count, mu=-5.0, sigma=0.5):
    """`count` Gaussians uniform in the occupied cells: log-normal axis scales,
    uniformly random unit quaternions (normalised Gaussian 4-vectors, xyzw)."""
    cx, cy, cz = lat.cell_dims
    pick = lat.cells[rng.integers(0, len(lat.cells), count)].astype(np.int64)
    ci = np.stack([pick % cx, (pick // cx) % cy, pick // (cx * cy)], axis=1)
    mean = (np.asarray(lat.origin) + (ci + rng.random((count, 3))) * lat.h).astype(np.float32)
    scale = np.exp(rng.normal(mu, sigma, (count, 3))).astype(np.float32)
    q = rng.normal(size=(count, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return mean, q.astype(np.float32), scale


# --------------------------------------------------------------- workloads
C1_MATERIAL = dict(youngs_modulus=1e5, poisson_ratio=0.3, density=1000.0)


def jelly_block(side, h=0.016, splats=0, seed=0, pinned=True, origin=(0.0, 0.0, 0.0), material=None,
                tank_margin=0.1, substeps=10, fps=60.0, name=None):
    """A `side`^3 block at rest on its pinned bottom node layer (the body of
    the 16^3 single-body workload, at any size)."""
    rng = np.random.default_rng(seed)
    mat = Material(**(material or C1_MATERIAL))
    sides = side if isinstance(side, (tuple, list)) else (side, side, side)
    lat = block_lattice(sides, origin, h, mat.density)
    static = None
    if pinned:
        nx, ny, nz = lat.node_dims
        k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        static = np.nonzero(j3.reshape(-1) == 0)[0].astype(np.uint32)
    body = Body(lat, mat, static)
    if splats:
        body.splat_mean, body.splat_quat, body.splat_scale = seed_splats(rng, lat, splats)
    ext = [sides[a] * h for a in range(3)]
    tmin = tuple(origin[a] - tank_margin for a in range(3))
    tmax = tuple(origin[a] + ext[a] + tank_margin for a in range(3))
    return Scene([body], dt=float(np.float32(1.0 / (fps * substeps))), substeps=substeps,
                 tank_min=tmin, tank_max=tmax, name=name or f"block{side}")


def c1_scene(splats=50_000, seed=1):
    """16^3 cells, h 0.016 m, E 1e5 / nu 0.3 / rho 1000, bottom layer pinned,
    gravity, 1/60 s frames of 10 substeps, 50k Gaussians."""
    return jelly_block(16, splats=splats, seed=seed, name="c1")


def handle_drive(scene, seed, law=0, amplitude=0.1, frequency=1.0):
    """The 3x3x3 node block at the top centre of body 0, driven along a
    seeded sinusoid of the grab target."""
    rng = np.random.default_rng(seed + 7919)
    lat = scene.bodies[0].lattice
    nx, ny, nz = lat.node_dims
    ci, cj, ck = nx // 2, ny - 2, nz // 2
    nodes, offs = [], []
    anchor = lat.rest[ci + nx * (cj + ny * ck)].astype(np.float64)
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                n = (ci + di) + nx * ((cj + dj) + ny * (ck + dk))
                nodes.append(n)
                offs.append(lat.rest[n].astype(np.float64) - anchor)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    return Drive(np.array(nodes, np.uint32), np.array(offs, np.float32), anchor, direction, amplitude, frequency,
                 float(rng.uniform(0, 2 * math.pi)), law)


def c2_scene(side=64, splats=2_000_000, seed=2, sphere=True):
    """64^3 block resting on the tank floor (ground plane y = min), 2M
    Gaussians, a 27-node handle on a seeded 0.1 m / 1 Hz sinusoid, hard law,
    and a sphere collider under the middle of the block: a ball of a quarter
    of the block's width whose cap stands two cells above the floor, into the
    block's underside."""
    sc = jelly_block(side, splats=splats, seed=seed, pinned=False, tank_margin=0.5, name="c2")
    sc.tank_min = (sc.tank_min[0], 0.0, sc.tank_min[2])
    sc.drive = handle_drive(sc, seed)
    if sphere:
        lat = sc.bodies[0].lattice
        width = side * lat.h
        r = 0.25 * width
        sc.sphere = (lat.origin[0] + 0.5 * width, 2 * lat.h - r, lat.origin[2] + 0.5 * width, r)
    return sc


def c3_scene(bodies=512, side=8, h=0.01, box=1.0, seed=3, bucket_voxels=1.0):
    """`bodies` blocks of side^3 cells at h (assumed 0.01 m), seeded
    non-overlapping placement in a `box` m tank, seeded velocities
    U(-0.5, 0.5) m/s, E log-uniform in [1e4, 1e6]."""
    rng = np.random.default_rng(seed)
    ext = side * h
    g = int(math.floor(box / (ext * 1.15)))
    slots = [(i, j, k) for k in range(g) for j in range(g) for i in range(g)]
    pick = rng.permutation(len(slots))[:bodies]
    pitch = box / g
    out = []
    for n in pick:
        i, j, k = slots[n]
        jitter = rng.uniform(0, pitch - ext, 3)
        origin = (i * pitch + jitter[0], j * pitch + jitter[1], k * pitch + jitter[2])
        e = float(10 ** rng.uniform(4, 6))
        mat = Material(youngs_modulus=e, poisson_ratio=0.3, density=1000.0)
        lat = block_lattice((side, side, side), origin, h, mat.density)
        out.append(Body(lat, mat, None, tuple(float(v) for v in rng.uniform(-0.5, 0.5, 3))))
    rate = max(wave_speed(b.material) / b.lattice.h for b in out)
    dt = stable_step_duration(rate)
    return Scene(out, dt=float(np.float32(dt)), substeps=min(MAX_STEPS_PER_FRAME, int(math.floor((1 / 60) / dt + 1e-9))),
                 tank_min=(0.0, 0.0, 0.0), tank_max=(box, box, box), contact=1,
                 bucket_size=float(np.float32(bucket_voxels * h)) if bucket_voxels != 1.0 else 0.0, name="c3")


def contact_stack(count=3, side=4, h=0.01, gap=-0.25, seed=9, friction=0.4):
    """`count` blocks of side^3 cells stacked along y, each overlapping the one
    below by -gap cells (gap < 0) and shifted sideways by a seeded fraction
    of a cell, falling onto the tank floor: contact from the first substep."""
    rng = np.random.default_rng(seed)
    bodies = []
    for k in range(count):
        shift = rng.uniform(-0.5, 0.5, 2) * h
        origin = (shift[0], k * (side + gap) * h + 0.001, shift[1])
        mat = Material(youngs_modulus=5e4, poisson_ratio=0.3, density=1000.0, friction=friction)
        bodies.append(Body(block_lattice((side, side, side), origin, h, mat.density), mat))
    rate = max(wave_speed(b.material) / b.lattice.h for b in bodies)
    ext = count * side * h
    return Scene(bodies, dt=float(np.float32(stable_step_duration(rate))), substeps=8,
                 tank_min=(-0.1, 0.0, -0.1), tank_max=(0.1 + side * h, ext + 0.1, 0.1 + side * h), contact=1,
                 name="stack")


def c4_scenes(count=256, seed=4, min_side=8, max_side=48, min_splats=10_000, max_splats=500_000):
    """`count` independent jelly blocks, side ~ U{8..48}, splats log-uniform."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(min_side, max_side + 1, count)
    splats = np.exp(rng.uniform(math.log(min_splats), math.log(max_splats), count)).astype(np.int64)
    return [jelly_block(int(s), splats=int(g), seed=seed * 100003 + i, name=f"c4-{i}")
            for i, (s, g) in enumerate(zip(sides, splats))]


def c5_requests(count, seed=5, rate=50.0, min_nodes=4_000, max_nodes=500_000, frames=120, splats_per_node=0.0):
    """Poisson arrivals (rate per second) of scenes with log-uniform node
    counts, each with a seeded drag handle; returns (arrival_s, scene, frames)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    nodes = np.exp(rng.uniform(math.log(min_nodes), math.log(max_nodes), count))
    out = []
    for i in range(count):
        side = max(2, int(round(nodes[i] ** (1 / 3))) - 1)
        sc = jelly_block(side, splats=int(splats_per_node * (side + 1) ** 3), seed=seed * 7 + i, pinned=True,
                         name=f"c5-{i}")
        sc.drive = handle_drive(sc, seed * 7 + i)
        out.append((float(t[i]), sc, frames))
    return out
