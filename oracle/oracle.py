"""ctypes binding of the CPU oracle (oracle.c), f32 and f64 builds.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
solver never calls it.

`OracleScene` runs a `vsim.scenes.Scene` through the restatement with the same
scene-local arrays the GPU receives (node order, colour-grouped cells, surface
records body by body).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}


class OBody(C.Structure):
    _fields_ = [
        ("cell_dims", C.c_uint32 * 3), ("node_offset", C.c_uint32),
        ("node_dims", C.c_uint32 * 3), ("node_count", C.c_uint32),
        ("origin", C.c_double * 3),
        ("h", C.c_double), ("alpha_dev", C.c_double), ("alpha_hyd", C.c_double), ("gamma", C.c_double),
        ("relaxation", C.c_double), ("max_stretch", C.c_double), ("friction", C.c_double), ("damping", C.c_double),
    ]


class OScene(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("n_bodies", C.c_int32),
        ("px", C.c_void_p), ("py", C.c_void_p), ("pz", C.c_void_p),
        ("qx", C.c_void_p), ("qy", C.c_void_p), ("qz", C.c_void_p),
        ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
        ("w", C.c_void_p), ("bodies", C.c_void_p), ("cells", C.c_void_p),
        ("color_start", C.c_int32 * 8), ("color_count", C.c_int32 * 8),
        ("dt", C.c_double), ("gravity", C.c_double * 3), ("tank_min", C.c_double * 3), ("tank_max", C.c_double * 3),
        ("n_surf", C.c_int32), ("contact", C.c_int32),
        ("s_node", C.c_void_p), ("s_body", C.c_void_p), ("s_radius", C.c_void_p),
        ("grid_origin", C.c_double * 3), ("bucket_size", C.c_double), ("gain", C.c_double),
        ("grid_dims", C.c_uint32 * 3), ("bucket_slots", C.c_uint32),
        ("n_driven", C.c_int32), ("drive_law", C.c_int32),
        ("d_node", C.c_void_p), ("d_off", C.c_void_p),
        ("target", C.c_double * 3), ("target_vel", C.c_double * 3), ("response_time", C.c_double),
        ("key_out", C.c_void_p), ("pair_count", C.c_void_p), ("pairs", C.c_void_p),
        ("pair_cap", C.c_int32), ("literal", C.c_int32), ("pad1", C.c_int32), ("sphere", C.c_double * 4),
    ]


class OSplats(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("pad0", C.c_int32), ("body", C.c_void_p),
        ("mx", C.c_void_p), ("my", C.c_void_p), ("mz", C.c_void_p), ("quat", C.c_void_p),
        ("sx", C.c_void_p), ("sy", C.c_void_p), ("sz", C.c_void_p),
        ("ox", C.c_void_p), ("oy", C.c_void_p), ("oz", C.c_void_p), ("cov", C.c_void_p),
        ("rot", C.c_void_p), ("axis", C.c_void_p),
    ]


class OShape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_offsets", C.c_uint32), ("size", C.c_double * 3),
                ("offset", (C.c_double * 3) * 4)]


SHAPES = {"sphere": 0, "box": 1, "torus": 2}


class OGrid(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("origin", C.c_double * 3), ("h", C.c_double)]


def load(precision="f32"):
    if precision in _libs:
        return _libs[precision]
    path = os.path.join(HERE, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(path)
    sp = C.POINTER(OScene)
    for name in ("o_predict", "o_drive", "o_contact", "o_finalize", "o_substep"):
        getattr(lib, name).argtypes = [sp]
        getattr(lib, name).restype = None
    lib.o_substeps.argtypes = [sp, C.c_int]
    lib.o_elastic_color.argtypes = [sp, C.c_int, C.c_int]
    lib.o_deform_splats.argtypes = [sp, C.POINTER(OSplats)]
    lib.o_ellipsoid_axes.argtypes = [C.c_void_p] * 3
    lib.o_voxelize_cloud.argtypes = [C.c_void_p] * 3 + [C.c_uint64, C.c_double, C.c_uint32, C.c_double, C.c_uint32,
                                                         C.c_int, C.POINTER(OGrid), C.POINTER(C.c_void_p)]
    lib.o_voxelize_mesh.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_double,
                                    C.c_uint32, C.c_int, C.POINTER(OGrid), C.POINTER(C.c_void_p)]
    lib.o_grid_for_bounds.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_uint32, C.POINTER(OGrid)]
    lib.o_neighbourhood.argtypes = [C.POINTER(OGrid), C.c_void_p, C.c_void_p, C.c_int]
    lib.o_fill_enclosed.argtypes = [C.POINTER(OGrid), C.c_void_p]
    lib.o_fill_enclosed.restype = C.c_uint64
    lib.o_free.argtypes = [C.c_void_p]
    lib.o_voxelize_shape.argtypes = [C.POINTER(OShape), C.c_double, C.c_uint32, C.POINTER(OGrid),
                                     C.POINTER(C.c_void_p)]
    lib.o_pick_ray.argtypes = [C.c_void_p] * 4 + [C.c_uint32, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
    lib.o_pick_ray.restype = C.c_int64
    lib.o_grip_radius.argtypes = [C.c_void_p] * 4 + [C.c_uint32] * 3 + [C.c_double, C.c_uint32] + [C.c_void_p] * 3
    lib.o_grip_radius.restype = C.c_uint32
    lib.o_project_tet.argtypes = [C.c_void_p] * 4 + [C.c_double] * 4 + [C.c_void_p]
    lib.o_project_kuhn_tet.argtypes = [C.c_void_p] * 3 + [C.c_int] + [C.c_double] * 5 + [C.c_int, C.c_void_p]
    lib.o_cell_residuals.argtypes = [sp, C.c_uint32, C.c_uint32, C.c_void_p]
    lib.o_stretch_edges.argtypes = [C.c_void_p] * 3
    lib.o_real_size.restype = C.c_int
    assert lib.o_real_size() == (4 if precision == "f32" else 8)
    _libs[precision] = lib
    return lib


def _planes(pos):
    p = np.ascontiguousarray(np.asarray(pos, np.float32).reshape(3, -1))
    return p, [p[k].ctypes.data for k in range(3)]


def pick_ray(pos, mass, origin, direction, radius):
    """nearestNodeToRay (pick.js:41-68) over positions (3, N) and the grab
    mass (material inverse mass, static nodes zeroed): (node or -1, depth)."""
    p, ptrs = _planes(pos)
    m = np.ascontiguousarray(mass, np.float32)
    o = np.ascontiguousarray(origin, np.float64)
    d = np.ascontiguousarray(direction, np.float64)
    depth = C.c_double()
    node = load("f64").o_pick_ray(*ptrs, m.ctypes.data, len(m), o.ctypes.data, d.ctypes.data, float(radius),
                                  C.byref(depth))
    return int(node), depth.value


def grip_radius(pos, mass, anchor, begin, end, radius, limit):
    """withinRadius (grip.js:99-138): (nodes, offsets (k, 3) f64, found)."""
    p, ptrs = _planes(pos)
    m = np.ascontiguousarray(mass, np.float32)
    nodes = np.zeros(max(limit, 1), np.uint32)
    offs = np.zeros((max(limit, 1), 3), np.float64)
    found = C.c_uint32()
    k = load("f64").o_grip_radius(*ptrs, m.ctypes.data, begin, end, anchor, float(radius), limit,
                                  nodes.ctypes.data, offs.ctypes.data, C.byref(found))
    return nodes[:k].copy(), offs[:k].copy(), found.value


def stretch_edges(precision="f64"):
    """The tensile limit's corner pairs and unit-cell lengths as the
    oracle's cell sweep reads them (STRETCH_EDGES, stretch.js:5-7)."""
    a = np.zeros(28, np.int32)
    b = np.zeros(28, np.int32)
    length = np.zeros(28, np.float64)
    load(precision).o_stretch_edges(a.ctypes.data, b.ctypes.data, length.ctypes.data)
    return a, b, length


class Grid:
    """An occupancy grid (OccupancyGrid, voxelize.js:1-16): dims (x, y, z),
    f64 origin and edge, and one byte per cell with x fastest."""

    def __init__(self, dims, origin, h, occupancy):
        self.dims, self.origin, self.h = tuple(int(d) for d in dims), tuple(float(o) for o in origin), float(h)
        self.occupancy = occupancy

    def volume(self):
        return int(self.occupancy.sum()) * self.h ** 3

    def at(self, point):
        """Occupancy of the cell holding `point` (locate, sample.js:42-60)."""
        c = [min(max(int(np.floor((point[a] - self.origin[a]) / self.h)), 0), self.dims[a] - 1) for a in range(3)]
        return int(self.occupancy[c[0] + self.dims[0] * (c[1] + self.dims[1] * c[2])])

    def array(self):
        """(z, y, x) view."""
        return self.occupancy.reshape(self.dims[2], self.dims[1], self.dims[0])


def _grid_from(g, ptr):
    n = g.dims[0] * g.dims[1] * g.dims[2]
    occ = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), shape=(n,)).copy()
    load("f64").o_free(ptr)
    return Grid(list(g.dims), list(g.origin), g.h, occ)


VOXEL_ERRORS = {
    -1: "voxelizeCloud: a cloud with no Gaussians has no volume",
    -2: "voxelizeCloud: a cloud with no extent has no volume",
    -3: "gridForBounds: elementSize must be a positive length in metres",
    -4: "gridForBounds: the bounds need more cells than one grid holds",
    -5: "voxelizeMesh: vertex coordinates must be finite and indices inside the vertices",
}


def _raise(rc):
    if rc:
        raise ValueError(VOXEL_ERRORS[rc])


def voxelize_cloud(positions, rotations, scales, element_size, padding=2, sigma=2.0, seal=2, fill=True):
    """voxelizeCloud (cloudVolume.js:95-121) over xyz / xyzw / xyz arrays."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1)
    r = np.ascontiguousarray(rotations, np.float32).reshape(-1)
    s = np.ascontiguousarray(scales, np.float32).reshape(-1)
    g = OGrid()
    ptr = C.c_void_p()
    rc = load("f64").o_voxelize_cloud(p.ctypes.data, r.ctypes.data, s.ctypes.data, len(p) // 3, element_size,
                                      padding, sigma, seal, int(fill), C.byref(g), C.byref(ptr))
    _raise(rc)
    return _grid_from(g, ptr)


def voxelize_mesh(positions, indices, element_size, padding=2, close=True, bounds=None):
    """voxelizeMesh (voxelize.js:174-203) over `bounds` ((min, max)) or
    meshBounds."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1)
    ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    b = None if bounds is None else np.ascontiguousarray(np.concatenate(bounds), np.float64)
    g = OGrid()
    ptr = C.c_void_p()
    rc = load("f64").o_voxelize_mesh(p.ctypes.data, len(p) // 3, ix.ctypes.data, len(ix) // 3,
                                     None if b is None else b.ctypes.data, element_size, padding, int(close),
                                     C.byref(g), C.byref(ptr))
    _raise(rc)
    return _grid_from(g, ptr)


def voxelize_shape(kind, size, element_size, offset=(0.0, 0.0, 0.0), padding=2):
    """voxelizeSdf (voxelize.js:118-148) of sphere (radius) / box (half
    extents) / torus (major, minor). `offset` is one translation or a list of
    up to four, applied in order (translateShape of translateShape ...)."""
    sh = OShape(kind=SHAPES[kind])
    sh.size[:] = list(size) + [0.0] * (3 - len(size))
    offsets = [offset] if np.ndim(offset) == 1 else list(offset)
    sh.n_offsets = len(offsets)
    for t, o in enumerate(offsets):
        sh.offset[t][:] = [float(v) for v in o]
    g = OGrid()
    ptr = C.c_void_p()
    _raise(load("f64").o_voxelize_shape(C.byref(sh), element_size, padding, C.byref(g), C.byref(ptr)))
    return _grid_from(g, ptr)


def grid_for_bounds(mn, mx, element_size, padding=2):
    g = OGrid()
    lo = np.ascontiguousarray(mn, np.float64)
    hi = np.ascontiguousarray(mx, np.float64)
    _raise(load("f64").o_grid_for_bounds(lo.ctypes.data, hi.ctypes.data, element_size, padding, C.byref(g)))
    return Grid(list(g.dims), list(g.origin), g.h, np.zeros(g.dims[0] * g.dims[1] * g.dims[2], np.uint8))


def close_grid(grid):
    """morphologicalClose (voxelize.js:386-394), in place."""
    g = OGrid()
    g.dims[:] = grid.dims
    g.origin[:] = grid.origin
    g.h = grid.h
    tmp = np.empty_like(grid.occupancy)
    lib = load("f64")
    lib.o_neighbourhood(C.byref(g), grid.occupancy.ctypes.data, tmp.ctypes.data, 1)
    lib.o_neighbourhood(C.byref(g), tmp.ctypes.data, grid.occupancy.ctypes.data, 0)


def ellipsoid_axes(cov6, precision="f64"):
    """Rotation (xyzw) and axis lengths of a covariance given as (xx, xy, xz,
    yy, yz, zz): the oracle's four-sweep cyclic Jacobi
    (apps/simulator/src/render/ellipsoid.js:40-149)."""
    lib = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    c = np.ascontiguousarray(cov6, dt)
    rot = np.zeros(4, dt)
    scale = np.zeros(3, dt)
    lib.o_ellipsoid_axes(c.ctypes.data, rot.ctypes.data, scale.ctypes.data)
    return rot, scale


def project_tet(p, prev, w, B, alpha_dev, alpha_hyd, gamma, relax=0.0, precision="f64"):
    """One tetrahedron of the 2x2 block projection (project.wgsl:75-153);
    p, prev (4, 3); B column-major 3x3 as 9 values; alphas over dt^2."""
    dt = np.float32 if precision == "f32" else np.float64
    lib = load(precision)
    p = np.ascontiguousarray(p, dt).reshape(12)
    prev = np.ascontiguousarray(prev, dt).reshape(12)
    w = np.ascontiguousarray(w, dt)
    B = np.ascontiguousarray(B, dt).reshape(9)
    out = np.empty(12, dt)
    lib.o_project_tet(p.ctypes.data, prev.ctypes.data, w.ctypes.data, B.ctypes.data, alpha_dev, alpha_hyd, gamma,
                      relax, out.ctypes.data)
    return out.reshape(4, 3)


def project_kuhn_tet(p, prev, w, t, h, alpha_dev, alpha_hyd, gamma, relax=0.0, literal=False, precision="f64"):
    """Kuhn tetrahedron t (corners in KUHN_TETS order) in the literal WGSL
    formulas or in the fused order the GPU kernel evaluates."""
    dt = np.float32 if precision == "f32" else np.float64
    lib = load(precision)
    p = np.ascontiguousarray(p, dt).reshape(12)
    prev = np.ascontiguousarray(prev, dt).reshape(12)
    w = np.ascontiguousarray(w, dt)
    out = np.empty(12, dt)
    lib.o_project_kuhn_tet(p.ctypes.data, prev.ctypes.data, w.ctypes.data, t, h, alpha_dev, alpha_hyd, gamma, relax,
                           int(literal), out.ctypes.data)
    return out.reshape(4, 3)


class OracleScene:
    """A Scene's state and tables in the oracle's layout, in f32 or f64."""

    def __init__(self, scene, precision="f32", pair_cap=0, literal=False):
        self.scene = scene
        self.precision = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.lib = load(precision)
        a = scene.arrays()
        self.a = a
        dt = self.dtype
        n = len(a["w"])
        self.p = np.ascontiguousarray(a["pos"].T, dt)         # (3, N)
        self.q = self.p.copy()
        self.v = np.ascontiguousarray(a["vel"].T, dt)
        self.w = np.ascontiguousarray(a["w"], dt)
        self.cells = np.ascontiguousarray(a["cells"], np.uint32)
        bodies = (OBody * len(scene.bodies))()
        for i, (b, off) in enumerate(zip(scene.bodies, a["node_offsets"])):
            lat = b.lattice
            dev, hyd, gamma = b.constants()
            ob = bodies[i]
            ob.cell_dims[:] = lat.cell_dims
            ob.node_offset = int(off)
            ob.node_dims[:] = lat.node_dims
            ob.node_count = lat.node_count
            ob.origin[:] = [float(np.float32(v)) for v in lat.origin]
            ob.h, ob.alpha_dev, ob.alpha_hyd, ob.gamma = lat.h, dev, hyd, gamma
            ob.relaxation, ob.max_stretch = b.material.relaxation, b.material.max_stretch
            ob.friction, ob.damping = b.material.friction, b.material.damping
        self.bodies = bodies
        self.s_node = a["s_node"]
        self.s_body = a["s_body"]
        self.s_radius = np.ascontiguousarray(a["s_radius"], dt)
        S = len(self.s_node)
        self.keys = np.zeros(S, np.uint32)
        self.pair_cap = pair_cap
        self.pair_count = np.zeros(S, np.int32)
        self.pairs = np.zeros((S, max(pair_cap, 1)), np.uint32)
        sc = OScene()
        sc.n_nodes, sc.n_bodies = n, len(scene.bodies)
        sc.px, sc.py, sc.pz = (self.p[k].ctypes.data for k in range(3))
        sc.qx, sc.qy, sc.qz = (self.q[k].ctypes.data for k in range(3))
        sc.vx, sc.vy, sc.vz = (self.v[k].ctypes.data for k in range(3))
        sc.w = self.w.ctypes.data
        sc.bodies = C.cast(bodies, C.c_void_p)
        sc.cells = self.cells.ctypes.data
        sc.color_start[:] = [int(v) for v in a["color_start"]]
        sc.color_count[:] = [int(v) for v in a["color_count"]]
        sc.dt = scene.dt
        sc.gravity[:] = scene.gravity
        sc.tank_min[:] = scene.tank_min
        sc.tank_max[:] = scene.tank_max
        sc.n_surf = S if scene.contact else 0
        sc.contact = 1 if scene.contact_active else 0
        sc.s_node, sc.s_body, sc.s_radius = self.s_node.ctypes.data, self.s_body.ctypes.data, self.s_radius.ctypes.data
        origin, edge, dims = scene.grid()
        sc.grid_origin[:] = origin
        sc.bucket_size = edge
        sc.gain = 1.0
        sc.grid_dims[:] = dims
        sc.bucket_slots = 32
        sc.key_out = self.keys.ctypes.data
        sc.literal = int(literal)
        if scene.sphere is not None:
            sc.sphere[:] = [float(np.float32(v)) for v in scene.sphere]
        if pair_cap:
            sc.pair_count = self.pair_count.ctypes.data
            sc.pairs = self.pairs.ctypes.data
            sc.pair_cap = pair_cap
        drv = scene.drive
        if drv is not None:
            self.d_node = np.ascontiguousarray(drv.nodes, np.uint32)
            self.d_off = np.ascontiguousarray(drv.offsets, dt).reshape(-1)
            sc.n_driven = len(self.d_node)
            sc.drive_law = drv.law
            sc.d_node, sc.d_off = self.d_node.ctypes.data, self.d_off.ctypes.data
            sc.response_time = drv.response_time
        self.sc = sc
        self.frame_index = 0
        # splats
        means, quats, scales, owner = [], [], [], []
        for i, b in enumerate(scene.bodies):
            if b.n_splats:
                means.append(b.splat_mean)
                quats.append(b.splat_quat)
                scales.append(b.splat_scale)
                owner.append(np.full(b.n_splats, i, np.uint32))
        G = scene.n_splats
        self.G = G
        if G:
            m = np.concatenate(means).T
            s = np.concatenate(scales).T
            self.sm = np.ascontiguousarray(m, dt)
            self.ss = np.ascontiguousarray(s, dt)
            self.sq = np.ascontiguousarray(np.concatenate(quats), dt).reshape(-1)
            self.sb = np.concatenate(owner)
            self.out_mean = np.zeros((3, G), dt)
            self.out_cov = np.zeros((6, G), dt)
            sp = OSplats()
            sp.n = G
            sp.body = self.sb.ctypes.data
            sp.mx, sp.my, sp.mz = (self.sm[k].ctypes.data for k in range(3))
            sp.quat = self.sq.ctypes.data
            sp.sx, sp.sy, sp.sz = (self.ss[k].ctypes.data for k in range(3))
            sp.ox, sp.oy, sp.oz = (self.out_mean[k].ctypes.data for k in range(3))
            sp.cov = self.out_cov.ctypes.data
            self.out_rot = np.zeros((4, G), dt)
            self.out_axis = np.zeros((3, G), dt)
            self.sp = sp

    # passes
    def set_target(self, target, velocity):
        self.sc.target[:] = [float(v) for v in target]
        self.sc.target_vel[:] = [float(v) for v in velocity]

    def predict(self):
        self.lib.o_predict(C.byref(self.sc))

    def drive(self):
        self.lib.o_drive(C.byref(self.sc))

    def elastic(self, color, limit_only=0):
        self.lib.o_elastic_color(C.byref(self.sc), color, limit_only)

    def contact(self):
        self.lib.o_contact(C.byref(self.sc))

    def finalize(self):
        self.lib.o_finalize(C.byref(self.sc))

    def substep(self):
        self.lib.o_substep(C.byref(self.sc))

    def substeps(self, count):
        self.lib.o_substeps(C.byref(self.sc), count)

    def frame(self, substeps=None, splats=False):
        """One frame as Simulator.dispatch runs it (simulator.js:292-318)."""
        n = substeps or self.scene.substeps
        drv = self.scene.drive
        if drv is not None:
            t, v = drv.frame_target(self.frame_index, n * self.scene.dt)
            self.set_target(t, v)
        self.substeps(n)
        self.frame_index += 1
        if splats:
            self.deform_splats()

    def deform_splats(self, axes=False):
        """Deformed means and covariances (6 planes); with `axes`, also the
        engine format, returned as self.out_rot (xyzw planes) and
        self.out_axis (axis lengths)."""
        if self.G:
            self.sp.rot = self.out_rot.ctypes.data if axes else None
            self.sp.axis = self.out_axis.ctypes.data if axes else None
            self.lib.o_deform_splats(C.byref(self.sc), C.byref(self.sp))
        return getattr(self, "out_mean", None), getattr(self, "out_cov", None)

    def residuals(self, body, linear):
        out = np.empty(12, self.dtype)
        self.lib.o_cell_residuals(C.byref(self.sc), body, linear, out.ctypes.data)
        return out.reshape(6, 2)
