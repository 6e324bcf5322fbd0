"""Capture ingest: PLY clouds and occupancy grids, through libvsim.so.

Mirrors the reference's host-side entry points with the same argument
meaning, defaults and error messages (raised as ValueError, as the JS throws
Error):

* ``read_ply`` / ``cloud_from_ply``: readPly (apps/simulator/src/assets/
  splats/ply.js:65-104) and cloudFromPly (apps/simulator/src/assets/splats/
  cloud.js:52-81), parsed natively on the host;
* ``voxelize_cloud`` / ``voxelize_mesh`` / ``voxelize_shape``: voxelizeCloud
  (packages/visionary-sim/src/sim/discretize/lattice/cloudVolume.js:95-121),
  voxelizeMesh (voxelize.js:174-203) and voxelizeSdf (voxelize.js:118-148),
  computed on the GPU;
* ``Occupancy.body``: the grid as a lattice body, ready for the device
  lattice build (``World.add(scene, device_lattice=True)``).
"""
import ctypes as C

import numpy as np

from . import native, scenes

SHAPES = {"sphere": native.SHAPE_SPHERE, "box": native.SHAPE_BOX, "torus": native.SHAPE_TORUS}

# REQUIRED_PROPERTIES (cloud.js:28-34).
REQUIRED_PROPERTIES = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
                       "rot_0", "rot_1", "rot_2", "rot_3"]


def _fail(lib):
    raise ValueError(lib.vs_ingest_error().decode())


class Occupancy:
    """OccupancyGrid (voxelize.js:21-29): dims (x, y, z), f64 origin and cell
    edge, one byte per cell with x fastest (1 = material)."""

    def __init__(self, dims, origin, h, cells):
        self.dims = tuple(int(d) for d in dims)
        self.origin = tuple(float(o) for o in origin)
        self.h = float(h)
        self.cells = cells

    def array(self):
        """(z, y, x) view."""
        return self.cells.reshape(self.dims[2], self.dims[1], self.dims[0])

    def count(self):
        """countOccupied (voxelize.js:84-90)."""
        return int(np.count_nonzero(self.cells))

    def volume(self):
        return self.count() * self.h ** 3

    def body(self, material, static_nodes=None, velocity=(0.0, 0.0, 0.0), splats=None):
        """A scenes.Body over this grid: the lattice is described by its
        occupancy (buildLattice, lattice.js:166-239), so it can be built on
        the device. `splats` = (mean, quat xyzw, scale) arrays or None."""
        lat = scenes.build_lattice(self.array(), self.origin, self.h, material.density)
        body = scenes.Body(lat, material, static_nodes, tuple(velocity))
        if splats is not None:
            body.splat_mean, body.splat_quat, body.splat_scale = (np.ascontiguousarray(a, np.float32) for a in splats)
        return body


def _take(lib, occ):
    n = occ.dims[0] * occ.dims[1] * occ.dims[2]
    try:
        cells = np.ctypeslib.as_array(occ.cells, shape=(n,)).copy() if n else np.zeros(0, np.uint8)
    finally:
        lib.vs_occupancy_free(C.byref(occ))
    return Occupancy(list(occ.dims), list(occ.origin), occ.h, cells)


def voxelize_cloud(positions, rotations, scales, element_size, padding=2, sigma=2.0, seal=2, fill=True, device=0):
    """voxelizeCloud over xyz positions, xyzw unit rotations and scales (m)."""
    lib = native.load()
    p = np.ascontiguousarray(positions, np.float32).reshape(-1)
    r = np.ascontiguousarray(rotations, np.float32).reshape(-1)
    s = np.ascontiguousarray(scales, np.float32).reshape(-1)
    n = len(p) // 3
    if len(r) != 4 * n or len(s) != 3 * n:
        raise ValueError("voxelizeCloud: positions, rotations and scales disagree on the Gaussian count")
    occ = native.VsOccupancy()
    rc = lib.vs_voxelize_cloud(device, p.ctypes.data, r.ctypes.data, s.ctypes.data, n, float(element_size),
                               int(padding), float(sigma), int(seal), int(bool(fill)), C.byref(occ))
    if rc != native.VS_OK:
        _fail(lib)
    return _take(lib, occ)


def voxelize_mesh(positions, indices, element_size, padding=2, close=True, bounds=None, device=0):
    """voxelizeMesh over a triangle mesh; `bounds` = (min, max), default the
    vertices' bounds (meshBounds, restGeometry.js:29-40)."""
    lib = native.load()
    p = np.ascontiguousarray(positions, np.float32).reshape(-1)
    ix = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    b = None if bounds is None else np.ascontiguousarray(np.concatenate([bounds[0], bounds[1]]), np.float64)
    occ = native.VsOccupancy()
    rc = lib.vs_voxelize_mesh(device, p.ctypes.data, len(p) // 3, ix.ctypes.data, len(ix) // 3,
                              None if b is None else b.ctypes.data, float(element_size), int(padding),
                              int(bool(close)), C.byref(occ))
    if rc != native.VS_OK:
        _fail(lib)
    return _take(lib, occ)


def voxelize_shape(kind, size, element_size, offset=(0.0, 0.0, 0.0), padding=2, device=0):
    """voxelizeSdf of an analytic shape: "sphere" (radius), "box" (half
    extents) or "torus" (major, minor radius), translated by `offset`: one
    translation, or a list of up to four applied in order."""
    lib = native.load()
    sh = native.VsShape(kind=SHAPES[kind])
    sh.size[:] = [float(v) for v in size] + [0.0] * (3 - len(size))
    offsets = [offset] if np.ndim(offset) == 1 else list(offset)
    if len(offsets) > 4:
        raise ValueError("voxelizeSdf: at most four translations")
    sh.n_offsets = len(offsets)
    for t, o in enumerate(offsets):
        sh.offset[t][:] = [float(v) for v in o]
    occ = native.VsOccupancy()
    if lib.vs_voxelize_shape(device, C.byref(sh), float(element_size), int(padding), C.byref(occ)) != native.VS_OK:
        _fail(lib)
    return _take(lib, occ)


def ply_properties(data):
    """(count, property names of the first element)."""
    lib = native.load()
    count = C.c_uint64()
    need = len(data) + 1
    names = C.create_string_buffer(need)
    if lib.vs_ply_properties(data, len(data), names, need, C.byref(count)) != native.VS_OK:
        _fail(lib)
    text = names.value.decode()
    return count.value, text.split(" ") if text else []


def read_ply(data, properties=None):
    """readPly over the bytes of a file: (count, {name: float32 array}),
    every property when `properties` is None."""
    lib = native.load()
    if properties is None:
        properties = ply_properties(data)[1]
    names = (C.c_char_p * max(len(properties), 1))(*[n.encode() for n in properties])
    cols = (C.POINTER(C.c_float) * max(len(properties), 1))()
    count = C.c_uint64()
    if lib.vs_ply_read(data, len(data), names, len(properties), C.byref(count), cols) != native.VS_OK:
        _fail(lib)
    out = {}
    try:
        for i, name in enumerate(properties):
            out[name] = np.ctypeslib.as_array(cols[i], shape=(count.value,)).copy() if count.value else \
                np.zeros(0, np.float32)
    finally:
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        for i in range(len(properties)):
            libc.free(C.cast(cols[i], C.c_void_p))
    return count.value, out


class Cloud:
    """SplatCloud (cloud.js:20-22): positions (n, 3), xyzw unit rotations
    (n, 4) with w >= 0, scales (n, 3) in metres, opacity (n,), colours (n, 3)."""

    def __init__(self, positions, rotations, scales, opacity, colors):
        self.positions, self.rotations, self.scales = positions, rotations, scales
        self.opacity, self.colors = opacity, colors

    @property
    def count(self):
        return len(self.positions)


def cloud_from_ply(data):
    """loadCloud's decode step (cloud.js:45): readPly with the required
    properties, then cloudFromPly."""
    lib = native.load()
    c = native.VsCloud()
    if lib.vs_cloud_from_ply(data, len(data), C.byref(c)) != native.VS_OK:
        _fail(lib)
    return _take_cloud(lib, c)


def cloud_from_columns(count, properties):
    """cloudFromPly (cloud.js:52-81) over an element already read:
    `properties` maps each required name to a float32 array of `count`."""
    lib = native.load()
    cols = [np.ascontiguousarray(properties[n], np.float32) for n in REQUIRED_PROPERTIES]
    if any(len(a) != count for a in cols):
        raise ValueError("cloudFromPly: every property needs `count` values")
    ptrs = (C.POINTER(C.c_float) * 14)(*[native.ptr(a) for a in cols])
    c = native.VsCloud()
    if lib.vs_cloud_from_columns(count, ptrs, C.byref(c)) != native.VS_OK:
        _fail(lib)
    return _take_cloud(lib, c)


def _take_cloud(lib, c):
    n = c.count
    try:
        arr = lambda ptr, w: np.ctypeslib.as_array(ptr, shape=(n * w,)).reshape(n, w).copy() if n else \
            np.zeros((0, w), np.float32)
        return Cloud(arr(c.positions, 3), arr(c.rotations, 4), arr(c.scales, 3), arr(c.opacity, 1).reshape(-1),
                     arr(c.colors, 3))
    finally:
        lib.vs_cloud_free(C.byref(c))
