"""ctypes binding of ``include/vsim.h`` (libvsim.so).

This module loads the in-tree library and fails loudly when it is missing:
there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(os.path.dirname(HERE))
LIB_PATH = os.path.join(PKG, "lib", "libvsim.so")

VS_OK, VS_E_INVALID, VS_E_DEVICE, VS_E_NOMEM, VS_E_DESTROYED = 0, -1, -2, -3, -4

# Kernel ids and buffer slots (vsim.h).
K_PREDICT, K_FINALIZE, K_ELASTIC, K_BIN, K_SORT, K_COLLIDE, K_SETTLE, K_DRIVE_HARD, K_DRIVE_COMPLIANT, K_SPLAT = range(10)
B_POS, B_PREV, B_VEL, B_INVMASS, B_BODIES, B_SCENES, B_TILES, B_ITEMS, B_GRID, B_OUT = range(10)


class VsBody(C.Structure):
    _fields_ = [
        ("cell_dims", C.c_uint32 * 3), ("node_offset", C.c_uint32),
        ("node_dims", C.c_uint32 * 3), ("scene", C.c_uint32),
        ("h", C.c_float), ("alpha_dev", C.c_float), ("alpha_hyd", C.c_float), ("gamma", C.c_float),
        ("relaxation", C.c_float), ("max_stretch", C.c_float), ("friction", C.c_float), ("damp", C.c_float),
        ("origin", C.c_float * 3), ("node_count", C.c_uint32),
    ]


class VsSceneParams(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gravity", C.c_float * 3),
        ("tank_min", C.c_float * 3), ("gain", C.c_float),
        ("tank_max", C.c_float * 3), ("bucket_size", C.c_float),
        ("grid_origin", C.c_float * 3), ("bucket_count", C.c_uint32),
        ("grid_dims", C.c_uint32 * 3), ("grid_base", C.c_uint32),
        ("target", C.c_float * 3), ("response_time", C.c_float),
        ("target_vel", C.c_float * 3), ("drive_law", C.c_uint32),
        ("surf_base", C.c_uint32), ("surf_count", C.c_uint32), ("bucket_slots", C.c_uint32), ("flags", C.c_uint32),
        ("sphere", C.c_float * 4), ("reach_max", C.c_float), ("skip_margin", C.c_float), ("pad1", C.c_uint32 * 2),
    ]


class VsLaunch(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32), ("grid", C.c_uint32), ("block", C.c_uint32), ("plane", C.c_uint32),
        ("items", C.c_uint32), ("flags", C.c_uint32), ("buckets", C.c_uint32), ("pair_cap", C.c_uint32),
        ("buffers", C.c_uint64 * 10),
    ]


class VsWorldConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("max_scenes", C.c_uint32), ("max_bodies", C.c_uint32),
        ("max_nodes", C.c_uint64), ("max_cells", C.c_uint64), ("max_surface", C.c_uint64),
        ("max_buckets", C.c_uint64), ("max_driven", C.c_uint64), ("max_splats", C.c_uint64),
        ("tile_nodes", C.c_uint32), ("tile_cells", C.c_uint32), ("block_cells", C.c_uint32),
        ("persistent", C.c_uint32), ("xcd_chunked", C.c_uint32), ("separate_node_passes", C.c_uint32),
        ("separate_colours", C.c_uint32), ("colour_launches", C.c_uint32), ("branches", C.c_uint32),
    ]


class VsBodyDesc(C.Structure):
    _fields_ = [
        ("cell_dims", C.c_uint32 * 3), ("n_cells", C.c_uint32),
        ("cells", C.POINTER(C.c_uint32)), ("color_count", C.c_uint32 * 8),
        ("origin", C.c_float * 3), ("h", C.c_float),
        ("rest", C.POINTER(C.c_float)), ("inv_mass", C.POINTER(C.c_float)),
        ("velocity", C.c_float * 3),
        ("alpha_dev", C.c_float), ("alpha_hyd", C.c_float), ("gamma", C.c_float), ("relaxation", C.c_float),
        ("max_stretch", C.c_float), ("friction", C.c_float), ("damping", C.c_float),
        ("n_surface", C.c_uint32), ("surface", C.POINTER(C.c_uint32)), ("contact_radius", C.c_float),
        ("n_splats", C.c_uint32), ("splat_mean", C.POINTER(C.c_float)),
        ("splat_quat", C.POINTER(C.c_float)), ("splat_scale", C.POINTER(C.c_float)),
        ("occupancy", C.POINTER(C.c_uint8)), ("origin64", C.c_double * 3), ("h64", C.c_double),
        ("density", C.c_double), ("n_static", C.c_uint32), ("static_nodes", C.POINTER(C.c_uint32)),
    ]


class VsBodyInfo(C.Structure):
    _fields_ = [("node_count", C.c_uint32), ("n_cells", C.c_uint32), ("n_surface", C.c_uint32),
                ("pad0", C.c_uint32), ("color_count", C.c_uint32 * 8)]


class VsSceneDesc(C.Structure):
    _fields_ = [
        ("n_bodies", C.c_uint32), ("bodies", C.POINTER(VsBodyDesc)),
        ("dt", C.c_float), ("gravity", C.c_float * 3), ("tank_min", C.c_float * 3), ("tank_max", C.c_float * 3),
        ("contact", C.c_int32), ("bucket_size", C.c_float),
        ("grid_origin", C.c_float * 3), ("grid_dims", C.c_uint32 * 3),
        ("n_driven", C.c_uint32), ("driven", C.POINTER(C.c_uint32)), ("driven_offset", C.POINTER(C.c_float)),
        ("drive_law", C.c_uint32), ("response_time", C.c_float), ("sphere", C.c_float * 4),
    ]


class VsWorldStats(C.Structure):
    _fields_ = [
        ("scenes", C.c_uint64), ("bodies", C.c_uint64), ("nodes", C.c_uint64), ("cells", C.c_uint64),
        ("surface", C.c_uint64), ("buckets", C.c_uint64), ("splats", C.c_uint64), ("driven", C.c_uint64),
        ("bytes_reserved", C.c_uint64), ("graph_captures", C.c_uint32), ("pad0", C.c_uint32),
    ]


class VsServerConfig(C.Structure):
    _fields_ = [("substeps", C.c_uint32), ("splats", C.c_int32), ("graph", C.c_int32), ("keep_results", C.c_int32)]


class VsServerReport(C.Structure):
    _fields_ = [
        ("wall_s", C.c_double), ("node_substeps", C.c_uint64), ("frames", C.c_uint64),
        ("requests", C.c_uint32), ("completed", C.c_uint32), ("iterations", C.c_uint32), ("max_active", C.c_uint32),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
        ("mean_queue_s", C.c_double), ("max_queue_s", C.c_double),
    ]


EXPORTS = [
    "vs_backend_create", "vs_backend_destroy", "vs_last_error", "vs_buffer_alloc", "vs_buffer_free",
    "vs_buffer_write", "vs_buffer_read", "vs_kernel_id", "vs_dispatch", "vs_synchronize", "vs_grip_noalloc",
    "vs_world_create", "vs_world_destroy", "vs_world_error", "vs_world_add_scene", "vs_world_remove_scene",
    "vs_world_reset_scene", "vs_world_set_target", "vs_world_frame", "vs_world_frame_scene",
    "vs_world_synchronize", "vs_world_record", "vs_world_elapsed", "vs_world_read_nodes",
    "vs_world_read_splats", "vs_world_read_splat_axes", "vs_world_body_info", "vs_world_read_lattice", "vs_world_log_pairs", "vs_world_read_contacts", "vs_world_stats_get",
    "vs_world_autotune", "vs_world_profile", "vs_world_time_pass",
    "vs_server_create", "vs_server_destroy", "vs_server_error", "vs_server_submit", "vs_server_run",
    "vs_server_latencies", "vs_server_result", "vs_server_times",
    "vs_voxelize_cloud", "vs_voxelize_mesh", "vs_voxelize_shape", "vs_occupancy_free", "vs_ingest_error",
    "vs_ply_read", "vs_ply_properties", "vs_cloud_from_columns", "vs_cloud_from_ply", "vs_cloud_free",
]

SPLAT_COV, SPLAT_AXES = 1, 2  # VS_SPLAT_* output bits of a frame's `splats`

PASS_KINDS = ["predict", "drive", "elastic", "bin", "sort", "collide", "settle", "finalize", "splat",
              "finalize+predict", "pair", "pair+nodes"]

class VsOccupancy(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("pad0", C.c_uint32), ("origin", C.c_double * 3), ("h", C.c_double),
                ("cells", C.POINTER(C.c_uint8))]


SHAPE_SPHERE, SHAPE_BOX, SHAPE_TORUS = 0, 1, 2


class VsShape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_offsets", C.c_uint32), ("size", C.c_double * 3),
                ("offset", (C.c_double * 3) * 4)]


class VsCloud(C.Structure):
    _fields_ = [("count", C.c_uint64), ("positions", C.POINTER(C.c_float)), ("rotations", C.POINTER(C.c_float)),
                ("scales", C.POINTER(C.c_float)), ("opacity", C.POINTER(C.c_float)), ("colors", C.POINTER(C.c_float))]


_lib = None


def load(path=LIB_PATH):
    """Loads libvsim.so (built by visionary-sim_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python visionary-sim_amd/build.py` first")
    lib = C.CDLL(path)
    p, u32, u64, i32, f = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_float
    sig = {
        "vs_backend_create": (C.c_int, [C.c_int, C.POINTER(p)]),
        "vs_backend_destroy": (None, [p]),
        "vs_last_error": (C.c_char_p, [p]),
        "vs_buffer_alloc": (C.c_int, [p, u64, C.POINTER(u64)]),
        "vs_buffer_free": (C.c_int, [p, u64]),
        "vs_buffer_write": (C.c_int, [p, u64, u64, p, u64]),
        "vs_buffer_read": (C.c_int, [p, u64, u64, u64, p]),
        "vs_kernel_id": (C.c_int, [C.c_char_p]),
        "vs_dispatch": (C.c_int, [p, C.POINTER(VsLaunch), u32]),
        "vs_synchronize": (C.c_int, [p]),
        "vs_grip_noalloc": (C.c_int, [p, u64, u32, u64, C.POINTER(VsGrabParams), u32, u32, u64]),
        "vs_world_create": (C.c_int, [C.POINTER(VsWorldConfig), C.POINTER(p)]),
        "vs_world_destroy": (None, [p]),
        "vs_world_error": (C.c_char_p, [p]),
        "vs_world_add_scene": (C.c_int, [p, C.POINTER(VsSceneDesc), C.POINTER(u32)]),
        "vs_world_remove_scene": (C.c_int, [p, u32]),
        "vs_world_reset_scene": (C.c_int, [p, u32]),
        "vs_world_set_target": (C.c_int, [p, u32, C.POINTER(f), C.POINTER(f)]),
        "vs_world_frame": (C.c_int, [p, u32, i32, i32]),
        "vs_world_frame_scene": (C.c_int, [p, u32, u32, i32]),
        "vs_world_synchronize": (C.c_int, [p]),
        "vs_world_record": (C.c_int, [p, C.POINTER(u32)]),
        "vs_world_elapsed": (C.c_int, [p, u32, u32, C.POINTER(f)]),
        "vs_world_read_nodes": (C.c_int, [p, u32, p, p]),
        "vs_world_read_splats": (C.c_int, [p, u32, p, p]),
        "vs_world_read_splat_axes": (C.c_int, [p, u32, p, p, p]),
        "vs_world_body_info": (C.c_int, [p, u32, u32, C.POINTER(VsBodyInfo)]),
        "vs_world_read_lattice": (C.c_int, [p, u32, u32, p, p, p, p]),
        "vs_world_log_pairs": (C.c_int, [p, u32]),
        "vs_world_read_contacts": (C.c_int, [p, u32, p, p, p]),
        "vs_world_stats_get": (C.c_int, [p, C.POINTER(VsWorldStats)]),
        "vs_world_autotune": (C.c_int, [p, u32, C.c_char_p, u32]),
        "vs_world_profile": (C.c_int, [p, u32, i32, p, p]),
        "vs_world_time_pass": (C.c_int, [p, i32, u32, C.POINTER(C.c_float), C.POINTER(u32)]),
        "vs_server_create": (C.c_int, [p, C.POINTER(VsServerConfig), C.POINTER(p)]),
        "vs_server_destroy": (None, [p]),
        "vs_server_error": (C.c_char_p, [p]),
        "vs_server_submit": (C.c_int, [p, C.POINTER(VsSceneDesc), C.c_double, u32, p, C.POINTER(u32)]),
        "vs_server_run": (C.c_int, [p, i32, C.POINTER(VsServerReport)]),
        "vs_server_latencies": (C.c_int, [p, p, u64, C.POINTER(u64)]),
        "vs_server_result": (C.c_int, [p, u32, p]),
        "vs_server_times": (C.c_int, [p, u32, p, p, p]),
        "vs_voxelize_cloud": (C.c_int, [C.c_int, p, p, p, u64, C.c_double, u32, C.c_double, u32, i32,
                                        C.POINTER(VsOccupancy)]),
        "vs_voxelize_mesh": (C.c_int, [C.c_int, p, u64, p, u64, p, C.c_double, u32, i32, C.POINTER(VsOccupancy)]),
        "vs_voxelize_shape": (C.c_int, [C.c_int, C.POINTER(VsShape), C.c_double, u32, C.POINTER(VsOccupancy)]),
        "vs_occupancy_free": (None, [C.POINTER(VsOccupancy)]),
        "vs_ingest_error": (C.c_char_p, []),
        "vs_ply_read": (C.c_int, [p, u64, C.POINTER(C.c_char_p), u32, C.POINTER(u64), C.POINTER(C.POINTER(f))]),
        "vs_ply_properties": (C.c_int, [p, u64, C.c_char_p, u64, C.POINTER(u64)]),
        "vs_cloud_from_columns": (C.c_int, [u64, C.POINTER(C.POINTER(f)), C.POINTER(VsCloud)]),
        "vs_cloud_from_ply": (C.c_int, [p, u64, C.POINTER(VsCloud)]),
        "vs_cloud_free": (None, [C.POINTER(VsCloud)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a, ctype=C.c_float):
    return a.ctypes.data_as(C.POINTER(ctype))


class VsimError(RuntimeError):
    pass


SLOTS = ["nodePos", "nodePrev", "nodeVel", "nodeInvMass", "bodies", "params", "tiles", "items", "contactGrid", "out"]


class VsGrabParams(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("direction", C.c_double * 3), ("radius", C.c_double),
                ("anchor", C.c_uint32), ("begin", C.c_uint32), ("end", C.c_uint32), ("pad0", C.c_uint32)]


class Backend:
    """Layer 1 (ComputeBackend's seam): device buffers and named launches."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.vs_backend_create(device, C.byref(h))
        if rc != VS_OK:
            raise VsimError(f"vs_backend_create: {self.lib.vs_last_error(None).decode()} ({rc})")
        self.h = h

    def _check(self, rc, what):
        if rc != VS_OK:
            raise VsimError(f"{what}: {self.lib.vs_last_error(self.h).decode()} ({rc})")

    def close(self):
        if self.h:
            self.lib.vs_backend_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc(self, nbytes):
        out = C.c_uint64()
        self._check(self.lib.vs_buffer_alloc(self.h, int(nbytes), C.byref(out)), "alloc")
        return out.value

    def free(self, handle):
        self._check(self.lib.vs_buffer_free(self.h, handle), "free")

    def write(self, handle, data, offset=0):
        buf = bytes(data) if isinstance(data, C.Structure) else np.ascontiguousarray(data).tobytes()
        self._check(self.lib.vs_buffer_write(self.h, handle, offset, buf, len(buf)), "write")

    def read(self, handle, nbytes, offset=0):
        out = (C.c_char * nbytes)()
        self._check(self.lib.vs_buffer_read(self.h, handle, offset, nbytes, out), "read")
        return bytes(out)

    def dispatch(self, *launches):
        arr = (VsLaunch * len(launches))(*launches)
        self._check(self.lib.vs_dispatch(self.h, arr, len(launches)), "dispatch")

    def sync(self):
        self._check(self.lib.vs_synchronize(self.h), "synchronize")

    def grip_noalloc(self, pos, plane, mass, params, limit, grid, out):
        """vs_grip_noalloc: the grip with the library's static candidate
        array; `params` a VsGrabParams, buffers by handle."""
        self._check(self.lib.vs_grip_noalloc(self.h, pos, plane, mass, C.byref(params), limit, grid, out),
                    "grip_noalloc")

    def launch(self, kernel, grid, block=64, plane=0, items=0, flags=0, buckets=0, pair_cap=0, **buffers):
        """A vs_launch record; `buffers` by slot name (nodePos=..., out=...)."""
        kid = self.lib.vs_kernel_id(kernel.encode())
        if kid < 0:
            raise VsimError(f"unknown kernel {kernel}")
        l = VsLaunch(kernel=kid, grid=grid, block=block, plane=plane, items=items, flags=flags, buckets=buckets,
                     pair_cap=pair_cap)
        for name, handle in buffers.items():
            l.buffers[SLOTS.index(name)] = handle
        return l


class World:
    """A vs_world: scenes resident on one GPU, stepped batched or plain."""

    def __init__(self, *, max_scenes=16, max_bodies=16, max_nodes=1 << 20, max_cells=1 << 20,
                 max_surface=1 << 16, max_buckets=1 << 16, max_driven=4096, max_splats=1 << 20,
                 device=0, block_cells=256, tile_cells=0, persistent=0, xcd_chunked=0, separate_node_passes=0,
                 separate_colours=0, colour_launches=0, branches=0):
        self.lib = load()
        cfg = VsWorldConfig(device=device, max_scenes=max_scenes, max_bodies=max_bodies, max_nodes=max_nodes,
                            max_cells=max_cells, max_surface=max_surface, max_buckets=max_buckets,
                            max_driven=max_driven, max_splats=max_splats, block_cells=block_cells,
                            tile_cells=tile_cells, persistent=persistent, xcd_chunked=xcd_chunked,
                            separate_node_passes=separate_node_passes, separate_colours=separate_colours,
                            colour_launches=colour_launches, branches=branches)
        h = C.c_void_p()
        rc = self.lib.vs_world_create(C.byref(cfg), C.byref(h))
        if rc != VS_OK:
            raise VsimError(f"vs_world_create failed ({rc})")
        self.h = h
        self.scenes = {}

    def _check(self, rc, what):
        if rc != VS_OK:
            raise VsimError(f"{what}: {self.lib.vs_world_error(self.h).decode()} ({rc})")

    def close(self):
        if self.h:
            self.lib.vs_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, scene, device_lattice=False):
        """Admits a scenes.Scene; returns its id (raises MemoryError on VS_E_NOMEM).
        `device_lattice`: send occupancy grids and build the lattices on the device."""
        desc, keep = scene.to_desc(device_lattice)
        sid = C.c_uint32()
        rc = self.lib.vs_world_add_scene(self.h, C.byref(desc), C.byref(sid))
        del keep
        if rc == VS_E_NOMEM:
            raise MemoryError(self.lib.vs_world_error(self.h).decode())
        self._check(rc, "add_scene")
        self.scenes[sid.value] = scene
        return sid.value

    def try_add(self, scene, device_lattice=False):
        try:
            return self.add(scene, device_lattice)
        except MemoryError:
            return None

    def remove(self, sid):
        self._check(self.lib.vs_world_remove_scene(self.h, sid), "remove_scene")
        self.scenes.pop(sid, None)

    def reset(self, sid):
        self._check(self.lib.vs_world_reset_scene(self.h, sid), "reset_scene")

    def set_target(self, sid, target, velocity):
        t = (C.c_float * 3)(*target)
        v = (C.c_float * 3)(*velocity)
        self._check(self.lib.vs_world_set_target(self.h, sid, t, v), "set_target")

    def frame(self, substeps, splats=False, graph=False):
        self._check(self.lib.vs_world_frame(self.h, substeps, int(splats), int(graph)), "frame")

    def frame_scene(self, sid, substeps, splats=False):
        self._check(self.lib.vs_world_frame_scene(self.h, sid, substeps, int(splats)), "frame_scene")

    def sync(self):
        self._check(self.lib.vs_world_synchronize(self.h), "synchronize")

    def record(self):
        slot = C.c_uint32()
        self._check(self.lib.vs_world_record(self.h, C.byref(slot)), "record")
        return slot.value

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self._check(self.lib.vs_world_elapsed(self.h, a, b, C.byref(ms)), "elapsed")
        return ms.value

    def read_nodes(self, sid):
        n = self.scenes[sid].n_nodes
        pos = np.empty((3, n), np.float32)
        vel = np.empty((3, n), np.float32)
        self._check(self.lib.vs_world_read_nodes(self.h, sid, pos.ctypes.data, vel.ctypes.data), "read_nodes")
        return pos, vel

    def read_splats(self, sid):
        g = self.scenes[sid].n_splats
        mean = np.empty((3, g), np.float32)
        cov = np.empty((6, g), np.float32)
        self._check(self.lib.vs_world_read_splats(self.h, sid, mean.ctypes.data, cov.ctypes.data), "read_splats")
        return mean, cov

    def body_info(self, sid, body):
        info = VsBodyInfo()
        self._check(self.lib.vs_world_body_info(self.h, sid, body, C.byref(info)), "body_info")
        return {"node_count": info.node_count, "n_cells": info.n_cells, "n_surface": info.n_surface,
                "color_count": list(info.color_count)}

    def read_lattice(self, sid, body):
        """The admitted lattice of one body: rest (3, N), GPU inverse mass (N,),
        cells by colour, surface nodes (body-local, empty without contact)."""
        info = self.body_info(sid, body)
        n = info["node_count"]
        rest = np.empty((3, n), np.float32)
        w = np.empty(n, np.float32)
        cells = np.empty(info["n_cells"], np.uint32)
        surf = np.empty(info["n_surface"], np.uint32)
        self._check(self.lib.vs_world_read_lattice(self.h, sid, body, rest.ctypes.data, w.ctypes.data,
                                                   cells.ctypes.data, surf.ctypes.data), "read_lattice")
        return rest, w, cells, surf, info

    def read_splat_axes(self, sid):
        """Means (3, G), rotations (4, G, xyzw) and axis lengths (3, G) of the
        last frame run with splats=SPLAT_AXES."""
        g = self.scenes[sid].n_splats
        mean = np.empty((3, g), np.float32)
        rot = np.empty((4, g), np.float32)
        axis = np.empty((3, g), np.float32)
        self._check(self.lib.vs_world_read_splat_axes(self.h, sid, mean.ctypes.data, rot.ctypes.data,
                                                      axis.ctypes.data), "read_splat_axes")
        return mean, rot, axis

    def log_pairs(self, cap):
        self._check(self.lib.vs_world_log_pairs(self.h, cap), "log_pairs")
        self.pair_cap = cap

    def read_contacts(self, sid):
        s = self.scenes[sid].n_surface_active
        keys = np.empty(s, np.uint32)
        counts = np.zeros(s, np.int32)
        cap = getattr(self, "pair_cap", 0)
        pairs = np.zeros((s, max(cap, 1)), np.uint32)
        self._check(self.lib.vs_world_read_contacts(self.h, sid, keys.ctypes.data, counts.ctypes.data,
                                                    pairs.ctypes.data if cap else None), "read_contacts")
        return keys, counts, pairs[:, :cap]

    def stats(self):
        st = VsWorldStats()
        self._check(self.lib.vs_world_stats_get(self.h, C.byref(st)), "stats")
        return {k: getattr(st, k) for k, _ in VsWorldStats._fields_ if k != "pad0"}

    def profile(self, substeps, splats=False):
        """{kind: (total ms, launches)} of one instrumented batched frame."""
        ms = np.zeros(len(PASS_KINDS), np.float32)
        n = np.zeros(len(PASS_KINDS), np.uint32)
        self._check(self.lib.vs_world_profile(self.h, substeps, int(splats), ms.ctypes.data, n.ctypes.data), "profile")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(PASS_KINDS)}

    def time_pass(self, kind, rounds):
        """(total ms, launches) of `rounds` back-to-back batched launches of one
        pass kind (elastic: eight colours per round); state is restored."""
        ms, n = C.c_float(), C.c_uint32()
        self._check(self.lib.vs_world_time_pass(self.h, PASS_KINDS.index(kind), rounds, C.byref(ms), C.byref(n)),
                    "time_pass")
        return float(ms.value), int(n.value)

    def autotune(self, substeps):
        buf = C.create_string_buffer(1 << 16)
        self._check(self.lib.vs_world_autotune(self.h, substeps, buf, len(buf)), "autotune")
        return buf.value.decode()


class Server:
    """A vs_server: continuous batching of scene requests on a World."""

    def __init__(self, world, substeps=10, splats=False, graph=True, keep_results=False):
        self.lib = world.lib
        self.world = world
        cfg = VsServerConfig(substeps=substeps, splats=int(splats), graph=int(graph), keep_results=int(keep_results))
        h = C.c_void_p()
        rc = self.lib.vs_server_create(world.h, C.byref(cfg), C.byref(h))
        if rc != VS_OK:
            raise VsimError(f"vs_server_create failed ({rc})")
        self.h = h
        self.requests = []

    def close(self):
        if self.h:
            self.lib.vs_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, scene, arrival, frames, device_lattice=False):
        desc, keep = scene.to_desc(device_lattice)
        targets = None
        if scene.drive is not None:
            fdt = scene.substeps * scene.dt
            targets = np.array([scene.drive.target(f * fdt) for f in range(frames + 1)], np.float32)
        rid = C.c_uint32()
        rc = self.lib.vs_server_submit(self.h, C.byref(desc), float(arrival), frames,
                                       targets.ctypes.data if targets is not None else None, C.byref(rid))
        del keep
        if rc != VS_OK:
            raise VsimError(f"submit: {self.lib.vs_server_error(self.h).decode()} ({rc})")
        self.requests.append(scene)
        return rid.value

    def run(self, realtime=True):
        rep = VsServerReport()
        rc = self.lib.vs_server_run(self.h, int(realtime), C.byref(rep))
        if rc != VS_OK:
            raise VsimError(f"run: {self.lib.vs_server_error(self.h).decode()} / "
                            f"{self.lib.vs_world_error(self.world.h).decode()} ({rc})")
        return {k: getattr(rep, k) for k, _ in VsServerReport._fields_}

    def latencies(self):
        n = C.c_uint64()
        self.lib.vs_server_latencies(self.h, None, 0, C.byref(n))
        out = np.empty(n.value, np.float32)
        self.lib.vs_server_latencies(self.h, out.ctypes.data, n.value, C.byref(n))
        return out

    def result(self, rid):
        sc = self.requests[rid]
        pos = np.empty((3, sc.n_nodes), np.float32)
        rc = self.lib.vs_server_result(self.h, rid, pos.ctypes.data)
        if rc != VS_OK:
            raise VsimError("no result kept for request %d" % rid)
        return pos

    def times(self, rid):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.lib.vs_server_times(self.h, rid, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value
