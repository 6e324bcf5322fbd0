"""A handful of free nodes in the oracle's scene layout, for the per-pass
known-answer tests (no lattice, no cells)."""
import ctypes as C

import numpy as np

from oracle import oracle as O


class Raw:
    def __init__(self, pos, prev=None, w=None, bodies=((0.0, 0.0),), body_of=None, precision="f64", dt=1 / 1200,
                 vel=None):
        dt_ = np.float32 if precision == "f32" else np.float64
        self.dtype = dt_
        self.lib = O.load(precision)
        pos = np.asarray(pos, float)
        n = len(pos)
        self.p = np.array(pos.T, dt_, order="C", copy=True)
        self.q = np.array((pos if prev is None else np.asarray(prev, float)).T, dt_, order="C", copy=True)
        self.v = np.array(np.zeros((3, n)) if vel is None else np.asarray(vel, float).T, dt_, order="C", copy=True)
        self.w = np.ascontiguousarray(np.ones(n) if w is None else w, dt_)
        body_of = np.zeros(n, int) if body_of is None else np.asarray(body_of)
        ob = (O.OBody * len(bodies))()
        for b, (friction, damping) in enumerate(bodies):
            idx = np.nonzero(body_of == b)[0]
            ob[b].node_offset = int(idx.min()) if len(idx) else 0
            ob[b].node_count = len(idx)
            ob[b].friction, ob[b].damping = friction, damping
        self.bodies = ob
        sc = O.OScene()
        sc.n_nodes, sc.n_bodies = n, len(bodies)
        sc.px, sc.py, sc.pz = (self.p[k].ctypes.data for k in range(3))
        sc.qx, sc.qy, sc.qz = (self.q[k].ctypes.data for k in range(3))
        sc.vx, sc.vy, sc.vz = (self.v[k].ctypes.data for k in range(3))
        sc.w = self.w.ctypes.data
        sc.bodies = C.cast(ob, C.c_void_p)
        sc.dt = dt
        sc.tank_min[:] = (-1e3,) * 3
        sc.tank_max[:] = (1e3,) * 3
        sc.gain = 1.0
        sc.bucket_slots = 32
        self.sc = sc
        self.body_of = body_of

    def surface(self, nodes, radius, origin=(-1.0, -1.0, -1.0), size=0.02, dims=(100, 100, 100), pair_cap=64):
        self.s_node = np.ascontiguousarray(nodes, np.uint32)
        self.s_body = np.ascontiguousarray(self.body_of[list(nodes)], np.uint32)
        self.s_radius = np.ascontiguousarray(np.full(len(nodes), radius), self.dtype)
        S = len(nodes)
        self.keys = np.zeros(S, np.uint32)
        self.pair_count = np.zeros(S, np.int32)
        self.pairs = np.zeros((S, pair_cap), np.uint32)
        sc = self.sc
        sc.n_surf, sc.contact = S, 1
        sc.s_node, sc.s_body, sc.s_radius = self.s_node.ctypes.data, self.s_body.ctypes.data, self.s_radius.ctypes.data
        sc.grid_origin[:] = origin
        sc.bucket_size = size
        sc.grid_dims[:] = dims
        sc.key_out = self.keys.ctypes.data
        sc.pair_count = self.pair_count.ctypes.data
        sc.pairs = self.pairs.ctypes.data
        sc.pair_cap = pair_cap
        return self

    def drive(self, nodes, offsets, target, velocity, law=0, response_time=0.05):
        self.d_node = np.ascontiguousarray(nodes, np.uint32)
        self.d_off = np.ascontiguousarray(np.asarray(offsets, float).reshape(-1), self.dtype)
        sc = self.sc
        sc.n_driven, sc.drive_law = len(nodes), law
        sc.d_node, sc.d_off = self.d_node.ctypes.data, self.d_off.ctypes.data
        sc.target[:] = target
        sc.target_vel[:] = velocity
        sc.response_time = response_time
        return self

    def call(self, name):
        getattr(self.lib, name)(C.byref(self.sc))
        return self

    @property
    def pos(self):
        return self.p.T.copy()
