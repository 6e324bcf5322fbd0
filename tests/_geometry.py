"""Synthetic capture geometry shared by the voxeliser tests: the closed box
and subdivided icosahedron meshes and the Fibonacci sphere cloud the
reference's voxeliser tests build (packages/visionary-sim/test/unit/
voxelize.test.js, cloudVolume.test.js), restated."""
import math

import numpy as np


BOX_INDICES = np.array([0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                        3, 7, 6, 3, 6, 2, 0, 4, 7, 0, 7, 3, 1, 2, 6, 1, 6, 5], np.uint32)


def box_mesh(half):
    p = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float32) * half
    return p, BOX_INDICES


def icosphere(radius, level):
    t = (1 + 5 ** 0.5) / 2
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    v = [list(np.array(x) / np.linalg.norm(x)) for x in v]
    for _ in range(level):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (np.array(v[a]) + np.array(v[b])) / 2
                v.append(list(m / np.linalg.norm(m)))
                cache[key] = len(v) - 1
            return cache[key]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return (np.array(v) * radius).astype(np.float32), np.array(f, np.uint32).reshape(-1)


def sphere_shell(radius, count, cover=1.0):
    golden = math.pi * (3 - math.sqrt(5))
    pts = []
    for i in range(count):
        y = 1 - (2 * (i + 0.5)) / count
        if y > 2 * cover - 1:
            continue
        ring = math.sqrt(max(0.0, 1 - y * y))
        theta = golden * i
        pts.append([math.cos(theta) * ring * radius, y * radius, math.sin(theta) * ring * radius])
    return np.array(pts, np.float32).reshape(-1, 3)


def cloud_at(points, radius):
    n = len(points)
    rot = np.zeros((n, 4), np.float32)
    rot[:, 3] = 1
    return points, rot, np.full((n, 3), radius, np.float32)
