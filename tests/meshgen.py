"""Closed triangle meshes for the mesh-obstacle tests (not a test module):
(vertices, faces) generators, triangle arrays, and OBJ / STL writers."""
from __future__ import annotations

import struct

import numpy as np


def cube(lo, hi):
    """Axis-aligned box, 12 outward-wound triangles."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    v = np.array([[lo[0] if b & 1 == 0 else hi[0], lo[1] if b & 2 == 0 else hi[1],
                   lo[2] if b & 4 == 0 else hi[2]] for b in range(8)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    return v, quads


def icosphere(center, radius, level=2):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[key] = len(v) - 1
            return mid[key]
        nf = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    verts = np.asarray(center, float) + radius * np.asarray(v)
    return verts, f


def torus(center, major, minor, nu=32, nv=16):
    """Torus around the y axis, quads split into triangles."""
    verts = []
    for a in range(nu):
        u = 2 * np.pi * a / nu
        for b in range(nv):
            w = 2 * np.pi * b / nv
            rr = major + minor * np.cos(w)
            verts.append((rr * np.cos(u), minor * np.sin(w), rr * np.sin(u)))
    idx = lambda a, b: (a % nu) * nv + (b % nv)  # noqa: E731
    quads = [(idx(a, b), idx(a + 1, b), idx(a + 1, b + 1), idx(a, b + 1))
             for a in range(nu) for b in range(nv)]
    return np.asarray(center, float) + np.asarray(verts), quads


def triangles(verts, faces):
    """(n, 3, 3) FP64 triangles, polygons fanned from their first vertex."""
    tris = []
    for face in faces:
        for a in range(1, len(face) - 1):
            tris.append([verts[face[0]], verts[face[a]], verts[face[a + 1]]])
    return np.asarray(tris, float)


def rotation(seed):
    q = np.random.default_rng(seed).normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def write_obj(path, verts, faces):
    with open(path, "w") as fh:
        fh.write("# test mesh\n")
        for p in verts:
            fh.write(f"v {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")
        for face in faces:
            fh.write("f " + " ".join(f"{i + 1}/{i + 1}" for i in face) + "\n")


def write_stl_binary(path, tris):
    tris = np.asarray(tris, np.float32)
    with open(path, "wb") as fh:
        fh.write(b"binary test mesh".ljust(80, b" "))
        fh.write(struct.pack("<I", len(tris)))
        for t in tris:
            fh.write(struct.pack("<3f", 0.0, 0.0, 0.0))
            fh.write(t.astype("<f4").tobytes())
            fh.write(struct.pack("<H", 0))


def write_stl_ascii(path, tris):
    with open(path, "w") as fh:
        fh.write("solid test\n")
        for t in np.asarray(tris, float):
            fh.write("  facet normal 0 0 0\n    outer loop\n")
            for p in t:
                fh.write(f"      vertex {float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")
            fh.write("    endloop\n  endfacet\n")
        fh.write("endsolid test\n")
