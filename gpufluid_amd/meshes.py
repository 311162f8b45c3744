"""Triangle-mesh obstacles: file loading and the device inside test.

Mirrors the reference's ``schemes/mesh_marker.py`` surface
(``mark_solid_from_mesh_gpu``, ``mesh_indicator_sdf_gpu``,
``DEFAULT_BVH_THRESHOLD``) on the HIP kernels of ``gf_mesh.hip``: a cell is
solid when the +x ray from its centre crosses the mesh an odd number of
times (the reference's brute-force Moller-Trumbore test, mesh_marker.py:39-80),
with the triangles binned into grid columns first, which serves every mesh
size (the reference switches to a BVH winding-number query at 256 triangles;
for a closed mesh both give the inside cells).

The reference loads mesh files with trimesh (absent here);
:func:`load_triangles` reads the same Wavefront OBJ and STL files itself and
applies scale and translation in FP64 before rounding the triangles to FP32,
as trimesh's ``apply_scale`` / ``apply_translation`` followed by
``np.asarray(mesh.triangles, np.float32)`` do (solver3d.py:2607-2613).
"""
from __future__ import annotations

import ctypes
import struct
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .blocks import block
from .device import DeviceArray, as_tensor, default_device

DEFAULT_BVH_THRESHOLD = 256


def _parse_obj(text: str) -> np.ndarray:
    verts, tris = [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([float(x) for x in parts[1:4]])
        elif parts[0] == "f":
            idx = []
            for ref in parts[1:]:
                i = int(ref.split("/")[0])
                idx.append(i - 1 if i > 0 else len(verts) + i)
            for a in range(1, len(idx) - 1):       # polygons as a fan from the first vertex
                tris.append((idx[0], idx[a], idx[a + 1]))
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    if not tris:
        return np.zeros((0, 3, 3))
    t = np.asarray(tris, dtype=np.int64)
    if t.min() < 0 or t.max() >= len(v):
        raise ValueError("OBJ face refers to a missing vertex")
    return v[t]


def _parse_stl(data: bytes) -> np.ndarray:
    if len(data) >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if 84 + 50 * n == len(data):             # binary: 80-byte header, count, 50-byte records
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)),
                                                      ("a", "<u2")]), count=n, offset=84)
            return rec["v"].astype(np.float64)
    verts = [[float(x) for x in line.split()[1:4]]
             for line in data.decode("ascii", errors="replace").splitlines()
             if line.strip().startswith("vertex")]
    if len(verts) % 3:
        raise ValueError("ASCII STL with a vertex count that is not a multiple of 3")
    return np.asarray(verts, dtype=np.float64).reshape(-1, 3, 3)


def read_mesh(path) -> np.ndarray:
    """(n, 3, 3) FP64 triangles of an OBJ or STL file, as stored."""
    p = Path(path)
    suffix = p.suffix.lower()
    if suffix == ".obj":
        return _parse_obj(p.read_text())
    if suffix == ".stl":
        return _parse_stl(p.read_bytes())
    raise ValueError(f"unsupported mesh file type: {p.suffix!r} (OBJ and STL are read)")


def place(tri64, translate) -> np.ndarray:
    """FP64 triangles plus an FP32 translation, rounded to FP32."""
    return (np.asarray(tri64, np.float64)
            + np.asarray(translate, dtype=np.float32).astype(np.float64)).astype(np.float32)


def load_triangles(path, scale: float = 1.0, translate=(0.0, 0.0, 0.0)) -> np.ndarray:
    """(n, 3, 3) FP32 triangles of an OBJ or STL file: vertices in FP64,
    times ``scale``, plus ``translate`` (FP32 values), then rounded to FP32."""
    return place(read_mesh(path) * float(scale), translate)


def _regular_dx(cell_centers_np) -> float:
    """The cell size of a regular (i + 0.5) * dx centre grid, checked."""
    c = np.asarray(cell_centers_np, dtype=np.float32)
    if c.ndim != 4 or c.shape[3] != 3:
        raise ValueError("cell centres must have shape (nx, ny, nz, 3)")
    nx = c.shape[0]
    first = float(c[0, 0, 0, 0])
    cands = [2.0 * first]
    if nx > 1:
        cands.insert(0, (float(c[-1, 0, 0, 0]) - first) / (nx - 1))
    cands += [1.0 / round(1.0 / d) for d in cands if d > 0 and round(1.0 / d) > 0]
    from .animation import cell_centers
    for dx in cands:
        if dx > 0 and np.array_equal(cell_centers(*c.shape[:3], dx), c):
            return float(dx)
    raise ValueError("cell centres must be the regular grid (i + 0.5) * dx")


def mark_mesh(marker: torch.Tensor, triangles, dx: float, faces=None, velocity=(0.0, 0.0, 0.0),
              stream=None) -> None:
    """Set ``marker`` (int32 (nx, ny, nz), on the device) to 2 inside the
    mesh where it is not 2 already; with ``faces`` = (solid_u, solid_v,
    solid_w), write ``velocity`` into the six faces of each such cell."""
    dev = marker.device
    nx, ny, nz = (int(s) for s in marker.shape)
    tri = as_tensor(np.ascontiguousarray(np.asarray(triangles, np.float32).reshape(-1, 9)),
                    torch.float32, dev).contiguous()
    n = int(tri.shape[0])
    st = stream if stream is not None else _lib.stream_handle(dev)
    nws = ctypes.c_int64()
    _lib.call("D4.3.GPU", "gf_mesh_workspace", nx, ny, nz, ctypes.byref(nws))
    ws = torch.empty(nws.value, dtype=torch.int32, device=dev)
    n_ent = ctypes.c_int64()
    tp = _lib.ptr(tri) if n else None
    _lib.call("D4.3.GPU", "gf_mesh_bin_sync", tp, n, nx, ny, nz, float(dx), _lib.ptr(ws),
              ctypes.byref(n_ent), st)
    lst = torch.empty(max(int(n_ent.value), 1), dtype=torch.int32, device=dev)
    su = sv = sw = None
    if faces is not None:
        su, sv, sw = (_lib.ptr(f) for f in faces)
    vx, vy, vz = (float(x) for x in velocity)
    _lib.call("D4.3.GPU", "gf_mesh_mark", tp, n, nx, ny, nz, float(dx), _lib.ptr(ws),
              _lib.ptr(lst), int(n_ent.value), _lib.ptr(marker), su, sv, sw, vx, vy, vz, st)


@block("D4.3.GPU", "Compute solid-cell mask of a triangle mesh on GPU")
def mark_solid_from_mesh_gpu(marker_wp, cell_centers_np, triangles_np, device=None,
                             use_bvh=None, cache_key=None) -> None:
    """Set every cell of ``marker_wp`` (int (nx, ny, nz)) inside the closed
    mesh ``triangles_np`` ((n, 3, 3)) to 2 (mesh_marker.py:197-236).
    ``cell_centers_np`` must be the regular grid of centres; ``use_bvh`` and
    ``cache_key`` are accepted for the reference's signature (one method
    serves every mesh size)."""
    dx = _regular_dx(cell_centers_np)
    t = marker_wp.t if isinstance(marker_wp, DeviceArray) else marker_wp
    if not (isinstance(t, torch.Tensor) and t.device.type == "cuda" and t.dtype == torch.int32
            and t.is_contiguous()):
        raise TypeError("marker must be a contiguous int32 device array")
    if tuple(t.shape) != tuple(np.shape(cell_centers_np)[:3]):
        raise ValueError("marker and cell centres have different shapes")
    mark_mesh(t, triangles_np, dx)
    torch.cuda.current_stream(t.device).synchronize()


@block("D4.3.GPU.BVH", "Build a +/-dx signed indicator from a mesh (GPU inside test)")
def mesh_indicator_sdf_gpu(cell_centers_np, triangles_np, dx: float, device=None,
                           cache_key=None) -> np.ndarray:
    """(nx, ny, nz) FP32 array: -dx inside the mesh, +dx outside
    (mesh_marker.py:170-194), for ``add_solid_from_sdf``."""
    cdx = _regular_dx(cell_centers_np)
    dev = torch.device(device) if device is not None else default_device()
    shape = tuple(np.shape(cell_centers_np)[:3])
    marker = torch.zeros(shape, dtype=torch.int32, device=dev)
    mark_mesh(marker, triangles_np, cdx)
    inside = marker.cpu().numpy() == 2
    sdf = np.full(shape, float(dx), dtype=np.float32)
    sdf[inside] = -float(dx)
    return sdf
