"""CPU restatement of the between-frame particle passes — TEST INFRASTRUCTURE.

The checker, never the product: only ``tests/`` may import it.  numpy
restatements of the reference's reseed (``sim/reseed.py:43-155`` host path,
``161-396`` GPU path) and outflow compaction (``solvers/solver3d.py:1946-1982``,
``2659-2730``), with the one choice the reference leaves open fixed the way
the HIP path fixes it: an over-full cell keeps its first ``max_per_cell``
particles in index order (the reference keeps its atomics' first arrivals on
the GPU and ``rng.choice`` picks on the host).

Also the analytic animated obstacles of prepare_frame (solver3d.py:3918-3940,
2537-2596, SDFs primitives/sdf.py:30-66): marker stamping and the per-
obstacle face-velocity writes of k3_write_solid_face_vel (solver3d.py:818-836);
and the mesh obstacles after them (solver3d.py:2598-2631): the +x ray-parity
inside test of k3_mesh_to_solid (schemes/mesh_marker.py:39-80), in FP32 with
every operation rounded on its own, by brute force over all triangles.

How each part is pinned (tests/test_oracle_frame.py): a hand-worked 3-cell
case; invariants (every fluid cell ends with at least min_per_cell
particles, survivors keep their order, emitted particles lie in their cell,
counts add up); a per-particle loop for the kept set; the outflow mask
against a per-particle loop with boundary points; obstacle cells against an
FP64 per-cell distance loop, face velocities against a per-face rule
(last obstacle containing either cell).  The mesh test
(tests/test_oracle_mesh.py): a closed cube mesh gives the analytic box, a
polyhedral sphere and torus (non-convex: rays cross up to four times) the
analytic shapes away from their chord error, reversed winding changes
nothing, and a per-cell per-triangle scalar loop agrees cell for cell.
"""
from __future__ import annotations

import numpy as np


def bin_cells(pos, dx, shape):
    """clamp(int(p/dx), 0, n-1) per axis in FP32 (reseed.py:72-74,
    166-178) and the flat key i*ny*nz + j*nz + k."""
    q = np.asarray(pos, np.float32) / np.float32(dx)
    c = [np.clip(np.clip(q[:, a], -1, shape[a]).astype(np.int64), 0, shape[a] - 1)
         for a in range(3)]
    key = (c[0] * shape[1] + c[1]) * shape[2] + c[2]
    return c, key


def reseed_plan(pos, marker, dx, min_per_cell, max_per_cell, drop_solid):
    """(alive mask, per-cell emit counts, per-cell counts).  Counts from the
    binned positions; a cell is a reseed target when fluid in the marker or
    occupied, and not solid (reseed.py:84-94, 334-343); the first
    max_per_cell particles per non-solid cell in index order survive;
    particles in solid cells are dropped (GPU path, reseed.py:200-215) or kept
    (host path, reseed.py:108-136)."""
    marker = np.asarray(marker)
    shape = marker.shape
    _, key = bin_cells(pos, dx, shape)
    C = marker.size
    counts = np.bincount(key, minlength=C).astype(np.int64)
    mflat = marker.reshape(-1)
    solid = mflat == 2
    fluid = ((mflat == 1) | (counts > 0)) & ~solid
    deficit = np.where(fluid, np.maximum(min_per_cell - counts, 0), 0)
    order = np.argsort(key, kind="stable")
    rank = np.empty(len(key), dtype=np.int64)
    if len(key):
        ks = key[order]
        starts = np.r_[0, np.flatnonzero(np.diff(ks)) + 1]
        seg = np.repeat(starts, np.diff(np.r_[starts, len(ks)]))
        rank[order] = np.arange(len(ks)) - seg
    in_solid = solid[key] if len(key) else np.zeros(0, bool)
    alive = np.where(in_solid, not drop_solid, rank < max_per_cell)
    return alive, deficit.reshape(shape), counts.reshape(shape)


def emit_positions(deficit, dx, jitter):
    """Cells with a deficit in flat order, each repeated by its count, at
    (cell + 0.5 + jitter) * dx in FP32 (reseed.py:96-104, 345-352)."""
    cells = np.repeat(np.argwhere(deficit > 0), deficit[deficit > 0], axis=0)
    return (cells.astype(np.float32) + np.float32(0.5) + jitter) * np.float32(dx)


def reseed(pos, vel, marker, dx, min_per_cell, max_per_cell, jitter_amp, rng, drop_solid,
           attrs=(), emit_fill=()):
    """Whole reseed: survivors in order, then emitted particles with zero
    velocity; attrs compacted in step and filled for the emitted ones.
    Returns (pos, vel, [attrs], n_emit, n_dropped)."""
    alive, deficit, _ = reseed_plan(pos, marker, dx, min_per_cell, max_per_cell, drop_solid)
    n_emit = int(deficit.sum())
    epos = np.zeros((0, 3), np.float32)
    if n_emit:
        jit = (rng.random((n_emit, 3)) - 0.5).astype(np.float32) * jitter_amp
        epos = emit_positions(deficit, dx, jit)
    new_pos = np.concatenate([np.asarray(pos, np.float32)[alive], epos])
    new_vel = np.concatenate([np.asarray(vel, np.float32)[alive], np.zeros_like(epos)])
    out = []
    for a, fill in zip(attrs, emit_fill):
        if a is None:
            out.append(None)
            continue
        a = np.asarray(a, np.float32)
        pad = np.full((n_emit,) + a.shape[1:], fill, np.float32)
        out.append(np.concatenate([a[alive], pad]))
    return new_pos, new_vel, out, n_emit, int((~alive).sum())


def outflow_alive(pos, boxes):
    """k3_mark_alive_outflows (solver3d.py:1946-1963): 0 inside any box
    (bounds inclusive, FP32 comparisons)."""
    p = np.asarray(pos, np.float32)
    alive = np.ones(len(p), dtype=bool)
    for lo, hi in boxes:
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        alive &= ~np.all((p >= lo) & (p <= hi), axis=1)
    return alive


def cell_centers(shape, dx):
    """FP32 cell centres (primitives/sdf.py:30-37)."""
    axes = [((np.arange(n) + 0.5) * dx) for n in shape]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    return np.stack([X, Y, Z], axis=-1).astype(np.float32)


def obstacle_sdf(grid, kind, center, params):
    """sdf_sphere / sdf_box / sdf_cylinder_y (primitives/sdf.py:40-66)."""
    c = np.asarray(center, dtype=np.float32)
    if kind == "sphere":
        return np.linalg.norm(grid - c, axis=-1) - float(params["radius"])
    if kind == "box":
        h = np.asarray(params["half_size"], dtype=np.float32)
        d = np.abs(grid - c) - h
        return np.linalg.norm(np.maximum(d, 0.0), axis=-1) + np.minimum(np.max(d, axis=-1), 0.0)
    local = grid - c
    dr = np.sqrt(local[..., 0] ** 2 + local[..., 2] ** 2) - float(params["radius"])
    da = np.abs(local[..., 1]) - float(params["half_height"])
    return (np.sqrt(np.maximum(dr, 0.0) ** 2 + np.maximum(da, 0.0) ** 2)
            + np.minimum(np.maximum(dr, da), 0.0))


def _ray_hits(v, px, py, pz):
    """Moller-Trumbore with direction (1,0,0) for one triangle v (9 FP32
    values) and FP32 ray origins (mesh_marker.py:51-76): crossing mask."""
    f32 = np.float32
    v = np.asarray(v, np.float32).reshape(9)
    e1x, e1y, e1z = v[3] - v[0], v[4] - v[1], v[5] - v[2]
    e2x, e2y, e2z = v[6] - v[0], v[7] - v[1], v[8] - v[2]
    hx, hy, hz = f32(0.0), -e2z, e2y
    det = e1x * hx + e1y * hy + e1z * hz
    if not abs(det) > f32(1.0e-12):
        return np.zeros(np.shape(px), dtype=bool)
    f = f32(1.0) / det
    sx, sy, sz = px - v[0], py - v[1], pz - v[2]
    u = f * (sx * hx + sy * hy + sz * hz)
    qx = sy * e1z - sz * e1y
    w = f * qx
    qy = sz * e1x - sx * e1z
    qz = sx * e1y - sy * e1x
    t = f * (e2x * qx + e2y * qy + e2z * qz)
    return (u >= 0) & (u <= 1) & (w >= 0) & (u + w <= 1) & (t > f32(1.0e-6))


def mesh_inside(shape, dx, tris):
    """Cells whose FP32 centre's +x ray crosses an odd number of triangles
    (k3_mesh_to_solid, mesh_marker.py:39-80); tris (n, 3, 3)."""
    g = cell_centers(shape, dx)
    px, py, pz = g[..., 0], g[..., 1], g[..., 2]
    count = np.zeros(shape, dtype=np.int64)
    with np.errstate(all="ignore"):
        for v in np.asarray(tris, np.float32).reshape(-1, 9):
            count += _ray_hits(v, px, py, pz)
    return (count & 1) == 1


def mesh_inside_loop(shape, dx, tris):
    """The same test one cell and one triangle at a time (tiny grids)."""
    g = cell_centers(shape, dx)
    out = np.zeros(shape, dtype=bool)
    tv = np.asarray(tris, np.float32).reshape(-1, 9)
    with np.errstate(all="ignore"):
        for idx in np.ndindex(*shape):
            p = g[idx]
            n = sum(bool(_ray_hits(v, p[0], p[1], p[2])) for v in tv)
            out[idx] = n % 2 == 1
    return out


def stamp_obstacles(base_marker, dx, obstacles, meshes=()):
    """obstacles: [(kind, centre, params, velocity)] in list order; meshes:
    [(triangles, velocity)] applied after them.  Returns (marker, solid_u,
    solid_v, solid_w): cells with sdf <= 0 solid, faces cleared then written
    obstacle by obstacle around their inside cells (k3_clear_solid_vel /
    k3_write_solid_face_vel, solver3d.py:803-836); then each mesh's inside
    cells that were not yet solid become solid and write the mesh's velocity
    into their six faces (solver3d.py:2598-2631)."""
    marker = np.array(base_marker, copy=True)
    nx, ny, nz = marker.shape
    su = np.zeros((nx + 1, ny, nz), np.float32)
    sv = np.zeros((nx, ny + 1, nz), np.float32)
    sw = np.zeros((nx, ny, nz + 1), np.float32)
    grid = cell_centers(marker.shape, dx)
    for kind, center, params, vel in obstacles:
        inside = obstacle_sdf(grid, kind, center, params) <= 0.0
        marker[inside] = 2
        v = np.asarray(vel, np.float32)
        su[:-1][inside] = v[0]
        su[1:][inside] = v[0]
        sv[:, :-1][inside] = v[1]
        sv[:, 1:][inside] = v[1]
        sw[:, :, :-1][inside] = v[2]
        sw[:, :, 1:][inside] = v[2]
    for tris, vel in meshes:
        new = mesh_inside(marker.shape, dx, tris) & (marker != 2)
        marker[new] = 2
        v = np.asarray(vel, np.float32)
        su[:-1][new] = v[0]
        su[1:][new] = v[0]
        sv[:, :-1][new] = v[1]
        sv[:, 1:][new] = v[1]
        sw[:, :, :-1][new] = v[2]
        sw[:, :, 1:][new] = v[2]
    return marker, su, sv, sw
