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
obstacle face-velocity writes of k3_write_solid_face_vel (solver3d.py:818-836).

How each part is pinned (tests/test_oracle_frame.py): a hand-worked 3-cell
case; invariants (every fluid cell ends with at least min_per_cell
particles, survivors keep their order, emitted particles lie in their cell,
counts add up); a per-particle loop for the kept set; the outflow mask
against a per-particle loop with boundary points; obstacle cells against an
FP64 per-cell distance loop, face velocities against a per-face rule
(last obstacle containing either cell).
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


def stamp_obstacles(base_marker, dx, obstacles):
    """obstacles: [(kind, centre, params, velocity)] in list order.  Returns
    (marker, solid_u, solid_v, solid_w): cells with sdf <= 0 solid, faces
    cleared then written obstacle by obstacle around their inside cells
    (k3_clear_solid_vel / k3_write_solid_face_vel, solver3d.py:803-836)."""
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
    return marker, su, sv, sw
