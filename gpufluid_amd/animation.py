"""Obstacle motion specs and analytic SDFs (G1.8, G1.10-G1.16), host side.

Same dataclasses, fields and arithmetic as the reference's
``primitives/animation.py:26-71`` and ``primitives/sdf.py:30-88``: callers
build motions and static obstacle SDFs with them and pass those to
``FlipSolver3D.add_animated_obstacle`` / ``add_solid_from_sdf``.  The
per-frame stamping of animated obstacles into the marker and the solid face
velocities runs on the device (``gf_anim_obstacles``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Literal, Optional, Sequence, Tuple, Union

import numpy as np

from .blocks import block


@dataclass
class LinearMotion:
    kind: Literal["linear"] = "linear"
    velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    fps: int = 24


@dataclass
class KeyframeMotion:
    kind: Literal["keyframes"] = "keyframes"
    keyframes: List[Tuple[int, Tuple[float, float, float]]] = field(default_factory=list)


Motion = Union[LinearMotion, KeyframeMotion]

block("G1.15", "Motion spec dataclasses (LinearMotion + KeyframeMotion union)")(LinearMotion)


@block("G1.16", "Resolve animated centre at a given frame (pure host transform)")
def evaluate_center(base_center: Sequence[float], motion: Optional[Motion],
                    frame_idx: int) -> np.ndarray:
    """Centre at a frame: base + t*velocity (t = frame/fps) for linear motion,
    linear interpolation between keyframes (clamped to the first and last)
    for keyframed motion (animation.py:47-71)."""
    base = np.asarray(base_center, dtype=np.float32)
    if motion is None:
        return base
    if isinstance(motion, LinearMotion):
        t = frame_idx / float(motion.fps)
        return base + t * np.asarray(motion.velocity, dtype=np.float32)
    if isinstance(motion, KeyframeMotion):
        kfs = sorted(motion.keyframes, key=lambda k: k[0])
        if not kfs:
            return base
        if frame_idx <= kfs[0][0]:
            return np.asarray(kfs[0][1], dtype=np.float32)
        if frame_idx >= kfs[-1][0]:
            return np.asarray(kfs[-1][1], dtype=np.float32)
        for (fa, pa), (fb, pb) in zip(kfs, kfs[1:]):
            if fa <= frame_idx <= fb:
                if fb == fa:
                    return np.asarray(pa, dtype=np.float32)
                t = (frame_idx - fa) / float(fb - fa)
                return (np.asarray(pa, dtype=np.float32) * (1 - t)
                        + np.asarray(pb, dtype=np.float32) * t)
        return base
    raise ValueError(f"unknown motion kind: {motion!r}")


@block("G1.8", "Cell-centre world-position grid (host)")
def cell_centers(nx: int, ny: int, nz: int, dx: float) -> np.ndarray:
    """(nx, ny, nz, 3) FP32 cell centres ((i + 0.5) * dx)."""
    i = (np.arange(nx) + 0.5) * dx
    j = (np.arange(ny) + 0.5) * dx
    k = (np.arange(nz) + 0.5) * dx
    X, Y, Z = np.meshgrid(i, j, k, indexing="ij")
    return np.stack([X, Y, Z], axis=-1).astype(np.float32)


@block("G1.10", "SDF sphere (analytic, negative inside)")
def sdf_sphere(grid_xyz, center, radius: float) -> np.ndarray:
    c = np.asarray(center, dtype=np.float32)
    return np.linalg.norm(grid_xyz - c, axis=-1) - float(radius)


@block("G1.11", "SDF axis-aligned box (analytic, negative inside)")
def sdf_box(grid_xyz, center, half_size) -> np.ndarray:
    c = np.asarray(center, dtype=np.float32)
    h = np.asarray(half_size, dtype=np.float32)
    d = np.abs(grid_xyz - c) - h
    return np.linalg.norm(np.maximum(d, 0.0), axis=-1) + np.minimum(np.max(d, axis=-1), 0.0)


@block("G1.12", "SDF cylinder aligned with Y (gravity) axis")
def sdf_cylinder_y(grid_xyz, center, radius: float, half_height: float) -> np.ndarray:
    c = np.asarray(center, dtype=np.float32)
    local = grid_xyz - c
    dr = np.sqrt(local[..., 0] ** 2 + local[..., 2] ** 2) - float(radius)
    da = np.abs(local[..., 1]) - float(half_height)
    outside = np.sqrt(np.maximum(dr, 0.0) ** 2 + np.maximum(da, 0.0) ** 2)
    return outside + np.minimum(np.maximum(dr, da), 0.0)


@block("G1.13", "SDF plane (signed distance to half-space)")
def sdf_plane(grid_xyz, point, normal) -> np.ndarray:
    p = np.asarray(point, dtype=np.float32)
    n = np.asarray(normal, dtype=np.float32)
    n = n / (np.linalg.norm(n) + 1e-12)
    return ((grid_xyz - p) * n).sum(axis=-1)


@block("G1.14", "SDF union (min of components)")
def sdf_union(*sdfs) -> np.ndarray:
    out = sdfs[0]
    for s in sdfs[1:]:
        out = np.minimum(out, s)
    return out
