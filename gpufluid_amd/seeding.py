"""Particle seeders and the benchmark scene recipes.

``seed_box_positions`` is the reference box seeder (FlipSolver3D.seed_box,
src/gpufluid/solvers/solver3d.py:2826-2910) vectorised: cells
[int(lo/dx), int(hi/dx)) per axis, per_axis = round(ppc^(1/3)) sub-positions,
jitter 0.3*(U-0.5)/per_axis drawn from default_rng(0) as x, y, z per particle
in i, j, k, sx, sy, sz order.

``seed_cells`` is the scene recipe used by the benchmark configurations:
fluid cells are interior cells (1..n-2) whose centre satisfies a predicate;
each gets 2x2x2 particles at sub-positions (s+0.5)/2, jittered U(-0.25,0.25)*dx
per axis, drawn from default_rng(seed) as x, y, z per particle in
i, j, k, sx, sy, sz order.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def seed_box_positions(lo, hi, ppc: int, dx: float) -> np.ndarray:
    i0, i1 = int(lo[0] / dx), int(hi[0] / dx)
    j0, j1 = int(lo[1] / dx), int(hi[1] / dx)
    k0, k1 = int(lo[2] / dx), int(hi[2] / dx)
    per = int(round(ppc ** (1.0 / 3.0)))
    ni, nj, nk = max(i1 - i0, 0), max(j1 - j0, 0), max(k1 - k0, 0)
    if ni * nj * nk * per == 0:
        return np.zeros((0, 3), dtype=np.float32)
    rng = np.random.default_rng(0)
    u = rng.random((ni, nj, nk, per, per, per, 3))
    i = np.arange(i0, i1).reshape(-1, 1, 1, 1, 1, 1)
    j = np.arange(j0, j1).reshape(1, -1, 1, 1, 1, 1)
    k = np.arange(k0, k1).reshape(1, 1, -1, 1, 1, 1)
    s = np.arange(per)
    sx = s.reshape(1, 1, 1, -1, 1, 1)
    sy = s.reshape(1, 1, 1, 1, -1, 1)
    sz = s.reshape(1, 1, 1, 1, 1, -1)
    px = (i + (sx + 0.5 + 0.3 * (u[..., 0] - 0.5)) / per) * dx
    py = (j + (sy + 0.5 + 0.3 * (u[..., 1] - 0.5)) / per) * dx
    pz = (k + (sz + 0.5 + 0.3 * (u[..., 2] - 0.5)) / per) * dx
    return np.stack([px, py, pz], axis=-1).reshape(-1, 3).astype(np.float32)


def fluid_cells(shape, dx: float, predicate) -> np.ndarray:
    """(M,3) int cell indices, C order, of interior cells whose centre passes
    ``predicate(x, y, z)`` (vectorised over arrays)."""
    nx, ny, nz = shape
    i, j, k = np.meshgrid(np.arange(1, nx - 1), np.arange(1, ny - 1), np.arange(1, nz - 1),
                          indexing="ij")
    x, y, z = (i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx
    mask = predicate(x, y, z)
    return np.stack([i[mask], j[mask], k[mask]], axis=1)


def seed_cells(cells: np.ndarray, dx: float, rng: np.random.Generator, per_axis: int = 2,
               jitter: float = 0.25) -> np.ndarray:
    n = len(cells)
    if n == 0:
        return np.zeros((0, 3), dtype=np.float32)
    u = rng.random((n, per_axis, per_axis, per_axis, 3))
    s = (np.arange(per_axis) + 0.5) / per_axis
    sub = np.stack(np.meshgrid(s, s, s, indexing="ij"), axis=-1)  # (p,p,p,3)
    base = cells.reshape(n, 1, 1, 1, 3).astype(np.float64)
    pos = (base + sub[None] + (u - 0.5) * 2.0 * jitter) * dx
    return pos.reshape(-1, 3).astype(np.float32)


@dataclass
class Scene:
    """A seeded scene: grid, particles and the physics parameters."""
    name: str
    shape: tuple
    dx: float
    pos: np.ndarray
    vel: np.ndarray
    params: dict = field(default_factory=dict)

    @property
    def n_particles(self) -> int:
        return len(self.pos)


def dam_break(n: int, seed: int, x_max=0.4, y_max=0.6) -> Scene:
    """Dam break in a unit box: fluid x < x_max, y < y_max, full z."""
    dx = 1.0 / n
    cells = fluid_cells((n, n, n), dx, lambda x, y, z: (x < x_max) & (y < y_max))
    pos = seed_cells(cells, dx, np.random.default_rng(seed))
    return Scene(f"dam_break_{n}", (n, n, n), dx, pos, np.zeros_like(pos))


def config_c1() -> Scene:
    """32^3 FLIP dam break, seed 0, dt 1/240, FLIP 0.95, PCG to 1e-6, 100 substeps."""
    s = dam_break(32, seed=0)
    s.name = "c1_flip_dam_break_32"
    s.params = dict(transfer="flip", dt=1.0 / 240.0, flip_blend=0.95, tol=1e-6, substeps=100,
                    precision="fp32")
    return s


def config_c2() -> Scene:
    """128^3 APIC dam break, seed 1, dt 1/480, FP32 PCG + FP64 refinement to 1e-8
    (no iteration cap in the recipe: 2000 is a safety bound, never reached)."""
    s = dam_break(128, seed=1)
    s.name = "c2_apic_dam_break_128"
    s.params = dict(transfer="apic", dt=1.0 / 480.0, tol=1e-8, max_iter=2000, substeps=50,
                    precision="mixed")
    return s


def pool_and_sphere(n: int = 256, seed: int = 2) -> Scene:
    """Pool y < 0.3 plus a sphere r = 0.12 at (0.5, 0.7, 0.5) moving at (0,-2,0)."""
    dx = 1.0 / n

    def in_sphere(x, y, z):
        return (x - 0.5) ** 2 + (y - 0.7) ** 2 + (z - 0.5) ** 2 < 0.12 ** 2

    cells = fluid_cells((n, n, n), dx, lambda x, y, z: (y < 0.3) | in_sphere(x, y, z))
    pos = seed_cells(cells, dx, np.random.default_rng(seed))
    c = (cells + 0.5) * dx
    sph = in_sphere(c[:, 0], c[:, 1], c[:, 2]) & ~(c[:, 1] < 0.3)
    vel = np.zeros_like(pos)
    vel[np.repeat(sph, 8), 1] = -2.0
    return Scene(f"pool_sphere_{n}", (n, n, n), dx, pos, vel)


def config_c3() -> Scene:
    """256^3 pool + sphere, seed 2, mixed-precision PCG to 1e-8, cap 500."""
    s = pool_and_sphere(256, 2)
    s.name = "c3_pool_sphere_256"
    s.params = dict(transfer="flip", dt=1.0 / 240.0, flip_blend=0.95, tol=1e-8, max_iter=500,
                    precision="mixed")
    return s


def config_c4() -> Scene:
    """64^3 box x in [0.2,0.5], y < 0.5, seed 3; velocities N(0, 0.5^2) drawn from the
    same generator after the positions (the differentiable-rollout parameter)."""
    n = 64
    dx = 1.0 / n
    cells = fluid_cells((n, n, n), dx, lambda x, y, z: (x >= 0.2) & (x <= 0.5) & (y < 0.5))
    rng = np.random.default_rng(3)
    pos = seed_cells(cells, dx, rng)
    vel = rng.normal(0.0, 0.5, size=pos.shape).astype(np.float32)
    return Scene("c4_rollout_64", (n, n, n), dx, pos, vel,
                 dict(transfer="flip", dt=1.0 / 240.0, substeps=40, checkpoint_every=10,
                      target=(0.7, 0.3, 0.5)))


def config_c5(n_scenes: int = 256, seed: int = 4) -> list:
    """Batched scenes: edge uniform in {24,32,48,64,96}; a fluid box per scene with
    sides U[0.2,0.7] of the domain and its lower corner U[0, 1-side], all drawn
    from default_rng(seed) in that order per scene; particle jitter from
    default_rng(seed*100000 + scene index)."""
    rng = np.random.default_rng(seed)
    scenes = []
    for s in range(n_scenes):
        n = int(rng.choice([24, 32, 48, 64, 96]))
        side = rng.uniform(0.2, 0.7, 3)
        lo = rng.uniform(0.0, 1.0 - side)
        hi = lo + side
        dx = 1.0 / n
        cells = fluid_cells((n, n, n), dx,
                            lambda x, y, z: (x >= lo[0]) & (x < hi[0]) & (y >= lo[1])
                            & (y < hi[1]) & (z >= lo[2]) & (z < hi[2]))
        pos = seed_cells(cells, dx, np.random.default_rng(seed * 100000 + s))
        scenes.append(Scene(f"c5_scene_{s}", (n, n, n), dx, pos, np.zeros_like(pos),
                            dict(transfer="flip", dt=1.0 / 240.0, tol=1e-8, substeps=20)))
    return scenes
