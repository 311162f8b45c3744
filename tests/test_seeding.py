"""Seeders and benchmark scene recipes (no GPU).

The vectorised seeders are pinned against per-particle loops of the same
recipe, and the scene sizes against the counts the configurations are quoted
with (fluid cells N_f, particles N_p = 8 N_f, active 8^3 tiles).
"""
from __future__ import annotations

import numpy as np
import pytest

from gpufluid_amd import seeding
from oracle import flip_oracle as fo


def _active_tiles(cells, tile=8):
    return len({(i // tile, j // tile, k // tile) for i, j, k in map(tuple, cells)})


def test_seed_box_matches_reference_loop():
    lo, hi, dx = (0.1, 0.05, 0.2), (0.3, 0.2, 0.33), 1.0 / 24
    for ppc in (1, 8, 27):
        a = seeding.seed_box_positions(lo, hi, ppc, dx)
        b = fo.seed_box_reference(lo, hi, ppc, dx)
        assert a.shape == b.shape
        np.testing.assert_array_equal(a, b)


def test_seed_box_empty_range():
    assert seeding.seed_box_positions((0.5, 0.5, 0.5), (0.5, 0.9, 0.9), 8, 1 / 16).shape == (0, 3)


def test_seed_cells_matches_loop():
    cells = seeding.fluid_cells((12, 12, 12), 1 / 12, lambda x, y, z: (x < 0.4) & (y < 0.6))
    a = seeding.seed_cells(cells, 1 / 12, np.random.default_rng(7))
    b = fo.seed_cells_loop(cells, 1 / 12, 7)
    np.testing.assert_array_equal(a, b)


def test_seeded_particles_stay_in_their_cell():
    cells = seeding.fluid_cells((16, 16, 16), 1 / 16, lambda x, y, z: y < 0.5)
    pos = seeding.seed_cells(cells, 1 / 16, np.random.default_rng(0))
    c = np.floor(pos.astype(np.float64) * 16).astype(int)
    np.testing.assert_array_equal(c, np.repeat(cells, 8, axis=0))


def test_config_c1_counts():
    s = seeding.config_c1()
    assert s.shape == (32, 32, 32)
    assert s.n_particles == 51_840
    cells = np.unique(np.floor(s.pos.astype(np.float64) * 32).astype(int), axis=0)
    assert len(cells) == 6_480
    assert _active_tiles(cells) == 24


def test_config_c4_counts():
    s = seeding.config_c4()
    assert s.n_particles == 292_144
    cells = np.unique(np.floor(s.pos.astype(np.float64) * 64).astype(int), axis=0)
    assert len(cells) == 36_518
    assert _active_tiles(cells) == 96
    assert abs(float(s.vel.std()) - 0.5) < 0.01


def test_config_c2_cell_count():
    cells = seeding.fluid_cells((128,) * 3, 1 / 128, lambda x, y, z: (x < 0.4) & (y < 0.6))
    assert len(cells) == 478_800  # 3,830,400 particles


@pytest.mark.slow
def test_config_c3_cell_count():
    n, dx = 256, 1 / 256

    def sphere(x, y, z):
        return (x - 0.5) ** 2 + (y - 0.7) ** 2 + (z - 0.5) ** 2 < 0.12 ** 2

    cells = seeding.fluid_cells((n,) * 3, dx, lambda x, y, z: (y < 0.3) | sphere(x, y, z))
    assert len(cells) == 5_024_592  # 40,196,736 particles
    c = (cells + 0.5) * dx
    assert int((sphere(c[:, 0], c[:, 1], c[:, 2]) & ~(c[:, 1] < 0.3)).sum()) == 121_376


def test_config_c5_is_deterministic():
    a = seeding.config_c5(n_scenes=4)
    b = seeding.config_c5(n_scenes=4)
    assert [s.shape for s in a] == [s.shape for s in b]
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.pos, y.pos)
        assert x.shape[0] in (24, 32, 48, 64, 96)
