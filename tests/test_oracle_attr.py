"""Pins of the oracle's per-particle attribute transfer (colour S2.15,
scalar S2.18; solver3d.py:562-755 and the drivers at 3304-3359)."""
import numpy as np
import pytest

from oracle import flip_oracle as fo


def _line_case(dx):
    # Two particles on a 4x1x1 grid, y and z at the cell centre so only the
    # x weights matter: A at fx = 0.25, B at fx = 1.5.
    pos = np.array([[0.75 * dx, 0.5 * dx, 0.5 * dx],
                    [2.0 * dx, 0.5 * dx, 0.5 * dx]], dtype=np.float32)
    return pos


@pytest.mark.parametrize("mode", [fo.ATTR_COLOR, fo.ATTR_SCALAR])
def test_two_particle_line_closed_form(mode):
    """Hand-derived: cells get w = (.75, .75, .5, 0) and values
    (a, (.25a+.5b)/.75, b, 0); A gathers .75*c0 + .25*c1, B .5*c1 + .5*c2."""
    dx = 0.125
    pos = _line_case(dx)
    a, b = 0.2, 0.8
    attr = np.array([a, b], dtype=np.float32)
    vsum, wsum = fo.attr_p2g(pos, attr, dx, (4, 1, 1))
    np.testing.assert_allclose(wsum[:, 0, 0], [0.75, 0.75, 0.5, 0.0], rtol=0, atol=1e-7)
    grid, w = fo.attr_normalize(vsum, wsum, mode)
    c1 = (0.25 * a + 0.5 * b) / 0.75
    np.testing.assert_allclose(grid[0, :, 0, 0], [a, c1, b, 0.0], rtol=0, atol=1e-6)
    out = fo.attr_g2p(pos, attr, grid, w, dx)
    np.testing.assert_allclose(out, [0.75 * a + 0.25 * c1, 0.5 * c1 + 0.5 * b],
                               rtol=0, atol=1e-6)
    np.testing.assert_allclose(out, [0.3, 0.7], rtol=0, atol=1e-6)


def test_three_channels_transfer_independently():
    dx = 0.125
    pos = _line_case(dx)
    col = np.array([[0.2, 1.0, 0.0], [0.8, 0.0, 0.5]], dtype=np.float32)
    out = fo.attr_transfer(pos, col, dx, (4, 1, 1), fo.ATTR_COLOR)
    for c in range(3):
        one = fo.attr_transfer(pos, col[:, c], dx, (4, 1, 1), fo.ATTR_SCALAR)
        np.testing.assert_allclose(out[:, c], one, rtol=0, atol=1e-7)


def test_one_particle_per_cell_centre_is_identity():
    """A particle at a cell centre puts its whole weight on that cell, and
    sampling at the centre returns the cell: the transfer is the identity.
    Fails for a missing half-cell offset or a transposed grid index."""
    rng = np.random.default_rng(3)
    shape = (5, 4, 3)
    dx = 0.1
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).reshape(-1, 3)
    pos = ((idx + 0.5) * dx).astype(np.float32)
    perm = rng.permutation(len(pos))
    pos = pos[perm]
    col = rng.random((len(pos), 3)).astype(np.float32)
    out = fo.attr_transfer(pos, col, dx, shape, fo.ATTR_COLOR)
    np.testing.assert_allclose(out, col, rtol=0, atol=2e-6)


@pytest.mark.parametrize("mode", [fo.ATTR_COLOR, fo.ATTR_SCALAR])
def test_constant_attribute_is_preserved(mode):
    rng = np.random.default_rng(5)
    shape = (8, 7, 6)
    dx = 1.0 / 8
    pos = (rng.random((400, 3)) * np.array(shape) * dx).astype(np.float32)
    val = np.array([0.3, 0.6, 0.9], dtype=np.float32)
    col = np.tile(val, (len(pos), 1))
    out = fo.attr_transfer(pos, col, dx, shape, mode)
    np.testing.assert_allclose(out, col, rtol=0, atol=1e-6)


def test_weights_sum_to_particle_count_inside():
    rng = np.random.default_rng(6)
    shape = (6, 6, 6)
    dx = 0.5
    # every particle's eight corners inside the grid: fx in [0, 5)
    pos = ((rng.random((300, 3)) * 4.9 + 0.55) * dx).astype(np.float32)
    _, wsum = fo.attr_p2g(pos, np.ones(len(pos), np.float32), dx, shape)
    assert abs(wsum.sum() - len(pos)) < 1e-4


def test_particle_without_grid_weight_keeps_its_value():
    """Sampling clamps to the grid: a particle beyond the far corner reads
    cell (5,5,5), which no particle reached, so it keeps its value; one
    beyond the near corner reads cell (0,0,0) and takes that cell's value."""
    dx = 0.25
    pos = np.array([[0.5 * dx] * 3, [3.0] * 3, [-3.0] * 3], dtype=np.float32)
    attr = np.array([0.4, 7.0, 9.0], dtype=np.float32)
    out = fo.attr_transfer(pos, attr, dx, (6, 6, 6), fo.ATTR_SCALAR)
    assert out[1] == np.float32(7.0)
    np.testing.assert_allclose(out[[0, 2]], [0.4, 0.4], rtol=0, atol=1e-6)


@pytest.mark.parametrize("mode", [fo.ATTR_COLOR, fo.ATTR_SCALAR])
def test_vectorised_matches_per_particle_loop(mode):
    rng = np.random.default_rng(11)
    shape = (6, 5, 7)
    dx = 0.2
    pos = (rng.random((60, 3)) * (np.array(shape) + 1) * dx - 0.5 * dx).astype(np.float32)
    attr = (rng.random((60, 3)) * 4 - 1).astype(np.float32)
    out = fo.attr_transfer(pos, attr, dx, shape, mode)
    ref = fo.attr_transfer_loop(pos, attr, dx, shape, mode)
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-6)
