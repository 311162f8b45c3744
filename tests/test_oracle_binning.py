"""Pins the oracle's cell binning, flat key and stable order.

Binning follows solver3d.py:96-99 (clamp(int(p/dx),0,n-1) with an FP32 divide);
the flat key and stable in-cell order follow sim/reseed.py:72-75,114.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import flip_oracle as fo

DXS = [1 / 24, 1 / 32, 1 / 48, 1 / 64, 1 / 96, 1 / 128, 1 / 256]


def _boundary_positions(n, dx):
    """p = k*dx (as FP32), one ulp either side, and the clamp limits."""
    ks = np.arange(0, n + 1)
    base = np.float32(ks * dx)
    pts = np.concatenate([base, np.nextafter(base, np.float32(np.inf)),
                          np.nextafter(base, np.float32(-np.inf)),
                          np.float32([-1.0, -1e-8, -0.0, n * dx * 2, 1e30, -1e30])])
    return pts.astype(np.float32)


def _scalar_bin(p, dx, n):
    """Brute force on numpy float32 scalars: IEEE divide, truncate, clamp."""
    if np.isnan(p):
        return 0
    q = np.float32(p) / np.float32(dx)
    if q >= n:
        return n - 1
    if q <= -1:
        return 0
    c = int(q)           # Python int() of a float truncates toward zero
    return min(max(c, 0), n - 1)


@pytest.mark.parametrize("dx", DXS)
def test_binning_matches_scalar_brute_force_on_boundaries(dx):
    n = int(round(1 / dx))
    pts = _boundary_positions(n, dx)
    pos = np.stack([pts, pts[::-1], np.roll(pts, 3)], axis=1)
    ci, cj, ck = fo.cell_index(pos, dx, (n, n, n))
    for ax, got in enumerate((ci, cj, ck)):
        want = np.array([_scalar_bin(p, dx, n) for p in pos[:, ax]])
        np.testing.assert_array_equal(got, want)


def test_fp32_divide_is_correctly_rounded():
    """The oracle's quotient is the float32 nearest to the exact quotient
    (round-to-nearest), checked with exact rational arithmetic."""
    rng = np.random.default_rng(7)
    for dx in DXS:
        d = np.float32(dx)
        ps = rng.random(200).astype(np.float32)
        q = ps / d
        for p, qi in zip(ps, q):
            exact = Fraction(float(p)) / Fraction(float(d))
            cands = [np.nextafter(qi, np.float32(-np.inf)), qi, np.nextafter(qi, np.float32(np.inf))]
            errs = [abs(Fraction(float(c)) - exact) for c in cands]
            assert errs[1] == min(errs), (p, dx)


def test_integer_multiples_bin_to_their_cell():
    """p = k*dx exactly representable (dx a power of two) lands in cell k."""
    for n in (32, 64, 128, 256):
        dx = 1.0 / n
        k = np.arange(n)
        pos = np.stack([k * dx, k * dx, k * dx], axis=1).astype(np.float32)
        ci, _, _ = fo.cell_index(pos, dx, (n, n, n))
        np.testing.assert_array_equal(ci, k)


def test_nan_bins_to_zero():
    pos = np.array([[np.nan, 0.5, np.nan]], dtype=np.float32)
    ci, cj, ck = fo.cell_index(pos, 1 / 32, (32, 32, 32))
    assert ci[0] == 0 and cj[0] == 16 and ck[0] == 0


def test_flat_key_and_stable_order_brute_force():
    rng = np.random.default_rng(3)
    shape = (6, 5, 7)
    pos = (rng.random((500, 3)) * np.array(shape) / 8).astype(np.float32)
    ci, cj, ck = fo.cell_index(pos, 1 / 8, shape)
    key = fo.flat_key(ci, cj, ck, shape)
    for t in range(0, 500, 37):
        assert key[t] == np.ravel_multi_index((ci[t], cj[t], ck[t]), shape)
    order = fo.stable_order(key)
    want = [i for _, i in sorted((int(key[i]), i) for i in range(len(key)))]
    np.testing.assert_array_equal(order, want)


def test_marker_set_algebra():
    shape = (8, 8, 8)
    m = fo.wall_marker(*shape)
    m[4, 4, 4] = fo.SOLID
    m[2, 2, 2] = fo.FLUID               # stale fluid flag must be cleared
    pos = np.array([[4.5, 4.5, 4.5], [3.2, 3.7, 5.1], [0.5, 0.5, 0.5]], dtype=np.float32) / 8
    out = fo.update_marker(m, pos, 1 / 8)
    solid = (m == fo.SOLID)
    want = np.where(solid, fo.SOLID, fo.AIR)
    want[3, 3, 5] = fo.FLUID            # only non-solid occupied cell
    np.testing.assert_array_equal(out, want)
