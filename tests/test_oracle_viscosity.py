"""Pins of the oracle's viscosity sweep (k3_jacobi_visc, solver3d.py:178-197)
and of its place in the substep (3524-3544)."""
import numpy as np
import pytest

from oracle import flip_oracle as fo


def _assembled(shape, r):
    """(I - r L) with L u = sum over in-range neighbours of (u_nb - u)."""
    sx, sy, sz = shape
    n = sx * sy * sz
    a = np.eye(n)
    idx = np.arange(n).reshape(shape)
    for i in range(sx):
        for j in range(sy):
            for k in range(sz):
                c = idx[i, j, k]
                for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1),
                                   (0, 0, 1)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < sx and 0 <= jj < sy and 0 <= kk < sz:
                        a[c, c] += r
                        a[c, idx[ii, jj, kk]] -= r
    return a


def test_fixed_point_is_the_implicit_solve():
    rng = np.random.default_rng(1)
    shape = (4, 3, 5)
    r = 0.35
    u_old = rng.normal(size=shape)
    u = u_old.copy()
    for _ in range(400):
        u = fo.visc_sweep(u_old, u, r)
    ref = np.linalg.solve(_assembled(shape, np.float32(r)), u_old.ravel()).reshape(shape)
    np.testing.assert_allclose(u, ref, rtol=0, atol=1e-10)


def test_constant_field_and_zero_r():
    rng = np.random.default_rng(2)
    c = np.full((5, 4, 6), 1.75, dtype=np.float32)
    np.testing.assert_array_equal(fo.visc_sweep(c, c, 0.8), c)
    u_old = rng.normal(size=(5, 4, 6)).astype(np.float32)
    u_in = rng.normal(size=(5, 4, 6)).astype(np.float32)
    np.testing.assert_array_equal(fo.visc_sweep(u_old, u_in, 0.0), u_old)


def test_sweep_matches_per_entry_loop():
    rng = np.random.default_rng(3)
    shape = (4, 5, 3)
    r = np.float32(0.21)
    u_old = rng.normal(size=shape).astype(np.float32)
    u_in = rng.normal(size=shape).astype(np.float32)
    out = fo.visc_sweep(u_old, u_in, r)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                s = np.float32(0)
                n = np.float32(0)
                for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1),
                                   (0, 0, 1)):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < shape[0] and 0 <= jj < shape[1] and 0 <= kk < shape[2]:
                        s = np.float32(s + u_in[ii, jj, kk])
                        n = np.float32(n + 1)
                want = np.float32(np.float32(u_old[i, j, k] + np.float32(r * s)) /
                                  np.float32(1 + np.float32(r * n)))
                assert out[i, j, k] == want


def test_stage_swaps_like_the_reference():
    """After N sweeps us holds the last-but-one iterate; 0 sweeps leave us
    alone; the result is solid-BC'd."""
    rng = np.random.default_rng(4)
    m = fo.wall_marker(6, 5, 4)
    shp = [(7, 5, 4), (6, 6, 4), (6, 5, 5)]
    u, v, w = (rng.normal(size=s) for s in shp)
    us, vs, ws = (rng.normal(size=s) for s in shp)
    r = np.float32(0.1)
    out = fo.viscosity_stage(u, v, w, us, vs, ws, m, r, 3)
    it = [u]
    for _ in range(3):
        it.append(fo.visc_sweep(u, it[-1], r))
    bc = fo.enforce_solid_bc(it[3], v, w, m)[0]
    np.testing.assert_array_equal(out[0], bc)
    np.testing.assert_array_equal(out[3], it[2])
    out0 = fo.viscosity_stage(u, v, w, us, vs, ws, m, r, 0)
    np.testing.assert_array_equal(out0[3], us)


def test_viscous_substep_damps_shear_and_stays_divergence_free():
    """A shear flow loses kinetic energy to viscosity, and the projected field
    is divergence-free on fluid cells."""
    from gpufluid_amd.seeding import fluid_cells, seed_cells
    shape = (12, 12, 12)
    dx = 1.0 / 12
    cells = fluid_cells(shape, dx, lambda x, y, z: (y < 0.7))
    pos = seed_cells(cells, dx, np.random.default_rng(5))
    vel = np.zeros_like(pos)
    vel[:, 0] = np.sin(2 * np.pi * pos[:, 1]).astype(np.float32)
    st = {"pos": pos, "vel": vel, "marker": fo.wall_marker(*shape), "dx": dx}
    e = []
    for visc in (0.0, 0.05):
        o, g = fo.substep(st, 1 / 240, pressure="exact", viscosity=visc, viscosity_iters=20,
                          gravity=0.0, flip_blend=0.0)
        e.append(float((o["vel"] ** 2).sum()))
        div = fo.divergence(g["u"], g["v"], g["w"], g["marker"], dx, 1 / 240, 1.0)
        assert np.abs(div[g["marker"] == fo.FLUID]).max() < 1e-6 * max(np.abs(g["div"]).max(), 1)
    assert e[1] < e[0]
