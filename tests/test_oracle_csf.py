"""Pins of the oracle's surface-tension stage (S2.14, solver3d.py:203-549,
3230-3300)."""
import numpy as np
import pytest

from oracle import flip_oracle as fo


def test_blur_matches_brute_force_mean():
    rng = np.random.default_rng(1)
    src = rng.random((5, 4, 6)).astype(np.float32)
    out = fo.box_blur(src)
    for i in range(5):
        for j in range(4):
            for k in range(6):
                s = np.float32(0)
                n = np.float32(0)
                for di in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        for dk in (-1, 0, 1):
                            ii, jj, kk = i + di, j + dj, k + dk
                            if 0 <= ii < 5 and 0 <= jj < 4 and 0 <= kk < 6:
                                s = np.float32(s + src[ii, jj, kk])
                                n = np.float32(n + 1)
                assert out[i, j, k] == np.float32(s / n)


def _sphere_marker(n, radius_cells):
    m = fo.wall_marker(n, n, n)
    c = (n - 1) / 2.0
    idx = np.indices((n, n, n)).astype(np.float64)
    r = np.sqrt(((idx - c) ** 2).sum(0))
    m[(r < radius_cells) & (m != fo.SOLID)] = fo.FLUID
    return m


def test_sphere_normals_point_outward_and_curvature_is_about_2_over_r():
    """On a blurred ball the normal grad(chi)/|grad(chi)| points inward (chi
    falls outward), so -div n is positive, near 2/R at the surface."""
    n, rc = 40, 10.0
    dx = 1.0 / n
    m = _sphere_marker(n, rc)
    chi = fo.csf_indicator(m, np.float64)
    for _ in range(2):
        chi = fo.box_blur(chi)
    nrm = fo.csf_normal(chi, dx)
    kap = fo.csf_curvature(nrm, dx)
    c = (n - 1) / 2.0
    idx = np.indices((n, n, n)).astype(np.float64)
    rel = idx - c
    r = np.sqrt((rel ** 2).sum(0))
    shell = (np.abs(r - rc) < 1.0) & (r > 0)
    radial = sum(nrm[a] * rel[a] for a in range(3)) / np.maximum(r, 1e-9)
    assert (radial[shell] < -0.9).mean() > 0.95
    k_est = float(np.median(kap[shell]))
    k_true = 2.0 / (rc * dx)
    assert 0.7 * k_true < k_est < 1.3 * k_true, (k_est, k_true)


def test_force_balance_leaves_zero_mean_change():
    n = 24
    m = _sphere_marker(n, 6.0)
    rng = np.random.default_rng(3)
    shp = [(n + 1, n, n), (n, n + 1, n), (n, n, n + 1)]
    u, v, w = (rng.normal(size=s) for s in shp)
    out = fo.csf_apply(u, v, w, m, 1.0 / n, 0.3, 2)
    for comp, (a, b) in enumerate(zip((u, v, w), out)):
        _, fluid, _, _, _ = fo._face_masks(m, comp)
        d = (b - a)[fluid]
        assert abs(d.mean()) < 1e-12
        assert np.abs(b - a).max() > 1e-3      # the impulse is not trivially zero


def test_apply_matches_per_face_loop():
    n = 10
    m = _sphere_marker(n, 3.0)
    m[2, 2, 5] = fo.SOLID
    dx = 1.0 / n
    rng = np.random.default_rng(4)
    shp = [(n + 1, n, n), (n, n + 1, n), (n, n, n + 1)]
    vel = [rng.normal(size=s).astype(np.float32) for s in shp]
    coeff = np.float32(0.7)
    chi = fo.csf_indicator(m)
    for _ in range(2):
        chi = fo.box_blur(chi)
    kap = fo.csf_curvature(fo.csf_normal(chi, dx), dx)
    out = fo.csf_apply(*vel, m, dx, coeff, 2)
    for comp in range(3):
        g = vel[comp].copy()
        deltas = []
        fl = []
        for idx in np.ndindex(g.shape):
            ax = idx[comp]
            if ax <= 0 or ax >= m.shape[comp]:
                continue
            lo = list(idx)
            lo[comp] -= 1
            lo = tuple(lo)
            if m[lo] == fo.SOLID or m[idx] == fo.SOLID:
                continue
            grad = np.float32((chi[idx] - chi[lo]) / np.float32(dx))
            kf = np.float32(np.float32(0.5) * np.float32(kap[idx] + kap[lo]))
            new = np.float32(g[idx] + np.float32(np.float32(coeff * kf) * grad))
            if m[lo] == fo.FLUID or m[idx] == fo.FLUID:
                deltas.append(np.float32(new - g[idx]))
                fl.append(idx)
            g[idx] = new
        mean = np.float32(np.float32(np.sum(np.array(deltas, np.float64))) / np.float32(len(fl)))
        for idx in fl:
            g[idx] = np.float32(g[idx] - mean)
        np.testing.assert_array_equal(out[comp], g)


def test_surface_tension_substep_pulls_a_cube_toward_a_ball():
    """A cube of fluid at rest without gravity: the CSF impulse points inward
    at the corners, so corner particles gain speed toward the centre."""
    from gpufluid_amd.seeding import fluid_cells, seed_cells
    shape = (16, 16, 16)
    dx = 1.0 / 16
    cells = fluid_cells(shape, dx, lambda x, y, z: (np.abs(x - 0.5) < 0.25) &
                        (np.abs(y - 0.5) < 0.25) & (np.abs(z - 0.5) < 0.25))
    pos = seed_cells(cells, dx, np.random.default_rng(5))
    vel = np.zeros_like(pos)
    st = {"pos": pos, "vel": vel, "marker": fo.wall_marker(*shape), "dx": dx}
    o, _ = fo.substep(st, 1 / 240, pressure="exact", gravity=0.0, surface_tension=0.5,
                      flip_blend=0.0)
    rel = pos - 0.5
    corner = (np.abs(rel) > 0.2).all(1)
    radial = (o["vel"] * rel).sum(1)
    assert corner.any() and (radial[corner] < 0).mean() > 0.9


def test_csf_max_stable_dt_closed_form():
    """S2.14.5 host helper (solver3d.py:2284-2307): safety*sqrt(rho dx^3 /
    (2 pi sigma)); no bound without surface tension."""
    import math
    from gpufluid_amd.kernels import csf_max_stable_dt
    assert csf_max_stable_dt(1.0, 0.1, 0.0) == float("inf")
    got = csf_max_stable_dt(1000.0, 1.0 / 64, 0.0728, safety=0.5)
    assert abs(got - 0.5 * math.sqrt(1000.0 / 64 ** 3 / (2 * math.pi * 0.0728))) < 1e-15
    # doubling sigma shrinks the bound by sqrt(2); dx^(3/2) scaling
    assert abs(csf_max_stable_dt(1.0, 0.02, 0.2) * math.sqrt(2) - csf_max_stable_dt(1.0, 0.02, 0.1)) < 1e-15
    assert abs(csf_max_stable_dt(1.0, 0.04, 0.1) / csf_max_stable_dt(1.0, 0.01, 0.1) - 8.0) < 1e-12
