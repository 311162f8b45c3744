"""The reference's own solver-level scenarios, run on this build's
FlipSolver3D with the reference's parameters: grid shapes and a 50-substep
run without NaN inside the box (tests/test_f3_solver.py:18-40), gravity
lowering the centroid (42-51), PCG and GS-RB against many-sweep Jacobi
(tests/test_s2_6_3_pcg.py:23-37, tests/test_s2_6_2_gsrb.py:9-21) and the PCG
finishing well inside its budget at 16^3 (tests/test_s2_6_3_pcg.py:40-47).
These are the reference's self-consistency checks; the oracle tests pin the
arithmetic."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(**kw):
    from gpufluid_amd import FlipSolver3D
    return FlipSolver3D(nx=16, ny=16, nz=16, dx=1.0 / 16, device="cuda:0", **kw)


def test_grid_shapes():
    s = _solver()
    assert s.u.shape == (17, 16, 16) and s.v.shape == (16, 17, 16) and s.w.shape == (16, 16, 17)
    assert s.marker.shape == (16, 16, 16)


def test_fifty_substeps_stay_finite_and_inside():
    s = _solver()
    s.seed_box((0.10, 0.10, 0.10), (0.50, 0.90, 0.50), ppc=8)
    n0 = s.n_particles
    for _ in range(50):
        s.step(0.005, pressure_iters=20)
    pos, vel = s.get_particles()
    assert s.n_particles == n0
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    for a in range(3):
        assert pos[:, a].min() >= 0.0 and pos[:, a].max() <= s.dom[a]
    assert not s.has_diverged()


def test_gravity_lowers_the_centroid():
    s = _solver(flip_blend=0.95)
    s.seed_box((0.20, 0.55, 0.20), (0.80, 0.90, 0.80), ppc=8)
    y0 = s.get_particles()[0][:, 1].mean()
    for _ in range(60):
        s.step(0.005, pressure_iters=30)
    y1 = s.get_particles()[0][:, 1].mean()
    assert y1 < y0 - 0.05, (y0, y1)


def _centroid(solver, iters, hi=(0.80, 0.90, 0.80)):
    s = _solver()
    s.seed_box((0.20, 0.30, 0.20), hi, ppc=8)
    for _ in range(30):
        s.step(0.005, pressure_iters=iters, pressure_solver=solver)
    return s.get_particles()[0].mean(axis=0)


def test_pcg_and_jacobi_agree():
    cj = _centroid("jacobi", 200)
    cp = _centroid("pcg", 40)
    assert np.linalg.norm(cj - cp) < 0.05, (cj, cp)


def test_gsrb_and_jacobi_agree():
    hi = (0.50, 0.70, 0.50)
    cj = _centroid("jacobi", 200, hi)
    cg = _centroid("gsrb", 100, hi)
    assert np.linalg.norm(cj - cg) < 0.05, (cj, cg)


def test_pcg_finishes_inside_its_budget():
    s = _solver()
    s.seed_box((0.20, 0.30, 0.20), (0.80, 0.90, 0.80), ppc=8)
    for _ in range(2):
        s.step(0.005, pressure_iters=40, pressure_solver="pcg")
    assert s.last_pressure_iters < 40, s.last_pressure_iters
