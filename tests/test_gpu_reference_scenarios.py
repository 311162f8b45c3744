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


@pytest.mark.parametrize("n,scale,seed", [(1, 1.0, 0), (1000, 0.5, 1), (50_000, 5.0, 2),
                                          (200_000, 12.0, 3)])
def test_gpu_cfl_count_equals_host_count(n, scale, seed):
    """tests/test_s2_10_gpu_cfl.py:24-40 of the reference."""
    import torch
    from gpufluid_amd.solver3d import cfl_substep_count, cfl_substep_count_gpu
    vel = (np.random.default_rng(seed).standard_normal((n, 3)) * scale).astype(np.float32)
    got = cfl_substep_count_gpu(torch.tensor(vel, device="cuda:0"), dx=0.0125, target_dt=0.04)
    assert got == cfl_substep_count(vel, dx=0.0125, target_dt=0.04)


def test_gpu_cfl_empty_and_still():
    import torch
    from gpufluid_amd.solver3d import cfl_substep_count_gpu
    assert cfl_substep_count_gpu(torch.zeros((0, 3), device="cuda:0"), dx=0.0125,
                                 target_dt=0.04) == 1
    assert cfl_substep_count_gpu(torch.zeros((100, 3), device="cuda:0"), dx=0.0125,
                                 target_dt=0.04) == 1


def _drop_on_sphere(mode, frames):
    """A fluid column dropped on a solid sphere (tests/test_s2_12_apic_obstacle.py:20-37)."""
    from gpufluid_amd import FlipSolver3D
    from gpufluid_amd.animation import sdf_sphere
    s = FlipSolver3D(nx=48, ny=48, nz=48, dx=1.0 / 48.0, gravity=-9.81, transfer_mode=mode,
                     flip_blend=0.95, device="cuda:0")
    s.add_solid_from_sdf(sdf_sphere(s.cell_centers_np(), center=(0.5, 0.35, 0.5), radius=0.10))
    s.seed_box((0.40, 0.65, 0.40), (0.60, 0.90, 0.60))
    ke, vmax = [], []
    for _ in range(frames):
        s.step(0.005, pressure_iters=40)
        sp = np.linalg.norm(s.vel.numpy(), axis=1)
        vmax.append(float(sp.max()))
        ke.append(float(0.5 * (sp ** 2).sum()))
    return np.asarray(ke), np.asarray(vmax), s.vel.numpy()


def test_apic_near_an_obstacle_is_not_more_energetic_than_flip():
    """tests/test_s2_12_apic_obstacle.py:39-70 of the reference: no NaN,
    APIC kinetic energy below 1.2x FLIP's over the impact window, max speed
    below 2x FLIP's."""
    ke_f, vm_f, _ = _drop_on_sphere("flip", 40)
    ke_a, vm_a, vel = _drop_on_sphere("apic", 40)
    assert np.isfinite(vel).all()
    assert (ke_a[20:40] / np.maximum(ke_f[20:40], 1.0)).max() < 1.2
    assert (vm_a / np.maximum(vm_f, 1e-3)).max() < 2.0


def _reseed_input(n=80_000, N=32, seed=0):
    rng = np.random.default_rng(seed)
    dx = 1.0 / N
    pos = (rng.random((n, 3)) * (N * dx * 0.8) + N * dx * 0.1).astype(np.float32)
    vel = (rng.standard_normal((n, 3)) * 0.1).astype(np.float32)
    marker = np.zeros((N, N, N), np.int32)
    for a in range(3):
        idx = [slice(None)] * 3
        idx[a] = [0, -1]
        marker[tuple(idx)] = 2
    return pos, vel, marker, dx


def test_device_and_host_reseed_agree_on_counts():
    """tests/test_s2_11_gpu_reseed.py:41-56 of the reference: same emitted and
    culled counts from both forms; no cell keeps more than max_per_cell."""
    import torch
    from gpufluid_amd.reseed import ReseedConfig, reseed_particles, reseed_particles_gpu
    pos, vel, marker, dx = _reseed_input()
    cfg = ReseedConfig(min_per_cell=4, max_per_cell=16)
    dev = "cuda:0"
    p_g, _, _, _, ne_g, nc_g = reseed_particles_gpu(torch.tensor(pos, device=dev),
                                                    torch.tensor(vel, device=dev),
                                                    torch.tensor(marker, device=dev), dx, cfg,
                                                    np.random.default_rng(0))
    _, _, ne_c, nc_c = reseed_particles(pos, vel, marker, dx, cfg, np.random.default_rng(0))
    assert (ne_g, nc_g) == (ne_c, nc_c)
    assert ne_g > 0


def test_device_reseed_bounds_the_cell_counts():
    """tests/test_s2_11_gpu_reseed.py:59-78 of the reference."""
    import torch
    from gpufluid_amd.reseed import ReseedConfig, reseed_particles_gpu
    pos, vel, marker, dx = _reseed_input(n=120_000)
    cfg = ReseedConfig(min_per_cell=4, max_per_cell=12)
    dev = "cuda:0"
    p_g, _, _, _, _, n_cull = reseed_particles_gpu(torch.tensor(pos, device=dev),
                                                   torch.tensor(vel, device=dev),
                                                   torch.tensor(marker, device=dev), dx, cfg,
                                                   np.random.default_rng(0))
    q = np.clip((p_g.numpy() / dx).astype(np.int32), 0, 31)
    counts = np.bincount((q[:, 0] * 32 + q[:, 1]) * 32 + q[:, 2],
                         minlength=32 ** 3).reshape(32, 32, 32)
    assert n_cull > 0
    assert int(counts[marker != 2].max()) <= cfg.max_per_cell
