"""The reference's own assertions on the solver surface, restated against this
build: the step profiler's sections (tests/test_g1_9_step_profiler.py:13-96),
the PCG stop-flag surface (tests/test_b5_a_pcg_graph.py:34-97), the solver's
ValueError for an unknown pressure solver (solver3d.py:3634) and the
checkpoint's handling of files without attribute keys (solver3d.py:4200-4214).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solver(n=16, **kw):
    from gpufluid_amd import FlipSolver3D
    return FlipSolver3D(nx=n, ny=n, nz=n, dx=1.0 / n, **kw)


def _seed_one_step(enable_timing):
    sol = _solver(enable_timing=enable_timing)
    sol.seed_box((0.2, 0.2, 0.2), (0.6, 0.6, 0.6), ppc=2)
    sol.step(dt=0.01)
    return sol._prof.summarize(), sol._prof.call_counts


def test_disabled_profiler_is_a_noop():
    totals, counts = _seed_one_step(enable_timing=False)
    assert totals == {} and counts == {}


def test_enabled_profiler_collects_expected_sections():
    totals, counts = _seed_one_step(enable_timing=True)
    expected = {"clear", "p2g", "normalize", "gravity_bc", "divergence",
                "pressure", "grad_subtract_bc", "g2p_advect", "color"}
    assert expected.issubset(totals.keys()), (expected - totals.keys(), sorted(totals))
    assert "viscosity" not in totals
    assert "surface_tension" not in totals
    for name, ms in totals.items():
        assert ms >= 0.0, name
    for name in expected:
        assert counts[name] == 1, (name, counts[name])


def test_viscosity_and_surface_tension_sections_fire_when_active():
    sol = _solver(viscosity=0.05, viscosity_iters=2, enable_timing=True)
    sol.seed_box((0.3, 0.3, 0.3), (0.5, 0.5, 0.5), ppc=2)
    sol.step(dt=0.01)
    assert sol._prof.summarize()["viscosity"] >= 0.0
    sol = _solver(surface_tension=0.05, enable_timing=True)
    sol.seed_box((0.3, 0.3, 0.3), (0.5, 0.5, 0.5), ppc=2)
    sol.step(dt=0.005)
    assert "surface_tension" in sol._prof.summarize()


def test_timings_with_cuda_graphs_capture():
    sol = _solver(enable_timing=True, enable_cuda_graphs=True)
    sol.seed_box((0.2, 0.2, 0.2), (0.6, 0.6, 0.6), ppc=2)
    sol.step(dt=0.01)
    sol.step(dt=0.01)
    assert sol._cuda_graph is not None
    assert sol._cuda_graph_misses == 1 and sol._cuda_graph_hits == 1


def _seeded(N=32, **kw):
    s = _solver(N, gravity=-9.81, **kw)
    s.seed_box((0.30, 0.30, 0.30), (0.60, 0.60, 0.60), ppc=4)
    return s


def test_pcg_captures_on_first_step_and_replays():
    s = _seeded(enable_cuda_graphs=True)
    s.step(0.005, pressure_iters=10, pressure_solver="pcg")
    assert s._cuda_graph is not None
    assert s._cuda_graph_misses == 1 and s._cuda_graph_hits == 0
    s.step(0.005, pressure_iters=10, pressure_solver="pcg")
    assert s._cuda_graph_hits == 1


def test_pcg_graph_matches_direct_path():
    s_direct = _seeded(enable_cuda_graphs=False)
    s_graph = _seeded(enable_cuda_graphs=True)
    for _ in range(4):
        s_direct.step(0.005, pressure_iters=30, pressure_solver="pcg")
        s_graph.step(0.005, pressure_iters=30, pressure_solver="pcg")
    assert s_graph._cuda_graph_hits == 3
    u_d, u_g = s_direct.u.numpy(), s_graph.u.numpy()
    assert np.abs(u_d - u_g).max() / max(np.abs(u_d).max(), 1e-3) < 1e-4


def _flags(s):
    return int(s._pcg_done.numpy()[0]), float(s._pcg_tol_sq.numpy()[0])


def _dam(**kw):
    """A block resting on the floor wall: its first substep has a nonzero
    divergence (a box in free fall has none and the solve does nothing)."""
    s = _solver(32, gravity=-9.81, **kw)
    s.seed_box((0.05, 0.05, 0.05), (0.5, 0.6, 0.5), ppc=4)
    return s


def test_pcg_done_flag_surface_and_meaning():
    s = _seeded()
    s.step(0.005, pressure_iters=5, pressure_solver="pcg")
    assert s._pcg_done.shape == (1,) and s._pcg_done.dtype == np.int32
    assert s._pcg_tol_sq.shape == (1,) and s._pcg_tol_sq.dtype == np.float32
    s = _dam()
    s.step(0.005, pressure_iters=5, pressure_solver="pcg")
    # 5 iterations do not reach the default 1e-4: the flag stays 0
    done, tol_sq = _flags(s)
    assert s.last_pressure_iters == 5 and done == 0
    # tol_sq = tol^2 * ||r0||^2 with r0 = b = the divergence on fluid cells
    div = s.div.numpy().astype(np.float64)
    bb = float((div[s.marker.numpy() == 1] ** 2).sum())
    assert bb > 0
    assert tol_sq == pytest.approx(np.float32(1e-4) ** 2 * bb, rel=1e-5)
    # a solve that meets its tolerance raises the flag, in both paths
    for graphs in (False, True):
        s = _dam(enable_cuda_graphs=graphs)
        for _ in range(2):
            s.step(0.005, pressure_iters=200, pressure_solver="pcg")
            done, tol_sq = _flags(s)
            assert s.last_pressure_iters < 200
            assert done == 1 and tol_sq > 0.0


def test_pcg_done_is_zero_without_fluid_to_solve():
    s = _dam()
    s.step(0.005, pressure_iters=200, pressure_solver="pcg")
    assert _flags(s)[0] == 1
    s.set_particles(np.zeros((0, 3), np.float32))
    s.step(0.005, pressure_iters=200, pressure_solver="pcg")
    assert _flags(s) == (0, 0.0)


def test_unknown_pressure_solver_is_a_value_error():
    s = _seeded()
    with pytest.raises(ValueError, match="unknown pressure_solver"):
        s.step(0.005, pressure_solver="multigrid")


def test_checkpoint_without_attribute_keys_keeps_live_attributes(tmp_path):
    s = _seeded()
    n = s.n_particles
    col = np.random.default_rng(0).random((n, 3)).astype(np.float32)
    s.attr_color = col
    s.attr_temperature = np.full(n, 25.0, np.float32)
    s.step(0.005)
    full = tmp_path / "full.npz"
    s.save_checkpoint(full)
    data = dict(np.load(full, allow_pickle=False))
    del data["attr_color"], data["attr_temperature"]
    old = tmp_path / "old.npz"
    np.savez(old, **data)
    live_col = s.attr_color.numpy().copy()
    live_t = s.attr_temperature.numpy().copy()
    s.load_checkpoint(old)              # same particle count: kept
    np.testing.assert_array_equal(s.attr_color.numpy(), live_col)
    np.testing.assert_array_equal(s.attr_temperature.numpy(), live_t)
    data["pos"], data["vel"] = data["pos"][:10], data["vel"][:10]
    fewer = tmp_path / "fewer.npz"
    np.savez(fewer, **data)
    s.load_checkpoint(fewer)            # counts differ: dropped
    assert s.n_particles == 10
    assert s.attr_color is None and s.attr_temperature is None
    torch.cuda.synchronize()
