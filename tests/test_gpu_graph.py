"""Substeps replayed from captured hipGraphs (``enable_cuda_graphs=True``, the
graph cache of solver3d.py:3361-3469) against the same substeps launched
directly: bit-identical particles, pressure and statistics while the fluid
moves and the set of active tiles changes between substeps (the pressure
solvers read the tile count on the device, GF_NACT_DEVICE)."""
from __future__ import annotations

import numpy as np
import pytest

import gfutil as gu

pytestmark = pytest.mark.gpu

DT = 0.01


def _scene(seed=31, n=24):
    # the dam's +x face sits on a tile boundary (cell 16 at n = 24), so the
    # set of active tiles changes within the first substeps
    return gu.dam_scene(n=n, x_max=0.66, vel_sigma=1.5, seed=seed)


def _pair(shape, dx, pos, vel, **kw):
    a = gu.make_solver(shape, dx, pos, vel, **kw)
    g = gu.make_solver(shape, dx, pos, vel, enable_cuda_graphs=True, **kw)
    return a, g


def _assert_same(a, g):
    pa, va = a.get_particles()
    pg, vg = g.get_particles()
    np.testing.assert_array_equal(pg, pa)
    np.testing.assert_array_equal(vg, va)
    np.testing.assert_array_equal(g.p.numpy(), a.p.numpy())
    np.testing.assert_array_equal(g.u.numpy(), a.u.numpy())
    assert g.n_active_tiles == a.n_active_tiles
    assert g._sc_fluid_cells == a._sc_fluid_cells


@pytest.mark.parametrize("transfer", ["flip", "apic"])
@pytest.mark.parametrize("solver", ["jacobi", "gsrb", "pcg"])
def test_replay_is_bit_identical_to_direct(solver, transfer):
    shape, dx, pos, vel = _scene()
    print(f"[graph] {solver} {transfer} {shape} n={len(pos)}", flush=True)
    a, g = _pair(shape, dx, pos, vel, transfer_mode=transfer)
    kw = dict(pressure_iters=200 if solver == "pcg" else 40, pressure_solver=solver,
              pressure_tol=1e-6)
    tiles = []
    for k in range(6):
        a.step(DT, **kw)
        g.step(DT, **kw)
        tiles.append(a.n_active_tiles)
        assert g.last_pressure_iters == a.last_pressure_iters
        assert g.last_pressure_residual == a.last_pressure_residual
        assert g.last_pressure_replacements == a.last_pressure_replacements
        _assert_same(a, g)
    assert len(set(tiles)) > 1, tiles          # the active tiles changed on the way
    # one configuration: both ping-pong states captured at the first substep
    assert (g._cuda_graph_misses, g._cuda_graph_hits) == (1, 5)
    assert len(g._graphs) == 2
    assert a._cuda_graph_misses == 0 and a._cuda_graph is None
    if solver == "pcg":
        assert 0 < a.last_pressure_iters < 200


def test_replay_with_every_stage():
    """Viscosity with an odd sweep count (the face buffers end swapped, so the
    two captured states differ in more than the particle buffers), surface
    tension and both per-particle attributes."""
    shape, dx, pos, vel = _scene(seed=32)
    rng = np.random.default_rng(32)
    col = rng.random((len(pos), 3)).astype(np.float32)
    temp = (300.0 + 50.0 * rng.random(len(pos))).astype(np.float32)
    kw = dict(viscosity=0.05, viscosity_iters=3, surface_tension=0.02, transfer_mode="apic")
    a, g = _pair(shape, dx, pos, vel, **kw)
    for s in (a, g):
        s.set_particles(pos, vel, color=col, temperature=temp)
    print(f"[graph] all stages {shape} n={len(pos)}", flush=True)
    for _ in range(4):
        for s in (a, g):
            s.step(DT, pressure_iters=200, pressure_solver="pcg", pressure_tol=1e-6)
        _assert_same(a, g)
        np.testing.assert_array_equal(g.attr_color.numpy(), a.attr_color.numpy())
        np.testing.assert_array_equal(g.attr_temperature.numpy(), a.attr_temperature.numpy())
        np.testing.assert_array_equal(g.us.numpy(), a.us.numpy())
    assert (g._cuda_graph_misses, g._cuda_graph_hits) == (1, 3)


def test_key_changes_recapture():
    """A new dt, a new particle set (as prepare_frame's emission makes) and a
    new iteration cap each capture afresh; results stay identical to the
    direct launches throughout."""
    shape, dx, pos, vel = _scene(seed=33)
    a, g = _pair(shape, dx, pos, vel)
    extra = (np.array([0.7, 0.7, 0.5], np.float32)
             + 0.05 * np.random.default_rng(33).random((300, 3)).astype(np.float32))
    # each configuration is one miss (both ping-pong states captured together)
    plan = ([(DT, 100)] * 3 + [(DT / 2, 100)] * 2 + ["grow"] + [(DT / 2, 100)] * 3
            + [(DT / 2, 60)])
    print(f"[graph] key changes {shape} n={len(pos)}", flush=True)
    for item in plan:
        if item == "grow":
            for s in (a, g):
                p, v = s.get_particles()
                s.set_particles(np.concatenate([p, extra]),
                                np.concatenate([v, np.zeros_like(extra)]))
            assert not g._graphs                 # the particle set invalidates the cache
            continue
        dt, iters = item
        for s in (a, g):
            s.step(dt, pressure_iters=iters, pressure_solver="pcg", pressure_tol=1e-6)
        assert g.last_pressure_iters == a.last_pressure_iters
        _assert_same(a, g)
    assert (g._cuda_graph_misses, g._cuda_graph_hits) == (4, 5)
    assert len(g._graphs) <= g._GRAPH_CACHE


def test_ineligible_configurations_run_direct():
    shape, dx, pos, vel = _scene(seed=34, n=16)
    g = gu.make_solver(shape, dx, pos, vel, enable_cuda_graphs=True)
    g.pcg_kernel_timing = True
    g.step(DT, pressure_iters=100, pressure_solver="pcg")
    assert g._cuda_graph_misses == 0 and g.last_pcg_kernel_ms is not None
    g.pcg_kernel_timing = False
    g.step(DT, pressure_iters=100, pressure_solver="pcg")
    assert g._cuda_graph_misses == 1


def test_cache_counters_follow_the_reference():
    """tests/test_b5_2_graph_cache.py:33-74 of the reference: off by default;
    the first substep misses and the next four hit; prepare_frame drops the
    cache; changing the iteration cap recaptures once."""
    from gpufluid_amd import FlipSolver3D
    s = FlipSolver3D(16, 16, 16, dx=1 / 16, device="cuda:0")
    s.seed_box((0.2, 0.2, 0.2), (0.6, 0.6, 0.6), ppc=8)
    s.step(0.005, pressure_iters=20)
    assert s._cuda_graph is None and (s._cuda_graph_hits, s._cuda_graph_misses) == (0, 0)
    s = FlipSolver3D(16, 16, 16, dx=1 / 16, device="cuda:0", enable_cuda_graphs=True)
    s.seed_box((0.2, 0.2, 0.2), (0.6, 0.6, 0.6), ppc=8)
    s.step(0.005, pressure_iters=20)
    assert s._cuda_graph is not None and (s._cuda_graph_misses, s._cuda_graph_hits) == (1, 0)
    for _ in range(4):
        s.step(0.005, pressure_iters=20)
    assert (s._cuda_graph_misses, s._cuda_graph_hits) == (1, 4)
    s.prepare_frame(0, 1 / 24)
    assert s._cuda_graph is None
    s = FlipSolver3D(16, 16, 16, dx=1 / 16, device="cuda:0", enable_cuda_graphs=True)
    s.seed_box((0.2, 0.2, 0.2), (0.6, 0.6, 0.6), ppc=8)
    for iters in (10, 20, 20):
        s.step(0.005, pressure_iters=iters)
    assert (s._cuda_graph_misses, s._cuda_graph_hits) == (2, 1)
