"""Shared helpers for the parity tests (not a test module)."""
from __future__ import annotations

import ctypes

import numpy as np

from oracle import flip_oracle as fo


def rel_err(a, b) -> float:
    """max|a-b| / max|b| (the reference tests' pair-parity measure,
    tests/test_b5_a_pcg_graph.py:75-78), with a floor on the scale."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.abs(b).max()) if b.size else 0.0, 1e-12)
    return float(np.abs(a - b).max()) / scale if a.size else 0.0


def dam_scene(n=16, seed=0, x_max=0.4, y_max=0.6, vel_sigma=0.0, shape=None):
    """Seeded dam-break particles for an (n,n,n) or given grid with dx = 1/max(shape)."""
    from gpufluid_amd.seeding import fluid_cells, seed_cells
    shape = shape or (n, n, n)
    dx = 1.0 / max(shape)
    cells = fluid_cells(shape, dx, lambda x, y, z: (x < x_max) & (y < y_max))
    rng = np.random.default_rng(seed)
    pos = seed_cells(cells, dx, rng)
    vel = (rng.normal(0.0, vel_sigma, pos.shape) if vel_sigma else np.zeros_like(pos))
    return shape, dx, pos, vel.astype(np.float32)


def make_solver(shape, dx, pos, vel, **kw):
    from gpufluid_amd import FlipSolver3D
    s = FlipSolver3D(*shape, dx=dx, **kw)
    s.set_particles(pos, vel)
    return s


def run_sort(s):
    """Run only the binning/sort stage on a solver; returns n_active."""
    from gpufluid_amd import _lib
    sc = s._sc
    apic = s.transfer_mode == "apic"
    sc.ensure_particles(s.n_particles, apic)
    d = s._descriptor()
    na = (ctypes.c_int * 2)()
    _lib.call("F3.3", "gf_scene_sort_sync", ctypes.byref(d), int(apic), na,
              s._stream())
    if s.n_particles:
        sc.swap()
    sc.n_active_host = na[0]
    s._cache.clear()
    return na[0]


def run_p2g(s, dt, gravity=None, staged=False):
    """The fused P2G alone: one-pass (gf_scene_p2g) or staged (the substep's)."""
    from gpufluid_amd import _lib
    d = s._descriptor()
    mode = _lib.GF_TRANSFER_APIC if s.transfer_mode == "apic" else _lib.GF_TRANSFER_FLIP
    g = s.gravity if gravity is None else gravity
    if staged:
        stage, has = s._p2g_stage()
        _lib.call("S2.1", "gf_scene_p2g_staged", ctypes.byref(d), mode, g, dt, s.rho,
                  _lib.ptr(stage), _lib.ptr(has), s._stream())
    else:
        _lib.call("S2.1", "gf_scene_p2g", ctypes.byref(d), mode, g, dt, s.rho, s._stream())
    s._cache.clear()


def oracle_state(shape, dx, pos, vel, marker=None):
    m = fo.wall_marker(*shape) if marker is None else marker.copy()
    return {"pos": pos.copy(), "vel": vel.copy(), "marker": m, "dx": dx}


def compact_to_dense_host(s, vec):
    """A compact per-tile vector (512 entries per active slot, cell
    (li*8+lj)*8+lk) as a dense (nx,ny,nz) numpy array, on the host, from the
    solver's tile map; cells of inactive tiles are 0."""
    vec = vec.detach().cpu().numpy() if hasattr(vec, "detach") else np.asarray(vec)
    nx, ny, nz = s.nx, s.ny, s.nz
    tnx, tny, tnz = -(-nx // 8), -(-ny // 8), -(-nz // 8)
    tmap = s._sc.tile_map.cpu().numpy().reshape(tnx, tny, tnz)
    full = np.zeros((tnx * 8, tny * 8, tnz * 8), dtype=vec.dtype)
    ti, tj, tk = np.nonzero(tmap >= 0)
    blocks = vec[: s._sc.compact].reshape(-1, 8, 8, 8)[tmap[ti, tj, tk]]
    for t in range(len(ti)):
        full[ti[t] * 8:ti[t] * 8 + 8, tj[t] * 8:tj[t] * 8 + 8, tk[t] * 8:tk[t] * 8 + 8] = blocks[t]
    return full[:nx, :ny, :nz]
