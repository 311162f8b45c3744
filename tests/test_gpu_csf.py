"""Surface tension (S2.14) on the HIP path against the oracle
(solver3d.py:203-549, 3230-3300, 3548-3555)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import gfutil as gu
from oracle import flip_oracle as fo

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _marker(shape, seed):
    rng = np.random.default_rng(seed)
    m = fo.wall_marker(*shape)
    idx = np.indices(shape).astype(np.float64)
    c = (np.array(shape) - 1) / 2.0
    r = np.sqrt(sum((idx[a] - c[a]) ** 2 for a in range(3)))
    m[(r < min(shape) * 0.3) & (m != fo.SOLID)] = fo.FLUID
    m[(rng.random(shape) < 0.05) & (m == fo.FLUID)] = fo.SOLID
    return m


@pytest.mark.parametrize("shape,passes", [((20, 24, 16), 2), ((33, 17, 26), 1)])
def test_csf_apply_matches_fp32_oracle(shape, passes):
    from gpufluid_amd import _lib
    nx, ny, nz = shape
    m = _marker(shape, nx)
    rng = np.random.default_rng(ny)
    shp = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    vel = [rng.normal(size=s).astype(np.float32) for s in shp]
    dx, coeff = 1.0 / max(shape), 0.05
    print(f"[csf] apply {shape} passes={passes}", flush=True)
    dev = torch.device("cuda:0")
    tv = [torch.tensor(a, device=dev) for a in vel]
    tm = torch.tensor(m, device=dev)
    wf, rd = ctypes.c_int64(), ctypes.c_int64()
    _lib.call("S2.14", "gf_csf_workspace", nx, ny, nz, ctypes.byref(wf), ctypes.byref(rd))
    work = torch.empty(wf.value, dtype=torch.float32, device=dev)
    red = torch.empty(rd.value, dtype=torch.float64, device=dev)
    _lib.call("S2.14", "gf_csf_apply", *(_lib.ptr(t) for t in tv), _lib.ptr(tm), nx, ny, nz,
              dx, coeff, passes, _lib.ptr(work), _lib.ptr(red), _lib.stream_handle(dev))
    ref = fo.csf_apply(*vel, m, dx, np.float32(coeff), passes)
    for c in range(3):
        got = tv[c].cpu().numpy()
        assert gu.rel_err(got, ref[c]) < 1e-6, c
        # faces the balance does not touch match bit for bit
        _, fluid, _, _, _ = fo._face_masks(m, c)
        np.testing.assert_array_equal(got[~fluid], ref[c][~fluid])


@pytest.mark.parametrize("transfer,solver,visc", [("flip", "jacobi", 0.0), ("apic", "pcg", 0.0),
                                                  ("flip", "pcg", 0.01)])
def test_surface_tension_substep_matches_oracle(transfer, solver, visc):
    shape, dx, pos, vel = gu.dam_scene(shape=(20, 24, 16), vel_sigma=0.4, seed=41)
    sigma = 0.07
    print(f"[csf] substep {transfer} {solver} visc={visc}", flush=True)
    s = gu.make_solver(shape, dx, pos, vel, transfer_mode=transfer, surface_tension=sigma,
                       viscosity=visc, viscosity_iters=6)
    dt = 1.0 / 240
    kw = dict(transfer=transfer, surface_tension=sigma, csf_passes=2, viscosity=visc,
              viscosity_iters=6)
    if solver == "jacobi":
        s.step(dt, pressure_iters=30, pressure_solver="jacobi")
        st, grid = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure="jacobi",
                              iters=30, **kw)
    else:
        s.step(dt, pressure_iters=400, pressure_solver="pcg", pressure_tol=1e-8)
        st, grid = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure="exact", **kw)
    assert gu.rel_err(s.div.numpy(), grid["div"]) < TOL
    for name in ("u", "v", "w"):
        assert gu.rel_err(getattr(s, name).numpy(), grid[name]) < TOL, name
    pp, vv = s.get_particles()
    assert gu.rel_err(vv, st["vel"]) < TOL
    assert gu.rel_err(pp, st["pos"]) < TOL


def test_sub_blocks_one_launch_at_a_time():
    """The S2.14 block callables (solver3d.py:212-547) chained as the
    reference's driver launches them equal the oracle's pieces: indicator,
    blur, normal, curvature and the per-component apply bit for bit; the
    force-balance sum (added to the caller's accumulators) to FP32 rounding of
    the exact sum; the bias subtraction bit for bit given that sum."""
    from gpufluid_amd import kernels as K
    shape = (22, 19, 24)
    nx, ny, nz = shape
    m = _marker(shape, 5)
    rng = np.random.default_rng(5)
    dx, coeff = 1.0 / 24, 0.03
    dev = "cuda:0"
    T = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    mt = T(m)
    chi = torch.empty(shape, device=dev)
    tmp = torch.empty(shape, device=dev)
    print(f"[csf] sub-blocks {shape}", flush=True)
    K.k3_indicator_from_marker(mt, chi)
    want_chi = fo.csf_indicator(m, np.float32)
    np.testing.assert_array_equal(chi.cpu().numpy(), want_chi)
    for _ in range(2):
        K.k3_box_blur_centred(chi, tmp)
        chi, tmp = tmp, chi
        want_chi = fo.box_blur(want_chi)
    np.testing.assert_array_equal(chi.cpu().numpy(), want_chi)
    n = [torch.empty(shape, device=dev) for _ in range(3)]
    K.k3_csf_normal(chi, *n, dx)
    want_n = fo.csf_normal(want_chi, dx)
    for a in range(3):
        np.testing.assert_array_equal(n[a].cpu().numpy(), want_n[a])
    kappa = torch.empty(shape, device=dev)
    K.k3_csf_curvature(*n, kappa, dx)
    want_k = fo.csf_curvature(want_n, dx)
    np.testing.assert_array_equal(kappa.cpu().numpy(), want_k)
    shp = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    applies = (K.k3_csf_apply_u, K.k3_csf_apply_v, K.k3_csf_apply_w)
    sums = (K.k3_csf_sum_u, K.k3_csf_sum_v, K.k3_csf_sum_w)
    biases = (K.k3_csf_subtract_bias_u_dev, K.k3_csf_subtract_bias_v_dev,
              K.k3_csf_subtract_bias_w_dev)
    for comp in range(3):
        f0 = rng.normal(size=shp[comp]).astype(np.float32)
        before, vel = T(f0), T(f0)
        applies[comp](vel, chi, kappa, mt, coeff, dx, nx, ny, nz)
        want = fo.csf_face_apply(f0, want_chi, want_k, m, comp, dx, coeff)
        np.testing.assert_array_equal(vel.cpu().numpy(), want)
        acc = T(np.array([0.25], np.float32))
        cnt = T(np.array([1.0], np.float32))
        sums[comp](before, vel, mt, acc, cnt, nx, ny, nz)
        total, count = fo.csf_face_sum(f0, want, m, comp)
        assert count > 100
        assert float(cnt.cpu()[0]) == np.float32(1.0) + np.float32(count)
        np.testing.assert_allclose(float(acc.cpu()[0]), np.float32(0.25) + total,
                                   rtol=2e-7, atol=1e-12)
        # the bias pass with the oracle's own accumulators
        s_t, c_t = T(np.array([total], np.float32)), T(np.array([count], np.float32))
        biases[comp](vel, mt, s_t, c_t, nx, ny, nz)
        np.testing.assert_array_equal(vel.cpu().numpy(),
                                      fo.csf_face_subtract(want, m, comp, total, count))
