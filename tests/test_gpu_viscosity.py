"""Viscosity (S2.13) on the HIP path: the dense block kernel bit for bit
against the FP32 oracle sweep, and whole viscous substeps against the FP64
oracle (solver3d.py:178-197, 3524-3544)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import gfutil as gu
from oracle import flip_oracle as fo

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("shape", [(17, 13, 9), (8, 9, 70), (1, 1, 5)])
def test_visc_block_kernel_bit_exact(shape):
    from gpufluid_amd import kernels
    rng = np.random.default_rng(sum(shape))
    u_old = rng.normal(size=shape).astype(np.float32)
    u_in = rng.normal(size=shape).astype(np.float32)
    r = 0.37
    print(f"[visc] block sweep {shape}", flush=True)
    t_old, t_in = torch.tensor(u_old, device="cuda:0"), torch.tensor(u_in, device="cuda:0")
    t_out = torch.empty_like(t_in)
    kernels.k3_jacobi_visc(t_old, t_in, t_out, r)
    np.testing.assert_array_equal(t_out.cpu().numpy(), fo.visc_sweep(u_old, u_in, r))


@pytest.mark.parametrize("shape", [(20, 24, 16), (70, 5, 9)])
def test_visc3_kernel_bit_exact(shape):
    """The three-component marching sweep of the substep, bit for bit."""
    from gpufluid_amd import _lib
    nx, ny, nz = shape
    rng = np.random.default_rng(nx)
    shp = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    old = [rng.normal(size=x).astype(np.float32) for x in shp]
    inp = [rng.normal(size=x).astype(np.float32) for x in shp]
    r = 0.29
    print(f"[visc] three-component sweep {shape}", flush=True)
    t_old = [torch.tensor(a, device="cuda:0") for a in old]
    t_in = [torch.tensor(a, device="cuda:0") for a in inp]
    t_out = [torch.empty_like(t) for t in t_in]
    _lib.call("S2.13", "gf_k3_jacobi_visc3", *(_lib.ptr(t) for t in t_old + t_in + t_out),
              nx, ny, nz, r, _lib.stream_handle(torch.device("cuda:0")))
    for c in range(3):
        np.testing.assert_array_equal(t_out[c].cpu().numpy(), fo.visc_sweep(old[c], inp[c], r))


@pytest.mark.parametrize("transfer,solver,iters", [("flip", "jacobi", 12), ("flip", "pcg", 5),
                                                   ("apic", "pcg", 12), ("pic", "jacobi", 0)])
def test_viscous_substep_matches_oracle(transfer, solver, iters):
    shape, dx, pos, vel = gu.dam_scene(shape=(20, 24, 16), vel_sigma=0.4, seed=31)
    fb = 0.0 if transfer == "pic" else 0.95
    visc = 0.02
    print(f"[visc] substep {transfer} {solver} iters={iters}", flush=True)
    s = gu.make_solver(shape, dx, pos, vel, transfer_mode=transfer, flip_blend=fb,
                       viscosity=visc, viscosity_iters=iters)
    dt = 1.0 / 240
    kw = dict(transfer=transfer, flip_blend=fb, viscosity=visc, viscosity_iters=iters)
    if solver == "jacobi":
        s.step(dt, pressure_iters=30, pressure_solver="jacobi")
        st, grid = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure="jacobi",
                              iters=30, **kw)
    else:
        s.step(dt, pressure_iters=400, pressure_solver="pcg", pressure_tol=1e-8)
        st, grid = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure="exact", **kw)
    assert gu.rel_err(s.div.numpy(), grid["div"]) < TOL
    assert gu.rel_err(s.p.numpy(), grid["p"]) < TOL
    for name in ("u", "v", "w"):
        assert gu.rel_err(getattr(s, name).numpy(), grid[name]) < TOL, name
    pp, vv = s.get_particles()
    assert gu.rel_err(vv, st["vel"]) < TOL
    assert gu.rel_err(pp, st["pos"]) < TOL


def test_viscous_steps_run_and_adjoint_refuses():
    from gpufluid_amd.adjoint import FlipRollout
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=32)
    s = gu.make_solver(shape, dx, pos, vel, viscosity=0.01, viscosity_iters=7)
    for _ in range(5):
        s.step(1.0 / 240, pressure_iters=80, pressure_solver="pcg")
    assert not s.has_diverged()
    with pytest.raises(NotImplementedError):
        FlipRollout(s, 1.0 / 240, 2)
