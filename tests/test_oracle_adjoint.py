"""Pins the FP64 torch restatement used as the adjoint's reference gradient
(oracle/torch_rollout.py), on the CPU:

* its forward substep equals the numpy oracle's substep (exact pressure solve);
* its autograd gradient equals central finite differences of itself;
* its pressure Function's backward is the transpose solve (symmetric A).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from gpufluid_amd import seeding
from oracle import flip_oracle as fo
from oracle import torch_rollout as tr


def _scene(n=12, seed=5, sigma=0.5):
    dx = 1.0 / n
    cells = seeding.fluid_cells((n, n, n), dx, lambda x, y, z: (x < 0.5) & (y < 0.6))
    rng = np.random.default_rng(seed)
    pos = seeding.seed_cells(cells, dx, rng)
    vel = rng.normal(0.0, sigma, size=pos.shape).astype(np.float32)
    return fo.wall_marker(n, n, n), dx, pos, vel


def test_forward_matches_numpy_oracle():
    walls, dx, pos, vel = _scene()
    dt = 1.0 / 240
    st = dict(pos=pos, vel=vel, marker=walls, dx=dx)
    out, _ = fo.substep(st, dt, pressure="exact", dtype=np.float64)
    x = torch.from_numpy(pos.astype(np.float64))
    v = torch.from_numpy(vel.astype(np.float64))
    x1, v1 = tr.substep(x, v, walls, dx, dt)
    np.testing.assert_allclose(x1.numpy(), out["pos"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(v1.numpy(), out["vel"], rtol=0, atol=1e-7 * np.abs(out["vel"]).max())


def test_poisson_function_backward_is_transpose_solve():
    walls, dx, pos, _ = _scene(n=10)
    m = fo.update_marker(walls, pos, dx)
    A, cells, _ = fo.poisson_matrix(m)
    rng = np.random.default_rng(1)
    b = torch.from_numpy(rng.normal(size=len(cells))).requires_grad_()
    g = torch.from_numpy(rng.normal(size=len(cells)))
    p = tr._Poisson.apply(b, A, -1.0)
    # p = -A^-1 b  =>  A p = -b
    np.testing.assert_allclose(A @ p.detach().numpy(), -b.detach().numpy(), atol=1e-9)
    (p * g).sum().backward()
    # dL/db = -A^-T g = -A^-1 g
    np.testing.assert_allclose(A @ b.grad.numpy(), -g.numpy(), atol=1e-9)


@pytest.mark.parametrize("substeps", [1, 3])
def test_autograd_gradient_matches_finite_differences(substeps):
    walls, dx, pos, vel = _scene(n=12, seed=7)
    dt = 1.0 / 240
    target = (0.7, 0.3, 0.5)
    x0 = torch.from_numpy(pos.astype(np.float64))
    v0 = torch.from_numpy(vel.astype(np.float64)).requires_grad_()
    x, _ = tr.rollout(x0, v0, walls, dx, dt, substeps)
    loss = tr.com_loss(x, target)
    loss.backward()
    g = v0.grad.numpy()
    rng = np.random.default_rng(11)
    eps = 1e-3
    for _ in range(4):
        d = rng.normal(size=vel.shape)
        d /= np.linalg.norm(d)
        with torch.no_grad():
            lp = tr.com_loss(tr.rollout(x0, torch.from_numpy(vel + eps * d), walls, dx, dt,
                                        substeps, checkpoint=False)[0], target)
            lm = tr.com_loss(tr.rollout(x0, torch.from_numpy(vel - eps * d), walls, dx, dt,
                                        substeps, checkpoint=False)[0], target)
        fd = float(lp - lm) / (2 * eps)
        an = float((g * d).sum())
        assert abs(fd - an) <= 1e-5 * abs(fd), (fd, an)
