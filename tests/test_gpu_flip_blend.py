"""The FLIP/PIC blend on the device against the reference formula written out
by hand (solver3d.py:1414-1463): with post-projection grid velocity u and
pre-projection us, v_new = fb*(v_p + u(x_p) - us(x_p)) + (1-fb)*u(x_p),
x_new = x_p + v_new*dt.

* the S2.8 block kernel (gf_k3_g2p_and_advect) on constant and linear fields;
* the fused substep kernel (gf_scene_g2p): us linear, a linear pressure on an
  all-fluid interior, no gravity, so every face between two fluid cells gets
  u = us - dt/(rho dx) * grad p, a constant shift c per component, and the
  particles whose stencils hold only such faces must get
  v_new = fb*(v_p + c) + (1-fb)*(us(x_p) + c).

Tolerance 1e-5 relative (max-abs convention), FP32 arithmetic.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import gfutil as gu

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _linear_faces(coef, shape, dx32):
    nx, ny, nz = shape
    out = []
    for c, (sh, off) in enumerate((((nx + 1, ny, nz), (0.0, 0.5, 0.5)),
                                   ((nx, ny + 1, nz), (0.5, 0.0, 0.5)),
                                   ((nx, ny, nz + 1), (0.5, 0.5, 0.0)))):
        i, j, k = np.meshgrid(*(np.arange(s, dtype=np.float64) for s in sh), indexing="ij")
        x = np.stack([(i + off[0]) * dx32, (j + off[1]) * dx32, (k + off[2]) * dx32])
        out.append(coef[c, 0] + np.tensordot(coef[c, 1:], x, axes=1))
    return out


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


@pytest.mark.parametrize("fb", [0.5, 0.95, 1.0])
def test_block_kernel_flip_blend_closed_form(fb):
    from gpufluid_amd import kernels
    rng = np.random.default_rng(31)
    shape, dx, dt = (10, 12, 9), 1 / 12, 1e-3
    dx32 = float(np.float32(dx))
    dom = np.array([np.float32(n * dx) for n in shape], np.float32)
    n = 500
    lo, hi = 1.5 * dx, (np.array(shape) - 1.5) * dx
    pos = (lo + rng.random((n, 3)) * (hi - lo)).astype(np.float32)
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    x = pos.astype(np.float64)
    for kind in ("constant", "linear"):
        if kind == "constant":
            a, c = np.array([0.7, -1.3, 2.1]), np.array([-0.4, 0.9, 0.25])
            cu = np.zeros((3, 4)); cu[:, 0] = a
            cs = np.zeros((3, 4)); cs[:, 0] = a - c
        else:
            cu, cs = rng.normal(size=(3, 4)), rng.normal(size=(3, 4))
        f = [_dev(t) for t in _linear_faces(cu, shape, dx32) + _linear_faces(cs, shape, dx32)]
        p, v = _dev(pos), _dev(vel)
        kernels.k3_g2p_and_advect(p, v, *f, dx, dt, *shape, fb, dom)
        torch.cuda.synchronize()
        pic = cu[:, 0] + x @ cu[:, 1:].T
        old = cs[:, 0] + x @ cs[:, 1:].T
        want = fb * (vel.astype(np.float64) + pic - old) + (1 - fb) * pic
        assert gu.rel_err(v.cpu().numpy(), want) < TOL, kind
        assert gu.rel_err(p.cpu().numpy(), x + want * dt) < TOL, kind


def _compact_index(s, ii, jj, kk):
    sc = s._sc
    tny, tnz = -(-s.ny // 8), -(-s.nz // 8)
    tmap = sc.tile_map.cpu().numpy()
    slot = tmap[((ii // 8) * tny + jj // 8) * tnz + kk // 8]
    assert (slot >= 0).all()
    return slot * 512 + ((ii % 8) * 8 + jj % 8) * 8 + kk % 8


@pytest.mark.parametrize("fb", [0.5, 0.95, 1.0])
def test_fused_g2p_flip_blend_closed_form(fb):
    from gpufluid_amd import _lib
    rng = np.random.default_rng(32)
    N = 16
    shape, dx, dt, rho = (N, N, N), 1.0 / N, 1e-3, 1.0
    dx32 = float(np.float32(dx))
    # two particles in every interior cell: the whole interior is fluid
    cells = np.stack(np.meshgrid(*(np.arange(1, N - 1),) * 3, indexing="ij"), -1).reshape(-1, 3)
    cells = np.repeat(cells, 2, axis=0)
    pos = ((cells + 0.05 + 0.9 * rng.random(cells.shape)) * dx).astype(np.float32)
    vel = rng.normal(size=pos.shape).astype(np.float32)
    s = gu.make_solver(shape, dx, pos, vel, flip_blend=fb)
    gu.run_sort(s)
    marker = s.marker.numpy()
    assert (marker[1:-1, 1:-1, 1:-1] == 1).all()
    # pre-projection faces: linear fields; pressure: linear on the fluid cells
    cs = rng.normal(size=(3, 4))
    for key, t in zip(("us", "vs", "ws"), _linear_faces(cs, shape, dx32)):
        s._f[key].copy_(_dev(t))
    g = np.array([30.0, -20.0, 12.5])
    ii, jj, kk = np.nonzero(marker == 1)
    pc = np.zeros(s._sc.compact, np.float32)
    pc[_compact_index(s, ii, jj, kk)] = 3.0 + g[0] * ii + g[1] * jj + g[2] * kk
    s._sc.pc.copy_(_dev(pc))
    s._cache.clear()
    p_dense = s.p.numpy()
    np.testing.assert_array_equal(p_dense[ii, jj, kk], pc[_compact_index(s, ii, jj, kk)])
    d = s._descriptor()
    _lib.call("S2.8", "gf_scene_g2p", ctypes.byref(d), _lib.ptr(s._sc.pc), _lib.GF_TRANSFER_FLIP,
              0.0, dt, rho, fb, s._stream())
    s._cache.clear()
    shift = -dt / (rho * dx32) * g          # u - us on faces between two fluid cells
    # post-projection faces inside the fluid: us + shift
    for comp, key in enumerate(("u", "v", "w")):
        f = getattr(s, key).numpy()
        us = _linear_faces(cs, shape, dx32)[comp]
        sl = [slice(1, N - 1)] * 3
        sl[comp] = slice(2, N - 1)
        sl = tuple(sl)
        assert gu.rel_err(f[sl], us[sl] + shift[comp]) < TOL, key
    pp, vv = s.get_particles()
    inner = ((pos >= 2 * dx) & (pos < (N - 2) * dx)).all(axis=1)
    assert inner.sum() > 1000
    x = pos[inner].astype(np.float64)
    old = cs[:, 0] + x @ cs[:, 1:].T
    want = fb * (vel[inner].astype(np.float64) + shift) + (1 - fb) * (old + shift)
    assert gu.rel_err(vv[inner], want) < TOL
    assert gu.rel_err(pp[inner], x + want * dt) < TOL
