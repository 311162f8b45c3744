"""Block-sparse and per-tile S2 block callables (S2.16 k_mark_active_blocks,
S2.6.4 k3_jacobi_pressure_blocksparse / k3_jacobi_pressure_per_tile, S2.6.5
k3_gauss_seidel_rb_per_tile, S2.6.6 k3_*_per_tile; solver3d.py:952-1197,
1270-1357, 2020-2212) against the oracle's dense sweeps restricted to the
listed blocks: bit-exact inside them, untouched outside."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import flip_oracle as fo

pytestmark = pytest.mark.gpu

BS = 8
SENTINEL = np.float32(7.25)


def _scene(seed=3, shape=(20, 18, 22)):
    rng = np.random.default_rng(seed)
    m = fo.wall_marker(*shape)
    inner = m[1:-1, 1:-1, 1:-1]
    inner[...] = np.where(rng.random(inner.shape) < 0.6, fo.FLUID, fo.AIR)
    m[:, 9:, 12:] = np.where(m[:, 9:, 12:] == fo.SOLID, fo.SOLID, fo.AIR)   # empty blocks
    nb = tuple(-(-n // BS) for n in shape)
    act = np.zeros(nb, np.int32)
    for b in np.ndindex(*nb):
        sl = tuple(slice(c * BS, (c + 1) * BS) for c in b)
        act[b] = int((m[sl] == fo.FLUID).any())
    return rng, m, act


def _listed(act, rng, drop_last=True):
    """Active block coordinates in shuffled order and the live count (the
    last listed block is left out of the count)."""
    coords = np.argwhere(act == 1).astype(np.int32)
    coords = coords[rng.permutation(len(coords))]
    n = len(coords) - 1 if drop_last else len(coords)
    return coords, n


def _cellmask(mask_blocks, shape):
    big = np.kron(mask_blocks, np.ones((BS, BS, BS), bool))
    return big[:shape[0], :shape[1], :shape[2]]


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), device="cuda:0")


def test_mark_active_blocks():
    from gpufluid_amd import kernels
    _, m, act = _scene()
    out = _dev(np.zeros(act.shape, np.int32))
    kernels.k_mark_active_blocks(_dev(m), out, BS)
    np.testing.assert_array_equal(out.cpu().numpy(), act)
    assert 0 < act.sum() < act.size


def test_blocksparse_jacobi_equals_dense_sweep():
    from gpufluid_amd import kernels
    rng, m, act = _scene(4)
    div = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    p_in = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    out = _dev(np.full(m.shape, SENTINEL))
    kernels.k3_jacobi_pressure_blocksparse(_dev(p_in), out, _dev(div), _dev(m), _dev(act), BS)
    np.testing.assert_array_equal(out.cpu().numpy(), fo.jacobi_sweep(p_in, div, m))


@pytest.mark.parametrize("kind", ["jacobi", "gsrb", "apply_A", "invM", "axpy"])
def test_per_tile_forms_touch_only_listed_blocks(kind):
    from gpufluid_amd import kernels
    rng, m, act = _scene(5)
    coords, n = _listed(act, rng)
    mask = np.zeros(act.shape, bool)
    for c in coords[:n]:
        mask[tuple(c)] = True
    cells = _cellmask(mask, m.shape)
    print(f"[blocks] {kind}: {n} of {len(coords)} listed blocks live", flush=True)
    co, na, mt = _dev(coords), _dev(np.array([n], np.int32)), _dev(m)
    done = _dev(np.zeros(1, np.int32))
    x = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    div = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    out = _dev(np.full(m.shape, SENTINEL))
    if kind == "jacobi":
        kernels.k3_jacobi_pressure_per_tile(_dev(x), out, _dev(div), mt, co, BS, na)
        want = fo.jacobi_sweep(x, div, m)
    elif kind == "gsrb":
        p = _dev(x)
        for color in (0, 1):
            kernels.k3_gauss_seidel_rb_per_tile(p, _dev(div), mt, co, BS, color, na)
        out = p
        # the listed cells' sweep: red then black over those cells only
        want = x.copy()
        for color in (0, 1):
            half = fo.gsrb_half_sweep(want, div, m, color)
            want = np.where(cells, half, want)
        np.testing.assert_array_equal(out.cpu().numpy(), want)
        return
    elif kind == "apply_A":
        kernels.k3_apply_A_per_tile(_dev(x), out, mt, co, BS, na, done)
        want = fo.apply_A(x, m)
    elif kind == "invM":
        diag = fo.precond_diag(m).astype(np.float32)
        kernels.k3_apply_invM_per_tile(_dev(x), _dev(diag), out, mt, co, BS, na, done)
        want = np.where(m == fo.FLUID, x / diag, 0).astype(np.float32)
    else:
        a = _dev(np.array([0.375], np.float32))
        kernels.k3_axpy_devscalar_per_tile(a, -1.0, _dev(x), _dev(div), out, mt, co, BS, na, done)
        want = np.where(m == fo.FLUID, div + np.float32(-0.375) * x, 0).astype(np.float32)
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got[cells], want[cells])
    assert (got[~cells] == SENTINEL).all()


def test_per_tile_dot_and_done_flag():
    from gpufluid_amd import kernels
    rng, m, act = _scene(6)
    coords, n = _listed(act, rng)
    mask = np.zeros(act.shape, bool)
    for c in coords[:n]:
        mask[tuple(c)] = True
    cells = _cellmask(mask, m.shape)
    a = rng.normal(size=m.shape).astype(np.float32)
    b = rng.normal(size=m.shape).astype(np.float32)
    out = _dev(np.array([0.5], np.float32))
    co, na, mt = _dev(coords), _dev(np.array([n], np.int32)), _dev(m)
    kernels.k3_dot_fluid_per_tile(_dev(a), _dev(b), out, mt, co, BS, na,
                                  _dev(np.zeros(1, np.int32)))
    sel = cells & (m == fo.FLUID)
    want = 0.5 + float(np.sum(a[sel].astype(np.float64) * b[sel]))
    assert abs(float(out.cpu()[0]) - want) <= 1e-5 * max(1.0, abs(want))
    # a set stop flag makes the PCG kernels pass through
    y = _dev(np.full(m.shape, SENTINEL))
    kernels.k3_apply_A_per_tile(_dev(a), y, mt, co, BS, na, _dev(np.ones(1, np.int32)))
    assert (y.cpu().numpy() == SENTINEL).all()
