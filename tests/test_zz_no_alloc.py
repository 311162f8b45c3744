"""gf_scene_pcg_enqueue_noalloc (work buffers in a static device array) against
gf_scene_pcg_enqueue (work buffers from the descriptor), through the C-ABI.

The two run the same kernels on the same inputs, so the pressure and the
solver results must be bit-identical.  Checked on two calls in a row, the
second on a smaller scene than the first (the static array then holds the
first call's larger state), for both precisions, both iteration forms and the
device-side tile count; and the size limit is refused before any launch.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import gfutil as gu

pytestmark = pytest.mark.gpu


def _prepared(n, seed):
    """A dam-break scene after sort + P2G: b, flags and the tile tables set."""
    shape, dx, pos, vel = gu.dam_scene(n=n, seed=seed, vel_sigma=0.3)
    s = gu.make_solver(shape, dx, pos, vel)
    gu.run_sort(s)
    gu.run_p2g(s, 1.0 / 240, staged=True)
    torch.cuda.synchronize()
    return s


def _stats(s):
    from gpufluid_amd import _lib
    it, res, nrep = (ctypes.c_int * 1)(), (ctypes.c_double * 1)(), (ctypes.c_int * 1)()
    d = s._descriptor()
    _lib.call("S2.6.3", "gf_scene_pcg_stats", ctypes.byref(d), it, res, nrep, s._stream())
    return int(it[0]), float(res[0]), int(nrep[0])


def _solve(s, entry, nact, prec, tol=1e-8, max_iter=400, ce=10):
    from gpufluid_amd import _lib
    if prec == _lib.GF_PCG_FP64:
        s._sc.ensure_fp64()
    s._sc.pc.fill_(7.0)                 # a stale value the solve must overwrite
    s._sc.scalars.fill_(0.0)
    d = s._descriptor()
    chunks = -(-max_iter // ce)
    _lib.call("S2.6.3", entry, ctypes.byref(d), nact, prec, tol, max_iter, ce, -1, 0, chunks,
              s._stream())
    torch.cuda.synchronize()
    return s._sc.pc.clone(), _stats(s)


@pytest.mark.parametrize("variant", [1, 2], ids=["two_kernel", "one_kernel"])
@pytest.mark.parametrize("prec", [0, 1], ids=["mixed", "fp64"])
def test_noalloc_matches_enqueue_on_two_calls(variant, prec):
    from gpufluid_amd import _lib, autotune
    autotune.set_pcg_variant(variant)
    try:
        for n, seed in ((64, 3), (32, 4)):          # second call: a smaller scene
            s = _prepared(n, seed)
            nact = s.n_active_tiles
            print(f"[noalloc] {n}^3, {nact} tiles, variant {variant}, precision {prec}", flush=True)
            for count in (nact, _lib.GF_NACT_DEVICE):
                pa, sa = _solve(s, "gf_scene_pcg_enqueue", count, prec)
                pb, sb = _solve(s, "gf_scene_pcg_enqueue_noalloc", count, prec)
                assert sa[0] > 0 and sa[1] < 1e-8, sa
                assert sa == sb, (sa, sb)
                assert torch.equal(pa, pb)
    finally:
        autotune.set_pcg_variant(0)


def test_noalloc_refuses_more_tiles_than_its_array_holds():
    """Checked on the host before any launch: 2049 active tiles, or a grid of
    17^3 = 4913 tiles when the count is to be read on the device."""
    from gpufluid_amd import FlipSolver3D, _lib
    from gpufluid_amd.blocks import BlockError
    s = FlipSolver3D(136, 136, 136, device="cuda:0")
    d = s._descriptor()
    for count, what in ((2049, "2049 active tiles"), (_lib.GF_NACT_DEVICE, "4913 grid tiles")):
        with pytest.raises(BlockError, match=f"{what} exceed the 2048"):
            _lib.call("S2.6.3", "gf_scene_pcg_enqueue_noalloc", ctypes.byref(d), count, 0, 1e-8,
                      10, 10, -1, 0, 1, s._stream())
