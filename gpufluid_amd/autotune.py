"""Launch-shape autotuner.

The tile edge (8 cells) is fixed by the compact pressure layout; what varies
with the workload is the launch shape of the persistent PCG grid (workgroups),
the PCG iteration form (two kernels per iteration, or one), and, for
graph-captured batches, the PCG iteration budget inside the graph.
``tune_pcg_workgroups`` times a caller-supplied solve on each candidate grid
size, keeps the fastest, and caches the choice per (device, workload key) for
the life of the process.  No reference counterpart (Warp picks launch shapes
itself).
"""
from __future__ import annotations

import ctypes
import time

import torch

from . import _lib
from .blocks import block

_CACHE: dict = {}


def set_pcg_workgroups(n: int) -> None:
    """Persistent-grid size of the PCG kernels (0 = built-in choice)."""
    _lib.call("X3.1", "gf_tune_set", b"pcg_workgroups", int(n))


def get_pcg_workgroups() -> int:
    v = ctypes.c_int(0)
    _lib.call("X3.1", "gf_tune_get", b"pcg_workgroups", ctypes.byref(v))
    return v.value


PCG_TWO_KERNEL = 1
PCG_ONE_KERNEL = 2


def set_pcg_variant(v: int) -> None:
    """PCG iteration form: 1 two kernels per iteration, 2 one kernel
    (Chronopoulos-Gear), 0 built-in choice by active-tile count."""
    _lib.call("X3.1", "gf_tune_set", b"pcg_variant", int(v))


def get_pcg_variant() -> int:
    v = ctypes.c_int(0)
    _lib.call("X3.1", "gf_tune_get", b"pcg_variant", ctypes.byref(v))
    return v.value


def set_pcg_x_every(v: int) -> None:
    """Two-kernel PCG: the iterate x takes its terms alpha p on every n-th
    update kernel (n = 1..4; 0, the default, or above 4 means 4) and on each
    check chunk's last, the pending terms in iteration order (bit-identical
    results; n-1 reads and writes of x fewer per n iterations, for n-1 reads
    of the pending directions)."""
    _lib.call("X3.1", "gf_tune_set", b"pcg_x_every", int(v))


def get_pcg_x_every() -> int:
    v = ctypes.c_int(0)
    _lib.call("X3.1", "gf_tune_get", b"pcg_x_every", ctypes.byref(v))
    return v.value


@block("X3.1", "Autotune the PCG persistent-grid size on the live workload")
def tune_pcg_workgroups(solve, key, candidates=(512, 768, 1024, 1536, 2048), repeats: int = 2,
                        device=None):
    """Time ``solve()`` (one pressure solve, or a substep) for each candidate
    workgroup count, ``repeats`` times each, and keep the fastest.  Returns
    (best, {candidate: seconds}); the choice stays set and is cached under
    (device name, key)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    ck = (torch.cuda.get_device_name(dev), key)
    if ck in _CACHE:
        best, times = _CACHE[ck]
        set_pcg_workgroups(best)
        return best, times
    times = {}
    for g in candidates:
        set_pcg_workgroups(g)
        best_t = float("inf")
        for _ in range(repeats):
            torch.cuda.synchronize(dev)
            t0 = time.perf_counter()
            solve()
            torch.cuda.synchronize(dev)
            best_t = min(best_t, time.perf_counter() - t0)
        times[g] = best_t
    best = min(times, key=times.get)
    set_pcg_workgroups(best)
    _CACHE[ck] = (best, times)
    return best, times


@block("X3.1", "Autotune the PCG iteration form and grid size on the live workload")
def tune_pcg(solve, key, workgroups=(512, 1024, 2048), variants=(PCG_TWO_KERNEL, PCG_ONE_KERNEL),
             repeats: int = 2, device=None):
    """Time ``solve()`` over every (variant, workgroups) pair and keep the
    fastest.  Returns ((variant, workgroups), {pair: seconds}); the choice
    stays set and is cached under (device name, key)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    ck = (torch.cuda.get_device_name(dev), "variant", key)
    if ck in _CACHE:
        best, times = _CACHE[ck]
        set_pcg_variant(best[0])
        set_pcg_workgroups(best[1])
        return best, times
    times = {}
    for v in variants:
        set_pcg_variant(v)
        for g in workgroups:
            set_pcg_workgroups(g)
            best_t = float("inf")
            for _ in range(repeats):
                torch.cuda.synchronize(dev)
                t0 = time.perf_counter()
                solve()
                torch.cuda.synchronize(dev)
                best_t = min(best_t, time.perf_counter() - t0)
            times[(v, g)] = best_t
    best = min(times, key=times.get)
    set_pcg_variant(best[0])
    set_pcg_workgroups(best[1])
    _CACHE[ck] = (best, times)
    return best, times


def graph_budget(last_iters, check_every: int, cap: int, margin: float = 1.25) -> int:
    """PCG iterations to place inside a captured substep: the largest count
    seen last time plus a margin, in whole chunks, at most the cap."""
    m = int(max(last_iters)) if len(last_iters) else cap
    b = int(m * margin) + check_every
    b = (b + check_every - 1) // check_every * check_every
    return max(check_every, min(b, cap))
