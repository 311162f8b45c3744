"""Particle reseeding (S2.11 / S2.11.GPU) on the HIP device.

Mirrors the reference's ``sim/reseed.py`` surface: :class:`ReseedConfig`,
``RESEED_GPU_THRESHOLD``, :func:`reseed_particles_gpu` (device arrays in and
out, reseed.py:267-396) and :func:`reseed_particles` (numpy in and out,
reseed.py:43-155).  Both run the same kernels (``gf_reseed_*``): per-cell
counts from the binned positions, the first ``max_per_cell`` particles of
each cell kept in index order, and ``max(min_per_cell - count, 0)`` particles
emitted into every cell that is fluid in the marker or holds particles and is
not solid, at ``(cell + 0.5 + jitter) * dx`` with the jitter drawn from the
caller's generator exactly as the reference draws it.

Which particles of an over-full cell are dropped is implementation-defined in
the reference (the first arrivals of its atomics on the GPU, ``rng.choice`` on
the host); here it is always the highest indices, so the result is
deterministic and the generator is used for the jitter only.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from .blocks import block
from .device import DeviceArray, as_tensor, default_device

RESEED_GPU_THRESHOLD = 100_000


@dataclass
class ReseedConfig:
    """Per-cell density bounds (reseed.py:34-39)."""
    min_per_cell: int = 4
    max_per_cell: int = 16
    jitter: float = 0.3
    every_n_frames: int = 5


def _dev_tensor(a, width, device):
    if a is None:
        return None
    t = a.t if isinstance(a, DeviceArray) else a
    t = as_tensor(t, torch.float32, device)
    return t.reshape(-1, width).contiguous() if width > 1 else t.reshape(-1).contiguous()


def _run(pos, vel, marker, dx, cfg, rng, attrs, drop_solid, emit_fill, device):
    """Shared body: returns (pos, vel, [attrs...], n_emit, n_culled) as
    device tensors; attrs = [(tensor, width)]."""
    dev = torch.device(device) if device is not None else default_device()
    p = _dev_tensor(pos, 3, dev)
    v = _dev_tensor(vel, 3, dev)
    m = marker.t if isinstance(marker, DeviceArray) else marker
    m = as_tensor(m, torch.int32, dev).contiguous()
    if m.dim() != 3:
        raise ValueError("marker must be an (nx, ny, nz) int32 array")
    nx, ny, nz = (int(x) for x in m.shape)
    n = int(p.shape[0])
    if v.shape[0] != n:
        raise ValueError(f"vel holds {v.shape[0]} rows for {n} positions")
    st = _lib.stream_handle(dev)
    nws = ctypes.c_int64()
    _lib.call("S2.11.GPU", "gf_reseed_workspace", n, nx, ny, nz, ctypes.byref(nws))
    ws = torch.empty(nws.value, dtype=torch.int32, device=dev)
    counts = (ctypes.c_int * 2)()
    _lib.call("S2.11.GPU", "gf_reseed_plan_sync", _lib.ptr(p) if n else None, n, _lib.ptr(m),
              nx, ny, nz, float(dx), int(cfg.min_per_cell), int(cfg.max_per_cell),
              int(drop_solid), _lib.ptr(ws), counts, st)
    n_alive, n_emit = int(counts[0]), int(counts[1])
    jit = None
    if n_emit:
        jit = (rng.random((n_emit, 3)) - 0.5).astype(np.float32) * cfg.jitter
    total = n_alive + n_emit

    def compact(t, width):
        out = torch.empty((total, width) if width > 1 else (total,), dtype=torch.float32,
                          device=dev)
        _lib.call("S2.11.GPU", "gf_reseed_compact", _lib.ptr(t), width, n, nx, ny, nz,
                  _lib.ptr(ws), _lib.ptr(out), st)
        return out

    new_p, new_v = compact(p, 3), compact(v, 3)
    new_v[n_alive:] = 0.0
    if n_emit:
        jt = torch.as_tensor(np.ascontiguousarray(jit), device=dev)
        _lib.call("S2.11.GPU", "gf_reseed_emit", n, nx, ny, nz, float(dx), _lib.ptr(ws),
                  _lib.ptr(jt), _lib.ptr(new_p[n_alive:]), st)
    outs = []
    for (t, width), fill in zip(attrs, emit_fill):
        if t is None:
            outs.append(None)
            continue
        if t.shape[0] != n:
            raise ValueError(f"attribute holds {t.shape[0]} rows for {n} positions")
        o = compact(t, width)
        o[n_alive:] = fill
        outs.append(o)
    return new_p, new_v, outs, n_emit, n - n_alive


@block("S2.11.GPU", "GPU reseed orchestrator: count → rank-cull → compact → emit")
def reseed_particles_gpu(pos_wp, vel_wp, marker_wp, dx: float, cfg: ReseedConfig,
                         rng: np.random.Generator, attr_color_wp=None,
                         attr_temperature_wp=None, device: Optional[str] = None):
    """Return ``(new_pos, new_vel, new_color, new_temperature, n_emitted,
    n_culled)`` as in reseed.py:267-396: device arrays (fresh allocations),
    particles in solid cells dropped, emitted particles with zero velocity,
    white colour and temperature 0; ``n_culled`` counts every dropped
    particle."""
    dev = device
    if dev is None:
        for a in (pos_wp, marker_wp):
            t = a.t if isinstance(a, DeviceArray) else a
            if isinstance(t, torch.Tensor) and t.device.type == "cuda":
                dev = t.device
                break
    attrs = [(_dev_tensor(attr_color_wp, 3, dev or default_device()), 3),
             (_dev_tensor(attr_temperature_wp, 1, dev or default_device()), 1)]
    p, v, (c, t), n_emit, n_cull = _run(pos_wp, vel_wp, marker_wp, dx, cfg, rng, attrs,
                                        drop_solid=True, emit_fill=(1.0, 0.0), device=dev)
    wrap = lambda x, vec: None if x is None else DeviceArray(x, vec=vec)  # noqa: E731
    return wrap(p, 1), wrap(v, 1), wrap(c, 1), wrap(t, 0), n_emit, n_cull


@block("S2.11", "Reseed particles to maintain bounded per-cell density")
def reseed_particles(pos: np.ndarray, vel: np.ndarray, marker_host: np.ndarray, dx: float,
                     cfg: ReseedConfig, rng: np.random.Generator,
                     attr_color: np.ndarray | None = None,
                     attr_temperature: np.ndarray | None = None):
    """Host-array form (reseed.py:43-155), run on the HIP device: particles in
    solid cells are kept, ``n_culled`` counts the over-capacity drops only,
    emitted particles get colour 0 and temperature 20.  Returns
    ``(new_pos, new_vel, n_emitted, n_culled)``, or with
    ``(..., new_color, new_temperature)`` appended when an attribute was
    given, as numpy arrays."""
    dev = default_device()
    n = len(pos)
    attrs = [(_dev_tensor(attr_color, 3, dev), 3), (_dev_tensor(attr_temperature, 1, dev), 1)]
    p, v, (c, t), n_emit, n_cull = _run(np.asarray(pos, np.float32).reshape(n, 3),
                                        np.asarray(vel, np.float32).reshape(n, 3),
                                        np.asarray(marker_host, np.int32), dx, cfg, rng, attrs,
                                        drop_solid=False, emit_fill=(0.0, 20.0), device=dev)
    new_pos, new_vel = p.cpu().numpy(), v.cpu().numpy()
    if attr_color is None and attr_temperature is None:
        return new_pos, new_vel, n_emit, n_cull
    return (new_pos, new_vel, n_emit, n_cull, None if c is None else c.cpu().numpy(),
            None if t is None else t.cpu().numpy())
