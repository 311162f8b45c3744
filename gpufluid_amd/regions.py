"""Inflow and outflow regions (D4.7) and the per-frame event queue (G1.17).

Host-side data classes with the reference's fields and methods
(``domain/regions.py:23-151``, ``primitives/frame_events.py:51-134``): an
:class:`InflowBox` publishes a :class:`FluidEmitEvent` of uniformly random
positions each frame it is active, an :class:`OutflowBox` a
:class:`FluidOutflowEvent` describing its box; ``FlipSolver3D.prepare_frame``
drains the queue, compacts the particles on the device and appends the
emitted ones.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .blocks import block


@dataclass
class FluidEmitEvent:
    """Particles to append this frame (frame_events.py:51-73)."""
    positions: np.ndarray
    velocities: np.ndarray
    color: Optional[Tuple[float, float, float]] = None
    temperature: Optional[float] = None


@dataclass
class FluidOutflowEvent:
    """A box whose particles are removed this frame (frame_events.py:76-91)."""
    lo: Tuple[float, float, float]
    hi: Tuple[float, float, float]
    inside: bool = True


class FrameEventQueue:
    """Emit and outflow channels, drained in push order
    (frame_events.py:94-134)."""

    def __init__(self):
        self._emits: List[FluidEmitEvent] = []
        self._outflows: List[FluidOutflowEvent] = []

    def push_emit(self, ev: FluidEmitEvent) -> None:
        self._emits.append(ev)

    def push_outflow(self, ev: FluidOutflowEvent) -> None:
        self._outflows.append(ev)

    def drain_emits(self) -> List[FluidEmitEvent]:
        out, self._emits = self._emits, []
        return out

    def drain_outflows(self) -> List[FluidOutflowEvent]:
        out, self._outflows = self._outflows, []
        return out

    def clear(self) -> None:
        self._emits = []
        self._outflows = []


@dataclass
class InflowBox:
    """A box that emits round(rate_per_sec * frame_dt) particles per active
    frame (regions.py:23-72)."""
    lo: Tuple[float, float, float]
    hi: Tuple[float, float, float]
    velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    rate_per_sec: float = 5000.0
    frame_start: int = 0
    frame_end: int = 10_000
    color: Optional[Tuple[float, float, float]] = None
    temperature: Optional[float] = None

    def emit(self, frame_idx: int, frame_dt: float,
             rng: np.random.Generator) -> Optional[np.ndarray]:
        if frame_idx < self.frame_start or frame_idx > self.frame_end:
            return None
        n = int(round(self.rate_per_sec * frame_dt))
        if n <= 0:
            return None
        lo = np.asarray(self.lo, dtype=np.float32)
        hi = np.asarray(self.hi, dtype=np.float32)
        return lo + (hi - lo) * rng.random((n, 3)).astype(np.float32)

    def publish_for_frame(self, queue: FrameEventQueue, frame_idx: int, frame_dt: float,
                          rng: np.random.Generator) -> None:
        pos = self.emit(frame_idx, frame_dt, rng)
        if pos is None or len(pos) == 0:
            return
        vel = np.broadcast_to(np.asarray(self.velocity, dtype=np.float32), (len(pos), 3)).copy()
        queue.push_emit(FluidEmitEvent(positions=pos, velocities=vel, color=self.color,
                                       temperature=self.temperature))


@dataclass
class OutflowBox:
    """A box that removes the particles inside it (bounds inclusive) while
    active (regions.py:75-110)."""
    lo: Tuple[float, float, float]
    hi: Tuple[float, float, float]
    frame_start: int = 0
    frame_end: int = 10_000

    def __post_init__(self):
        for axis, (lo_v, hi_v) in enumerate(zip(self.lo, self.hi)):
            if hi_v < lo_v:
                name = "xyz"[axis]
                raise ValueError(f"OutflowBox: hi[{name}]={hi_v} < lo[{name}]={lo_v}. Outflow box "
                                 f"AABB must have hi >= lo on every axis (got lo={self.lo}, "
                                 f"hi={self.hi}).")

    def publish_for_frame(self, queue: FrameEventQueue, frame_idx: int) -> None:
        if frame_idx < self.frame_start or frame_idx > self.frame_end:
            return
        queue.push_outflow(FluidOutflowEvent(lo=tuple(float(x) for x in self.lo),
                                             hi=tuple(float(x) for x in self.hi), inside=True))


@block("D4.7", "Emit particles from all inflow boxes for this frame")
def apply_inflows(inflows: List[InflowBox], frame_idx: int, frame_dt: float,
                  rng: np.random.Generator):
    """(positions, velocities) of every active inflow, in list order
    (regions.py:131-151)."""
    q = FrameEventQueue()
    for inf in inflows or []:
        inf.publish_for_frame(q, frame_idx, frame_dt, rng)
    ev = q.drain_emits()
    if not ev:
        z = np.zeros((0, 3), dtype=np.float32)
        return z, z.copy()
    return (np.concatenate([e.positions for e in ev]),
            np.concatenate([e.velocities for e in ev]))
