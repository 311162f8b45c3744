"""Per-section step timing (same contract as the reference StepProfiler,
src/gpufluid/primitives/profiling.py:26-76).

Disabled: ``section()`` is a ``nullcontext`` and nothing is recorded.  Enabled:
each section is bracketed by HIP events on the current stream and synchronised
on exit; ``summarize()`` returns per-section totals in milliseconds and
``call_counts`` how often each fired.  Stages that the fused kernels execute
inside another section are recorded with ``fused(name)`` as zero-duration
entries so that every reference section name still fires once per step.
"""
from __future__ import annotations

from contextlib import contextmanager, nullcontext

from .blocks import block


@block("G1.9", "Step profiler: per-section timing over HIP events")
class StepProfiler:
    def __init__(self, enabled: bool = False, device=None):
        self.enabled = bool(enabled)
        self.device = device
        self._dict: dict[str, list[float]] = {}

    def section(self, name: str):
        if not self.enabled:
            return nullcontext()
        return self._timed(name)

    @contextmanager
    def _timed(self, name: str):
        import torch
        if torch.cuda.is_current_stream_capturing():
            yield
            return
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        try:
            yield
        finally:
            end.record()
            end.synchronize()
            self._dict.setdefault(name, []).append(float(start.elapsed_time(end)))

    def fused(self, name: str) -> None:
        """Record a stage that ran inside another section's kernel (nothing
        while a stream is being captured, like a section)."""
        if self.enabled:
            import torch
            if torch.cuda.is_current_stream_capturing():
                return
            self._dict.setdefault(name, []).append(0.0)

    def summarize(self) -> dict[str, float]:
        return {k: float(sum(v)) for k, v in self._dict.items()}

    def reset(self) -> None:
        self._dict.clear()

    @property
    def call_counts(self) -> dict[str, int]:
        return {k: len(v) for k, v in self._dict.items()}
