"""gpufluid_amd — MI355X (gfx950) FLIP/PIC/APIC substep for gpufluid.

Drop-in for the substep hot path of ArkBee/GPUFLUID (gpufluid 0.0.2):
``FlipSolver3D`` keeps the reference solver's Python surface and
``gpufluid_amd.kernels`` the S2 block interface; the compute runs in
hand-written HIP kernels (libgpufluid_hip.so, C ABI in include/gpufluid_hip.h)
loaded with ctypes.  Build the library with ``python -m gpufluid_amd.build``.
"""
from __future__ import annotations

__version__ = "0.1.0"

from .blocks import BlockError, block, by_layer, find, get_registry  # noqa: F401


def __getattr__(name):
    # Solver classes import torch; load them lazily so the registry and the
    # build script stay importable without initialising a device.
    if name in ("FlipSolver3D", "FlipDivergenceError", "cfl_substep_count",
                "cfl_substep_count_gpu"):
        from . import solver3d
        return getattr(solver3d, name)
    if name == "kernels":
        import importlib
        return importlib.import_module(".kernels", __name__)
    raise AttributeError(name)
