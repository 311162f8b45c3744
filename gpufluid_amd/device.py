"""Device buffers for the solver façade.

Fields are torch tensors on the HIP device (torch is the allocator and stream
provider only).  ``DeviceArray`` wraps one so callers written against the
reference's ``wp.array`` attributes keep working: ``.numpy()``, ``.shape``,
``.dtype``, ``.zero_()``, ``.flatten()`` (solver attributes read by callers,
see the reference tests e.g. tests/test_b5_a_pcg_graph.py:84-97).
"""
from __future__ import annotations

import numpy as np
import torch


class DeviceArray:
    """A torch device tensor with a numpy-style surface.

    ``vec`` > 0 marks arrays of vectors (vec3 positions, mat33 affine
    matrices): ``shape`` then reports only the leading element count, as a
    Warp array of vec3 does, while ``numpy()`` returns the full (n, 3) array.
    """

    __slots__ = ("t", "vec")

    def __init__(self, tensor: torch.Tensor, vec: int = 0):
        self.t = tensor
        self.vec = vec

    @property
    def shape(self):
        s = tuple(self.t.shape)
        return s[: len(s) - self.vec] if self.vec else s

    @property
    def dtype(self):
        return {torch.float32: np.float32, torch.float64: np.float64,
                torch.int32: np.int32, torch.uint8: np.uint8}.get(self.t.dtype, self.t.dtype)

    @property
    def device(self):
        return self.t.device

    def numpy(self) -> np.ndarray:
        return self.t.detach().cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a.astype(dtype) if dtype is not None else a

    def zero_(self):
        self.t.zero_()
        return self

    def flatten(self):
        return DeviceArray(self.t.reshape(-1))

    def torch(self) -> torch.Tensor:
        return self.t

    def __len__(self):
        return self.shape[0]

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.t.dtype}, device={self.t.device})"


def as_tensor(x, dtype, device) -> torch.Tensor:
    """Accept a DeviceArray, torch tensor or array-like and return a
    contiguous tensor of ``dtype`` on ``device``."""
    if isinstance(x, DeviceArray):
        x = x.t
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device).contiguous()


def default_device() -> torch.device:
    """The HIP device; raises when none is visible (there is no CPU path)."""
    if not torch.cuda.is_available():
        raise RuntimeError("gpufluid_amd needs a ROCm/HIP device (torch.cuda is unavailable)")
    return torch.device("cuda", torch.cuda.current_device())
