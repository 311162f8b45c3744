"""Sub-dense box bookkeeping (F3.7, reference solvers/solver3d.py:3669-3906).

The reference shrinks its dense grid arrays to the bounding box of the 8^3
tiles that hold fluid, dilated by ``sub_dilation`` cells, and runs every grid
kernel on that box with an offset.  On MI355X the grid fields stay full-size
(288 GB of HBM; the pressure vectors are already compacted to active tiles):
the box is a window.  The solver's grid attributes show the box, and the
viscosity sweep runs over the box with the box as its stencil bounds, as the
reference's does on its box-shaped arrays.  This module holds the
host-side part: the box, its rebuild policy and the two one-shot warnings.
It uses numpy only, so it can be exercised without a device.
"""
from __future__ import annotations

import sys

import numpy as np

TILE = 8


class SubDenseBox:
    """The box of cells [offset, offset + shape) and when to move it.

    Follows _compute_active_bbox (solver3d.py:3670-3702), _pos_to_sub_cell
    (3704-3726), _check_particles_in_sub_bbox (3728-3768),
    _should_rebuild_sub_dense (3770-3819) and the bookkeeping half of
    _rebuild_sub_dense (3821-3906).
    """

    def __init__(self, nx: int, ny: int, nz: int, enabled: bool, rebuild_every: int,
                 dilation: int):
        self.full = (int(nx), int(ny), int(nz))
        self.enabled = bool(enabled)
        self.rebuild_every = max(1, int(rebuild_every))
        self.dilation = max(0, int(dilation))
        self.offset = (0, 0, 0)
        self.shape = self.full
        self.last_rebuild_frame = -1      # never rebuilt
        self.oob_warned = False
        self.extent_warned = False

    @property
    def active(self) -> bool:
        """True when the box is smaller than the grid."""
        return self.offset != (0, 0, 0) or self.shape != self.full

    def bbox(self, marker: np.ndarray, dilation: int | None = None):
        """((lo), (hi)) in cells covering every fluid (marker == 1) cell's 8^3
        tile (or, when an extent is not a multiple of 8, every fluid cell),
        dilated and clamped to the grid; None without fluid cells."""
        d = self.dilation if dilation is None else int(dilation)
        nx, ny, nz = self.full
        m = np.asarray(marker)
        if nx % TILE == 0 and ny % TILE == 0 and nz % TILE == 0:
            t = (m == 1).reshape(nx // TILE, TILE, ny // TILE, TILE, nz // TILE, TILE)
            occ = t.any(axis=(1, 3, 5))
            if not occ.any():
                return None
            idx = np.nonzero(occ)
            lo = tuple(int(a.min()) * TILE for a in idx)
            hi = tuple((int(a.max()) + 1) * TILE for a in idx)
        else:
            idx = np.nonzero(m == 1)
            if idx[0].size == 0:
                return None
            lo = tuple(int(a.min()) for a in idx)
            hi = tuple(int(a.max()) + 1 for a in idx)
        lo = tuple(max(0, c - d) for c in lo)
        hi = tuple(min(n, c + d) for c, n in zip(hi, self.full))
        return lo, hi

    def pos_to_cell(self, pos, dx: float):
        """(local cell, inside) of one world position: floor(pos/dx) in FP32,
        minus the offset, inside when within the box."""
        p = np.asarray(pos, dtype=np.float32).reshape(3)
        g = [int(np.floor(p[a] / np.float32(dx))) for a in range(3)]
        loc = tuple(g[a] - self.offset[a] for a in range(3))
        inside = all(0 <= loc[a] < self.shape[a] for a in range(3))
        return loc, inside

    def count_outside(self, pos: np.ndarray, dx: float, frame_idx: int) -> int:
        """Particles whose cell lies outside the box; the first time any do,
        one warning on stderr naming the frame, the box and the two knobs."""
        if not self.enabled or pos is None or len(pos) == 0 or not self.active:
            return 0
        cells = np.floor(np.asarray(pos, dtype=np.float32) / np.float32(dx)).astype(np.int64)
        loc = cells - np.asarray(self.offset, dtype=np.int64)
        out = ((loc < 0) | (loc >= np.asarray(self.shape, dtype=np.int64))).any(axis=1)
        n_out = int(out.sum())
        if n_out > 0 and not self.oob_warned:
            print(f"[gpufluid] WARNING: frame {frame_idx}: {n_out} of {len(pos)} particles "
                  f"({100.0 * n_out / len(pos):.1f}%) are outside the sub-dense bbox "
                  f"(offset={self.offset}, shape={self.shape}); grid transfers for them are "
                  f"clamped to the box.  Increase sub_dilation (now {self.dilation}) or "
                  f"decrease sub_rebuild_every (now {self.rebuild_every}).",
                  file=sys.stderr, flush=True)
            self.oob_warned = True
        return n_out

    def should_rebuild(self, frame_idx: int, marker: np.ndarray | None) -> bool:
        """Rebuild on the first call, every rebuild_every frames, or when the
        undilated fluid bbox comes within dilation cells of the box edge."""
        if not self.enabled:
            return False
        if self.last_rebuild_frame < 0:
            return True
        if frame_idx - self.last_rebuild_frame >= self.rebuild_every:
            return True
        if marker is None:
            return False
        raw = self.bbox(marker, dilation=0)
        if raw is None:
            return False
        lo, hi = raw
        cur_lo = self.offset
        cur_hi = tuple(o + s for o, s in zip(self.offset, self.shape))
        d = self.dilation
        return any(lo[a] - cur_lo[a] < d or cur_hi[a] - hi[a] < d for a in range(3))

    def rebuild(self, lo, hi) -> None:
        """Move the box to [lo, hi); the first box covering more than 80% of
        the grid prints a warning (dispersed fluid: the box saves nothing)."""
        lo = tuple(int(c) for c in lo)
        hi = tuple(int(c) for c in hi)
        shape = tuple(h - l for l, h in zip(lo, hi))
        if any(s <= 0 for s in shape):
            raise ValueError(f"sub-dense shape {shape} non-positive")
        if any(l < 0 for l in lo) or any(h > n for h, n in zip(hi, self.full)):
            raise ValueError(f"sub-dense box {lo}..{hi} outside the {self.full} grid")
        self.offset = lo
        self.shape = shape
        ratio = float(np.prod(shape, dtype=np.float64)) / float(np.prod(self.full, dtype=np.float64))
        if ratio > 0.8 and not self.extent_warned:
            print(f"[gpufluid] WARNING: the sub-dense bbox {shape} covers {100.0 * ratio:.1f}% "
                  f"of the {self.full} grid (scattered topology: dispersed fluid leaves "
                  f"nothing to trim).  Run with enable_sub_dense=False, or lower "
                  f"`sub_dilation`={self.dilation} if it is larger than needed.",
                  file=sys.stderr, flush=True)
            self.extent_warned = True

    def face_slices(self, axis: int | None):
        """Index of the box's part of a face array (axis 0/1/2) or of a cell
        array (None)."""
        return tuple(slice(o, o + s + (1 if a == axis else 0))
                     for a, (o, s) in enumerate(zip(self.offset, self.shape)))
