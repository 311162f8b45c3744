"""Simulation cache directory (I6.2 manifest, I6.3 particle dumps).

The reference's on-disk interchange (``io/cache.py:23-93``,
``io/ply.py:185-237``): ``cache.json`` describes the frames a run produced,
``particles/frame_NNNN.npy`` holds each frame's (N, 3) FP32 positions and a
points-only binary PLY may carry the same positions for viewers.  Files
written here are read by the reference's readers and the other way round.
"""
from __future__ import annotations

import json
import time
from dataclasses import asdict, dataclass, field
from pathlib import Path
from typing import List, Optional, Union

import numpy as np

from .blocks import BlockError, block

CACHE_SCHEMA_VERSION = 1
# The reference package version recorded in manifests, so downstream readers
# that key on it see the format they expect.
FORMAT_VERSION = "0.0.2"


@dataclass
class CacheManifest:
    """cache.json fields, defaults and order as io/cache.py:27-65."""
    schema_version: int = CACHE_SCHEMA_VERSION
    gpufluid_version: str = ""
    created_unix: float = 0.0
    fps: int = 24
    frame_count: int = 0
    mesh_pattern: str = "mesh/frame_{:04d}.ply"
    particles_pattern: Optional[str] = None
    domain_size: List[float] = field(default_factory=lambda: [1.0, 1.0, 1.0])
    resolution: List[int] = field(default_factory=lambda: [64, 64, 64])
    dx: float = 1.0 / 64
    notes: str = ""
    solver: str = "flip"
    truncated_at_frame: Optional[int] = None
    truncation_reason: Optional[str] = None

    def mesh_path(self, root: Path, frame: int) -> Path:
        return Path(root) / self.mesh_pattern.format(frame)

    def particles_path(self, root: Path, frame: int) -> Optional[Path]:
        if not self.particles_pattern:
            return None
        return Path(root) / self.particles_pattern.format(frame)

    def to_json(self) -> str:
        return json.dumps(asdict(self), indent=2)


@block("I6.2", "Write cache manifest (cache.json)")
def write_cache_manifest(cache_dir, manifest: CacheManifest) -> Path:
    """Write cache.json (creation time and version filled when unset)."""
    cache_dir = Path(cache_dir)
    cache_dir.mkdir(parents=True, exist_ok=True)
    if manifest.created_unix == 0.0:
        manifest.created_unix = time.time()
    if not manifest.gpufluid_version:
        manifest.gpufluid_version = FORMAT_VERSION
    path = cache_dir / "cache.json"
    path.write_text(manifest.to_json(), encoding="utf-8")
    return path


@block("I6.2", "Read cache manifest (cache.json)")
def read_cache_manifest(cache_dir) -> CacheManifest:
    """Read cache.json; a missing file or another schema version raises
    BlockError("I6.2"), unknown keys are dropped (io/cache.py:81-93)."""
    path = Path(cache_dir) / "cache.json"
    if not path.exists():
        raise BlockError("I6.2", f"no cache.json in {cache_dir}")
    raw = json.loads(path.read_text(encoding="utf-8"))
    if raw.get("schema_version", 0) != CACHE_SCHEMA_VERSION:
        raise BlockError("I6.2", f"unsupported cache schema {raw.get('schema_version')}; "
                                 f"expected {CACHE_SCHEMA_VERSION}")
    fields = set(CacheManifest.__dataclass_fields__)
    return CacheManifest(**{k: v for k, v in raw.items() if k in fields})


@block("I6.3", "Particle dump (numpy .npy of positions)")
def write_particles_npy(path: Union[str, Path], positions) -> None:
    """(N, 3) FP32 positions to .npy (io/ply.py:231-237)."""
    arr = np.asarray(positions, dtype=np.float32)
    if arr.ndim != 2 or arr.shape[1] != 3:
        raise BlockError("I6.3", f"positions must be (N,3); got {arr.shape}")
    np.save(path, arr)


@block("S2.17.6", "Points-only binary PLY of particle positions")
def write_particles_ply(path: Union[str, Path], positions, selection_mask=None) -> int:
    """Binary little-endian PLY with x, y, z float vertices; rows whose
    selection_mask entry is nonzero are left out (io/ply.py:185-228).
    Returns the number of points written."""
    pos = np.asarray(positions, dtype=np.float32)
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise BlockError("S2.17.6", f"positions must be (N,3); got {pos.shape}")
    if selection_mask is not None:
        keep = np.asarray(selection_mask).ravel() == 0
        if keep.shape[0] != pos.shape[0]:
            raise BlockError("S2.17.6",
                             f"selection_mask shape {keep.shape} != positions {pos.shape[0]}")
        pos = pos[keep]
    n = len(pos)
    header = ("ply\nformat binary_little_endian 1.0\n"
              f"element vertex {n}\n"
              "property float x\nproperty float y\nproperty float z\n"
              "end_header\n")
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(np.ascontiguousarray(pos, dtype="<f4").tobytes())
    return n


def dump_frame(solver, cache_dir, frame: int,
               pattern: str = "particles/frame_{:04d}.npy") -> Path:
    """Write one frame's particle positions (caller order) under cache_dir
    with the manifest's particle pattern."""
    path = Path(cache_dir) / pattern.format(frame)
    path.parent.mkdir(parents=True, exist_ok=True)
    pos = solver.pos.numpy() if solver.n_particles else np.zeros((0, 3), np.float32)
    write_particles_npy(path, pos)
    return path


def manifest_for(solver, fps: int, frame_count: int, particles: bool = True,
                 notes: str = "") -> CacheManifest:
    """A manifest describing a solver's grid (domain, resolution, dx)."""
    return CacheManifest(fps=int(fps), frame_count=int(frame_count),
                         particles_pattern="particles/frame_{:04d}.npy" if particles else None,
                         domain_size=[float(d) for d in solver.dom],
                         resolution=[solver.nx, solver.ny, solver.nz], dx=float(solver.dx),
                         notes=notes)
