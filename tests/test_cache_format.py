"""The cache directory format (io/cache.py:23-93, io/ply.py:185-237): field
names and defaults, schema check, .npy and PLY layouts."""
import json
import struct

import numpy as np
import pytest

from gpufluid_amd import cache
from gpufluid_amd.blocks import BlockError


def test_manifest_round_trip_and_keys(tmp_path):
    m = cache.CacheManifest(fps=30, frame_count=12, particles_pattern="particles/frame_{:04d}.npy",
                            resolution=[32, 48, 32], dx=1 / 48, notes="x",
                            truncated_at_frame=9, truncation_reason="diverged")
    path = cache.write_cache_manifest(tmp_path, m)
    raw = json.loads(path.read_text())
    assert list(raw) == ["schema_version", "gpufluid_version", "created_unix", "fps",
                         "frame_count", "mesh_pattern", "particles_pattern", "domain_size",
                         "resolution", "dx", "notes", "solver", "truncated_at_frame",
                         "truncation_reason"]
    assert raw["schema_version"] == 1 and raw["gpufluid_version"] == "0.0.2"
    assert raw["created_unix"] > 0 and raw["solver"] == "flip"
    back = cache.read_cache_manifest(tmp_path)
    assert back == m
    assert back.particles_path(tmp_path, 7) == tmp_path / "particles/frame_0007.npy"
    assert back.mesh_path(tmp_path, 3) == tmp_path / "mesh/frame_0003.ply"


def test_reader_checks_schema_and_drops_unknown_keys(tmp_path):
    with pytest.raises(BlockError):
        cache.read_cache_manifest(tmp_path)
    (tmp_path / "cache.json").write_text(json.dumps({"schema_version": 2}))
    with pytest.raises(BlockError):
        cache.read_cache_manifest(tmp_path)
    (tmp_path / "cache.json").write_text(json.dumps({"schema_version": 1, "fps": 60,
                                                     "future_key": 1}))
    m = cache.read_cache_manifest(tmp_path)
    assert m.fps == 60 and m.frame_count == 0


def test_particle_npy_and_ply_layouts(tmp_path):
    pos = np.random.default_rng(0).random((17, 3)).astype(np.float32)
    cache.write_particles_npy(tmp_path / "p.npy", pos)
    back = np.load(tmp_path / "p.npy", allow_pickle=False)
    assert back.dtype == np.float32 and back.shape == (17, 3)
    np.testing.assert_array_equal(back, pos)
    with pytest.raises(BlockError):
        cache.write_particles_npy(tmp_path / "bad.npy", pos[:, :2])
    mask = np.zeros(17, np.int32)
    mask[[2, 5]] = 1
    n = cache.write_particles_ply(tmp_path / "p.ply", pos, mask)
    data = (tmp_path / "p.ply").read_bytes()
    head, body = data.split(b"end_header\n", 1)
    assert n == 15 and b"element vertex 15" in head and head.startswith(b"ply\nformat "
                                                                      b"binary_little_endian 1.0\n")
    pts = np.array(struct.unpack("<" + "f" * 45, body)).reshape(15, 3)
    np.testing.assert_array_equal(pts, pos[mask == 0])
