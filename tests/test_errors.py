"""Error conventions of the boundary, on the host (no device calls).

A failing C entry point surfaces as ``BlockError(<block id>, <message>)``
whose text is ``"[<block id>] <entry point> failed (<code>): <C message>"``
(reference src/gpufluid/blocks/__init__.py:129-150 renders "[id] message").
The entry points used here validate their arguments before touching HIP, so
they run without a GPU.
"""
from __future__ import annotations

import ctypes

import pytest

from gpufluid_amd import _lib
from gpufluid_amd.blocks import BlockError


def _lib_or_skip():
    from gpufluid_amd import build
    if not build.LIB.exists():
        build.build()
    return _lib.lib()


def test_bad_grid_raises_block_error_with_id_and_c_message():
    _lib_or_skip()
    out = _lib.GfSizes()
    with pytest.raises(BlockError) as ei:
        _lib.call("F3.2", "gf_scene_sizes", 0, 8, 8, ctypes.byref(out))
    e = ei.value
    assert e.block_id == "F3.2"
    text = str(e)
    assert text.startswith("[F3.2] ")
    assert "gf_scene_sizes failed (-1)" in text
    assert "grid 0x8x8" in text          # the C side's own message
    assert _lib.last_error() == "gf_scene_sizes: grid 0x8x8"


def test_oversized_grid_is_a_range_error():
    _lib_or_skip()
    out = _lib.GfSizes()
    with pytest.raises(BlockError) as ei:
        _lib.call("F3.2", "gf_scene_sizes", 2048, 2048, 1024, ctypes.byref(out))
    assert f"failed ({_lib.GF_ERANGE})" in str(ei.value)
    assert "exceed int32 indexing" in str(ei.value)


def test_null_output_and_unknown_tunable():
    _lib_or_skip()
    with pytest.raises(BlockError, match=r"^\[F3.2\] gf_scene_sizes failed \(-1\): "
                                         r"gf_scene_sizes: null output$"):
        _lib.call("F3.2", "gf_scene_sizes", 8, 8, 8, None)
    with pytest.raises(BlockError, match="unknown key 'no_such_knob'"):
        _lib.call("F3.2", "gf_tune_set", b"no_such_knob", 1)


def test_success_clears_nothing_and_returns_sizes():
    _lib_or_skip()
    out = _lib.GfSizes()
    _lib.call("F3.2", "gf_scene_sizes", 17, 8, 9, ctypes.byref(out))
    assert out.cells == 17 * 8 * 9
    assert out.ntiles == 3 * 1 * 2
    assert out.compact == out.ntiles * 512


def test_scene_entry_point_rejects_null_descriptor():
    """Every scene entry point checks its descriptor before any launch."""
    lib = _lib_or_skip()
    rc = lib.gf_scene_pcg_flags(None, ctypes.c_float(1e-4), None, None, None)
    assert rc == _lib.GF_EINVAL
    assert "gf_scene_pcg_flags" in _lib.last_error()
    with pytest.raises(BlockError) as ei:
        _lib.call("S2.6.3", "gf_scene_pcg_flags", None, 1e-4, None, None, None)
    assert ei.value.block_id == "S2.6.3"


def test_noalloc_pcg_limit_is_checked_before_any_launch():
    """The no-allocation PCG refuses more active tiles than its static work
    array holds with GF_ERANGE, from the descriptor's sizes alone (the
    buffer addresses below are placeholders that are never dereferenced)."""
    _lib_or_skip()
    d = _lib.GfScene()
    d.nx = d.ny = d.nz = 136
    d.tnx = d.tny = d.tnz = 17
    d.dx = 1.0 / 136
    for name in ("flags", "nbr", "b", "pc", "scalars", "n_active"):
        setattr(d, name, 4096)
    with pytest.raises(BlockError) as ei:
        _lib.call("S2.6.3", "gf_scene_pcg_enqueue_noalloc", ctypes.byref(d), 2049, 0, 1e-8, 10,
                  10, -1, 0, 1, None)
    assert f"failed ({_lib.GF_ERANGE})" in str(ei.value)
    assert "2049 active tiles exceed the 2048" in str(ei.value)
    with pytest.raises(BlockError, match="4913 grid tiles exceed the 2048"):
        _lib.call("S2.6.3", "gf_scene_pcg_enqueue_noalloc", ctypes.byref(d),
                  _lib.GF_NACT_DEVICE, 0, 1e-8, 10, 10, -1, 0, 1, None)
