"""F3.7 sub-dense box: bookkeeping (reference tests/test_b7_alt_2_sub_dense_storage.py,
test_b7_alt_4_particle_bounds.py, test_b7_alt_7_scattered_warning.py restated on
SubDenseBox) and the oracle's box viscosity stage."""
from __future__ import annotations

import numpy as np
import pytest

from gpufluid_amd.subdense import SubDenseBox
from oracle import flip_oracle as fo


def _walls(n):
    m = np.zeros((n, n, n), dtype=np.int32)
    m[0], m[-1] = 2, 2
    m[:, 0], m[:, -1] = 2, 2
    m[:, :, 0], m[:, :, -1] = 2, 2
    return m


def test_default_box_is_the_grid_and_never_rebuilds():
    b = SubDenseBox(32, 24, 16, enabled=False, rebuild_every=8, dilation=4)
    assert b.offset == (0, 0, 0) and b.shape == (32, 24, 16) and not b.active
    assert b.face_slices(0) == (slice(0, 33), slice(0, 24), slice(0, 16))
    assert b.should_rebuild(0, _walls(32)[:, :24, :16]) is False
    assert b.should_rebuild(10_000, None) is False


def test_bbox_empty_tile_snapped_dilated_and_clamped():
    b = SubDenseBox(32, 32, 32, True, 8, 0)
    m = _walls(32)
    assert b.bbox(m) is None
    m[10, 18, 27] = 1                       # one cell: its 8^3 tile
    assert b.bbox(m) == ((8, 16, 24), (16, 24, 32))
    b4 = SubDenseBox(32, 32, 32, True, 8, 4)
    m = np.zeros((32, 32, 32), np.int32)
    m[16:24, 16:24, 16:24] = 1
    assert b4.bbox(m) == ((12, 12, 12), (28, 28, 28))
    m[:] = 0
    m[0:8, 0:8, 0:8] = 1                    # corner: clamped at 0
    assert b4.bbox(m) == ((0, 0, 0), (12, 12, 12))
    m[:] = 0
    m[5, 5, 5] = 1
    m[20, 20, 20] = 1                       # two tiles
    assert b.bbox(m) == ((0, 0, 0), (24, 24, 24))


def test_bbox_cell_precise_when_extents_are_not_multiples_of_8():
    b = SubDenseBox(30, 20, 10, True, 8, 2)
    m = np.zeros((30, 20, 10), np.int32)
    m[11, 3, 9] = 1
    m[13, 5, 4] = 1
    assert b.bbox(m) == ((9, 1, 2), (16, 8, 10))


def test_rebuild_sets_box_and_face_slices():
    b = SubDenseBox(32, 32, 32, True, 8, 2)
    m = np.zeros((32, 32, 32), np.int32)
    m[16:24, 16:24, 16:24] = 1
    lo, hi = b.bbox(m)
    assert (lo, hi) == ((14, 14, 14), (26, 26, 26))
    b.rebuild(lo, hi)
    assert b.offset == (14, 14, 14) and b.shape == (12, 12, 12) and b.active
    u = np.zeros((33, 32, 32))
    v = np.zeros((32, 33, 32))
    w = np.zeros((32, 32, 33))
    assert u[b.face_slices(0)].shape == (13, 12, 12)
    assert v[b.face_slices(1)].shape == (12, 13, 12)
    assert w[b.face_slices(2)].shape == (12, 12, 13)
    assert np.zeros((32, 32, 32))[b.face_slices(None)].shape == (12, 12, 12)
    with pytest.raises(ValueError):
        b.rebuild((4, 4, 4), (4, 8, 8))
    with pytest.raises(ValueError):
        b.rebuild((28, 0, 0), (36, 8, 8))


def test_periodic_and_proximity_triggers():
    b = SubDenseBox(32, 32, 32, True, 4, 8)
    m = _walls(32)
    m[12:20, 12:20, 12:20] = 1
    assert b.should_rebuild(0, m) is True        # never rebuilt
    b.rebuild(*b.bbox(m))
    b.last_rebuild_frame = 0
    assert b.should_rebuild(1, m) is False
    assert b.should_rebuild(3, m) is False
    assert b.should_rebuild(4, m) is True        # period
    assert b.should_rebuild(100, m) is True
    p = SubDenseBox(64, 64, 64, True, 1000, 4)
    m = _walls(64)
    m[16:24, 16:24, 16:24] = 1
    p.rebuild(*p.bbox(m))
    p.last_rebuild_frame = 0
    m2 = _walls(64)
    m2[18:26, 18:26, 18:26] = 1                  # raw box 16..32 vs box 12..28: too close
    assert p.should_rebuild(1, m2) is True
    assert p.should_rebuild(1, m) is False


def test_pos_to_cell_and_outside_count_warn_once(capsys):
    dx = 1.0 / 32
    b = SubDenseBox(32, 32, 32, True, 8, 0)
    (li, lj, lk), inside = b.pos_to_cell((10.5 * dx, 14.2 * dx, 21.9 * dx), dx)
    assert (li, lj, lk) == (10, 14, 21) and inside is True
    assert b.pos_to_cell((-0.5 * dx, 0.0, 0.0), dx)[1] is False
    b.rebuild((14, 14, 14), (26, 26, 26))
    assert b.pos_to_cell((16.5 * dx, 16.5 * dx, 16.5 * dx), dx) == ((2, 2, 2), True)
    assert b.pos_to_cell((28.5 * dx, 16.5 * dx, 16.5 * dx), dx) == ((14, 2, 2), False)
    c = SubDenseBox(32, 32, 32, True, 8, 0)
    c.rebuild((8, 8, 8), (12, 12, 12))
    pos = np.array([[9.5, 9.5, 9.5], [10.5, 10.5, 10.5], [1.5, 1.5, 1.5], [20.5, 9.5, 9.5],
                    [9.5, 30.5, 9.5], [9.5, 9.5, 12.0]], np.float32) * np.float32(dx)
    assert c.count_outside(pos, dx, frame_idx=5) == 4
    err = capsys.readouterr().err
    assert "outside the sub-dense bbox" in err and "frame 5" in err and "sub_dilation" in err
    assert c.count_outside(pos, dx, frame_idx=6) == 4
    assert capsys.readouterr().err == ""
    assert c.count_outside(pos[:2], dx, 7) == 0
    off = SubDenseBox(32, 32, 32, False, 8, 0)
    assert off.count_outside(pos, dx, 0) == 0


def test_scattered_topology_warning(capsys):
    b = SubDenseBox(64, 64, 64, True, 8, 2)
    b.rebuild((16, 16, 16), (48, 48, 48))
    assert "scattered topology" not in capsys.readouterr().err
    s = SubDenseBox(32, 32, 32, True, 8, 4)
    s.rebuild((0, 0, 0), (30, 30, 30))             # 82.4 %
    err = capsys.readouterr().err
    assert "scattered topology" in err and "82.4%" in err and "sub_dilation`=4" in err
    s.rebuild((1, 1, 1), (31, 31, 31))
    assert capsys.readouterr().err == ""          # once
    t = SubDenseBox(10, 10, 10, True, 8, 0)
    t.rebuild((0, 0, 0), (8, 10, 10))              # exactly 0.8: strict
    assert "scattered topology" not in capsys.readouterr().err
    t2 = SubDenseBox(10, 10, 10, True, 8, 0)
    t2.rebuild((0, 0, 0), (9, 10, 10))
    assert "scattered topology" in capsys.readouterr().err


def _fields(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    f = [rng.normal(size=s).astype(np.float32)
         for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    return f, [a.copy() * np.float32(0.5) for a in f]


def test_oracle_box_viscosity_whole_grid_equals_dense():
    """The box stage over the whole grid is the dense stage (iters >= 1)."""
    shape = (9, 7, 6)
    m = np.ones(shape, np.int32)
    (u, v, w), (us, vs, ws) = _fields(shape, 1)
    r = np.float32(0.3)
    full = fo.viscosity_stage(u, v, w, us, vs, ws, m, r, 5)
    boxed = fo.viscosity_stage(u, v, w, us, vs, ws, m, r, 5, box=((0, 0, 0), shape))
    for a, b in zip(full, boxed):
        np.testing.assert_array_equal(a, b)


def test_oracle_box_viscosity_domain_of_dependence():
    """Inside the box, faces at least `iters` cells from the box edge equal
    the dense result (the sweep reads one neighbour per direction); faces
    outside keep their pre-sweep values; a constant field stays constant."""
    shape = (20, 18, 16)
    m = np.ones(shape, np.int32)
    (u, v, w), (us, vs, ws) = _fields(shape, 2)
    r, iters = np.float32(0.4), 3
    lo, sz = (5, 4, 3), (11, 10, 9)
    full = fo.viscosity_stage(u, v, w, us, vs, ws, m, r, iters)
    boxed = fo.viscosity_stage(u, v, w, us, vs, ws, m, r, iters, box=(lo, sz))
    bc = fo.enforce_solid_bc(u.copy(), v.copy(), w.copy(), m)   # the stage ends with the BC
    for c, (a, b, orig) in enumerate(zip(full[:3], boxed[:3], (u, v, w))):
        inner = tuple(slice(lo[k] + iters, lo[k] + sz[k] + (k == c) - iters) for k in range(3))
        np.testing.assert_array_equal(a[inner], b[inner])
        mask = np.ones(orig.shape, bool)
        mask[fo.box_face_slices((lo, sz), c)] = False
        np.testing.assert_array_equal(b[mask], bc[c][mask])
        np.testing.assert_array_equal(boxed[3 + c][mask], orig[mask])
        assert not np.array_equal(a, b)     # the box edge changes the stencil
    k = np.float32(0.75)
    const = [np.full(x.shape, k, np.float32) for x in (u, v, w)]
    out = fo.viscosity_stage(*const, *const, m, r, iters, box=(lo, sz))
    for c in range(3):
        np.testing.assert_allclose(out[c][fo.box_face_slices((lo, sz), c)], k, rtol=0,
                                   atol=2e-7)
