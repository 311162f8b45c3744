"""F3.7 sub-dense box on the HIP path: the box viscosity sweep bit for bit
against the oracle's sweep on the box-shaped arrays (the reference runs
k3_jacobi_visc on its box arrays, solver3d.py:178-200, 3530-3541), a whole
sub-dense substep against the oracle, and the box views / rebuild policy
through FlipSolver3D (solver3d.py:3669-3906, 4050-4072)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import gfutil as gu
from oracle import flip_oracle as fo

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _log(msg):
    print(f"[subdense] {msg}", flush=True)


@pytest.mark.parametrize("box", [((3, 5, 2), (9, 7, 6)), ((0, 0, 0), (20, 24, 16)),
                                 ((12, 0, 10), (8, 24, 6)), ((4, 4, 4), (1, 1, 1))])
def test_visc3_box_kernel_bit_exact(box):
    from gpufluid_amd import _lib
    nx, ny, nz = 20, 24, 16
    rng = np.random.default_rng(sum(box[0]) + 7)
    shp = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]
    old = [rng.normal(size=x).astype(np.float32) for x in shp]
    inp = [rng.normal(size=x).astype(np.float32) for x in shp]
    r = 0.31
    _log(f"box sweep {box}")
    t_old = [torch.tensor(a, device="cuda:0") for a in old]
    t_in = [torch.tensor(a, device="cuda:0") for a in inp]
    t_out = [torch.full_like(t, 123.0) for t in t_in]
    _lib.call("S2.13", "gf_k3_jacobi_visc3_box", *(_lib.ptr(t) for t in t_old + t_in + t_out),
              nx, ny, nz, *box[0], *box[1], r, _lib.stream_handle(torch.device("cuda:0")))
    for c in range(3):
        got = t_out[c].cpu().numpy()
        sl = fo.box_face_slices(box, c)
        np.testing.assert_array_equal(got[sl], fo.visc_sweep(old[c][sl], inp[c][sl], r))
        outside = np.ones(got.shape, bool)
        outside[sl] = False
        assert np.all(got[outside] == 123.0)


def test_visc3_box_rejects_a_box_outside_the_grid():
    from gpufluid_amd import _lib
    from gpufluid_amd.blocks import BlockError
    t = [torch.zeros(4, device="cuda:0") for _ in range(9)]
    with pytest.raises(BlockError, match="not inside"):
        _lib.call("S2.13", "gf_k3_jacobi_visc3_box", *(_lib.ptr(x) for x in t), 8, 8, 8,
                  4, 0, 0, 5, 8, 8, 0.1, _lib.stream_handle(torch.device("cuda:0")))


def _blob(n=32, seed=5, ylo=0.2, yhi=0.5):
    from gpufluid_amd.seeding import fluid_cells, seed_cells
    dx = 1.0 / n
    cells = fluid_cells((n, n, n), dx, lambda x, y, z: (x > 0.3) & (x < 0.6) & (y > ylo) &
                        (y < yhi) & (z > 0.35) & (z < 0.65))
    rng = np.random.default_rng(seed)
    pos = seed_cells(cells, dx, rng)
    vel = rng.normal(0.0, 0.4, pos.shape).astype(np.float32)
    return (n, n, n), dx, pos, vel


@pytest.mark.parametrize("iters", [6, 0])
def test_subdense_substep_matches_oracle(iters):
    """A substep with the box around the blob, PCG to 1e-8 and viscosity:
    grid fields (box views) and particles against the oracle substep whose
    viscosity runs on the same box."""
    shape, dx, pos, vel = _blob()
    visc = 0.05
    s = gu.make_solver(shape, dx, pos, vel, viscosity=visc, viscosity_iters=iters,
                       enable_sub_dense=True, sub_dilation=4)
    m = fo.update_marker(s._wall_marker.copy(), pos, dx)
    lo, hi = s._compute_active_bbox(m)
    s._rebuild_sub_dense(lo, hi)
    box = (s._sub_offset, s._sub_shape)
    _log(f"substep in box {box}, {iters} viscosity sweeps")
    assert s._sub_offset != (0, 0, 0) and s._sub_shape != shape
    dt = 1.0 / 240
    s.step(dt, pressure_iters=400, pressure_solver="pcg", pressure_tol=1e-8)
    st, grid = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure="exact",
                          viscosity=visc, viscosity_iters=iters, visc_box=box)
    for c, name in enumerate(("u", "v", "w")):
        got = getattr(s, name).numpy()
        ref = grid[name][fo.box_face_slices(box, c)]
        assert got.shape == ref.shape
        assert gu.rel_err(got, ref) < TOL, name
    assert s.p.shape == s._sub_shape and s.div.shape == s._sub_shape
    assert gu.rel_err(s.p.numpy(), grid["p"][fo.box_face_slices(box, None)]) < TOL
    pp, vv = s.get_particles()
    assert gu.rel_err(vv, st["vel"]) < TOL
    assert gu.rel_err(pp, st["pos"]) < TOL


def test_subdense_without_viscosity_is_bit_identical_to_dense():
    """Every stage but the viscosity sweep runs on the whole grid, so with
    no viscosity the box changes nothing; with viscosity the box edge cuts
    the stencil far from the fluid, and the result moves by FP32 noise only
    (the reference's own bound for its sub-dense path, tests/
    test_b7_alt_3_jacobi_dense_flip.py:165-187, is 1e-4)."""
    shape, dx, pos, vel = _blob(seed=9)
    dt = 1.0 / 240
    out = {}
    for visc in (0.0, 0.05):
        runs = []
        for sub in (False, True):
            s = gu.make_solver(shape, dx, pos, vel, viscosity=visc, viscosity_iters=8,
                               enable_sub_dense=sub, sub_dilation=4)
            if sub:
                s._rebuild_sub_dense(*s._compute_active_bbox(
                    fo.update_marker(s._wall_marker.copy(), pos, dx)))
            for _ in range(3):
                s.step(dt, pressure_iters=300, pressure_solver="pcg", pressure_tol=1e-8)
            runs.append(s.get_particles())
        out[visc] = runs
        _log(f"viscosity {visc}: sub-dense vs dense done")
    (pa, va), (pb, vb) = out[0.0]
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)
    (pa, va), (pb, vb) = out[0.05]
    assert gu.rel_err(pb, pa) < 1e-4 and gu.rel_err(vb, va) < 1e-4


def test_prepare_frame_rebuilds_the_box_and_counts_no_escapes():
    # a blob clear of the walls: its box is not clamped at the domain edge,
    # where the proximity rule (solver3d.py:3803-3818) fires on every frame
    shape, dx, pos, vel = _blob(seed=3, ylo=0.4, yhi=0.7)
    s = gu.make_solver(shape, dx, pos, vel, enable_sub_dense=True, sub_dilation=4,
                       sub_rebuild_every=8)
    assert s._should_rebuild_sub_dense(0) is True
    s.prepare_frame(0, 1.0 / 60)                     # no step yet: walls only, no box
    assert s._sub_offset == (0, 0, 0) and s._last_sub_rebuild_frame == -1
    s.step(1.0 / 240, pressure_iters=200, pressure_solver="pcg")
    s.prepare_frame(1, 1.0 / 60)
    _log(f"box after prepare_frame: {s._sub_offset} {s._sub_shape}")
    assert s._last_sub_rebuild_frame == 1
    assert s._sub_offset != (0, 0, 0)
    ox, oy, oz = s._sub_offset
    sx, sy, sz = s._sub_shape
    assert s.u.shape == (sx + 1, sy, sz) and s.v.shape == (sx, sy + 1, sz)
    assert s.w.shape == (sx, sy, sz + 1)
    np.testing.assert_array_equal(s.u.numpy(), s._f["u"][ox:ox + sx + 1, oy:oy + sy,
                                                           oz:oz + sz].cpu().numpy())
    assert s._check_particles_in_sub_bbox(1) == 0
    assert s._should_rebuild_sub_dense(2) is False
    assert s._should_rebuild_sub_dense(9) is True
    # clamped at y = 0: the undilated box starts at the box edge, so every
    # frame asks for a rebuild, as the reference's rule does
    shape, dx, pos, vel = _blob(seed=3)
    c = gu.make_solver(shape, dx, pos, vel, enable_sub_dense=True, sub_dilation=4)
    c.step(1.0 / 240, pressure_iters=200, pressure_solver="pcg")
    c.prepare_frame(1, 1.0 / 60)
    assert c._sub_offset[1] == 0 and c._should_rebuild_sub_dense(2) is True


def test_step_refuses_an_unknown_solver_while_the_box_is_active():
    shape, dx, pos, vel = _blob(seed=1)
    s = gu.make_solver(shape, dx, pos, vel, enable_sub_dense=True, sub_dilation=0)
    s._rebuild_sub_dense((0, 0, 8), (8, 8, 16))
    with pytest.raises(NotImplementedError, match="pressure_solver"):
        s.step(0.01, pressure_iters=1, pressure_solver="multigrid_v_cycle")
