"""GPU parity of the fused substep stages against the oracle.

Integer work (cell binning, sort permutation, cell ranges, marker, tiles) is
checked bit-exact; FP32 fields against the FP64 oracle at 1e-5 relative
(max|a-b|/max|b|), the tolerance the substep's float accumulation order
allows.  The Jacobi sweep is bit-exact against the FP32-faithful oracle
because its summation order is the reference's.
"""
import ctypes

import numpy as np
import pytest

from oracle import flip_oracle as fo
import gfutil as gu

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _keys(pos, dx, shape):
    ci, cj, ck = fo.cell_index(pos, dx, shape)
    return fo.flat_key(ci, cj, ck, shape)


@pytest.mark.parametrize("shape,seed", [((24, 24, 24), 0), ((20, 13, 28), 1)])
def test_sort_bit_exact(shape, seed):
    import torch
    dx = 1.0 / max(shape)
    rng = np.random.default_rng(seed)
    n = 5000
    pos = (rng.random((n, 3)) * np.array(shape) * dx).astype(np.float32)
    # exact cell boundaries, one ulp around them and out-of-domain values
    edge = np.float32(np.arange(0, max(shape) + 1) * dx)
    pos[:len(edge), 0] = edge
    pos[len(edge):2 * len(edge), 1] = np.nextafter(edge, np.float32(-1))
    pos[2 * len(edge):3 * len(edge), 2] = np.nextafter(edge, np.float32(9))
    pos[-4:] = np.float32([[-0.3, 0.2, 0.2], [2.0, 0.1, 0.3], [0.2, -1e-7, 5.0], [0.1, 0.1, 0.1]])
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    s = gu.make_solver(shape, dx, pos, vel)
    na = gu.run_sort(s)
    key = _keys(pos, dx, shape)
    order = fo.stable_order(key)
    sc = s._sc
    live = sc.live(n)
    np.testing.assert_array_equal(live["pid"].cpu().numpy(), order)
    np.testing.assert_array_equal(live["pos"].cpu().numpy(), pos[order])
    np.testing.assert_array_equal(live["vel"].cpu().numpy(), vel[order])
    C = int(np.prod(shape))
    cs = np.concatenate([[0], np.cumsum(np.bincount(key, minlength=C))])
    np.testing.assert_array_equal(sc.cell_start.cpu().numpy(), cs)
    assert int(sc.counts.abs().sum()) == 0, "per-cell counts must return to zero"
    m_want = fo.update_marker(fo.wall_marker(*shape), pos, dx)
    np.testing.assert_array_equal(s.marker.numpy(), m_want)
    # tiles: active = tiles holding a fluid cell, in tile-id order
    tshape = tuple((d + 7) // 8 for d in shape)
    fl = np.zeros(tshape, dtype=bool)
    ii, jj, kk = np.nonzero(m_want == 1)
    fl[ii // 8, jj // 8, kk // 8] = True
    assert na == int(fl.sum())
    tm = sc.tile_map.cpu().numpy()
    want_tm = np.where(fl.ravel(), np.cumsum(fl.ravel()) - 1, -1)
    np.testing.assert_array_equal(tm, want_tm)
    # the unsorted views give the particles back in input order
    np.testing.assert_array_equal(s.pos.numpy(), pos)
    np.testing.assert_array_equal(s.vel.numpy(), vel)
    torch.cuda.synchronize()


@pytest.mark.parametrize("shape", [(16, 16, 16), (20, 13, 28)])
@pytest.mark.parametrize("transfer", ["flip", "apic"])
@pytest.mark.parametrize("staged", [False, True], ids=["one_pass", "staged"])
def test_p2g_fused_matches_oracle(shape, transfer, staged):
    shape, dx, pos, vel = gu.dam_scene(shape=shape, vel_sigma=0.7, seed=5)
    rng = np.random.default_rng(9)
    # a few particles on and beyond the domain boundary: clamped into boundary
    # cells, their in-grid corners still land on faces.  (Offsets are whole
    # fractions of a cell: a particle within rounding of a face gives a
    # weight near the 1e-8 normalisation guard, where FP32 and the FP64
    # oracle may fall on different sides.)
    dom = np.array(shape) * dx
    extra = np.array([[-0.3 * dx, 0.5, 0.5], [dom[0] + 0.4 * dx, 0.3, 0.2],
                      [0.4, -1.7 * dx, 0.6], [0.2, 0.3, dom[2] + 0.1 * dx],
                      [0.0, 0.0, 0.0],
                      [dom[0] + 0.25 * dx, dom[1] + 0.35 * dx, dom[2] + 0.15 * dx]], np.float32)
    pos = np.concatenate([pos, extra])
    vel = np.concatenate([vel, rng.normal(0, 0.7, extra.shape).astype(np.float32)])
    aff = rng.normal(0, 3.0, (len(pos), 3, 3)).astype(np.float32)
    print(f"[p2g] {shape} {transfer} staged={staged}", flush=True)
    s = gu.make_solver(shape, dx, pos, vel, transfer_mode=transfer)
    if transfer == "apic":
        s.set_particles(pos, vel, aff)
    dt = 1.0 / 240
    gu.run_sort(s)
    gu.run_p2g(s, dt, staged=staged)
    m = fo.update_marker(fo.wall_marker(*shape), pos, dx)
    u, v, w, uw, vw, ww = fo.p2g(pos, vel, m, dx, affine=aff if transfer == "apic" else None)
    u, us = fo.normalize(u, uw)
    v, vs = fo.normalize(v, vw)
    w, ws = fo.normalize(w, ww)
    v = fo.add_gravity(v, -9.81, dt)
    u, v, w = fo.enforce_solid_bc(u, v, w, m)
    div = fo.divergence(u, v, w, m, dx, dt, 1.0)
    for name, want in (("u", u), ("v", v), ("w", w), ("us", us), ("vs", vs), ("ws", ws)):
        got = getattr(s, name).numpy()
        assert got.shape == want.shape
        assert gu.rel_err(got, want) < TOL, name
    assert gu.rel_err(s.div.numpy(), div) < TOL
    np.testing.assert_array_equal(s.marker.numpy(), m)


def _prepared(shape=(24, 24, 24), seed=2, sigma=0.5, **kw):
    shape, dx, pos, vel = gu.dam_scene(shape=shape, vel_sigma=sigma, seed=seed)
    s = gu.make_solver(shape, dx, pos, vel, **kw)
    gu.run_sort(s)
    gu.run_p2g(s, 1.0 / 240)
    return s


def test_jacobi_bit_exact_against_fp32_oracle():
    from gpufluid_amd import _lib
    s = _prepared()
    div = s.div.numpy()
    m = s.marker.numpy()
    d = s._descriptor()
    _lib.call("S2.6.1", "gf_scene_jacobi", ctypes.byref(d), s.n_active_tiles, 25, s._stream())
    s._cache.clear()
    want = fo.jacobi(div, m, 25, dtype=np.float32)
    got = s.p.numpy()
    np.testing.assert_array_equal(got, want)


def test_gsrb_bit_exact_against_fp32_oracle():
    """Red-black Gauss-Seidel on the compact tiles equals the FP32 oracle bit
    for bit (in-place sweeps, reference sum order)."""
    from gpufluid_amd import _lib
    s = _prepared(shape=(20, 24, 28), seed=6)
    div = s.div.numpy()
    m = s.marker.numpy()
    d = s._descriptor()
    _lib.call("S2.6.2", "gf_scene_gsrb", ctypes.byref(d), s.n_active_tiles, 17, s._stream())
    s._cache.clear()
    want = fo.gsrb(div, m, 17, dtype=np.float32)
    np.testing.assert_array_equal(s.p.numpy(), want)


def test_gsrb_block_kernel_bit_exact():
    """The S2.6.2 block callable (dense, in place) equals the oracle half sweep."""
    import torch
    from gpufluid_amd import kernels
    rng = np.random.default_rng(13)
    m = fo.wall_marker(10, 12, 9)
    inner = m[1:-1, 1:-1, 1:-1]
    inner[...] = np.where(rng.random(inner.shape) < 0.7, fo.FLUID, fo.AIR)
    div = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    p0 = np.where(m == fo.FLUID, rng.normal(size=m.shape), 0).astype(np.float32)
    p = torch.tensor(p0, device="cuda:0")
    dt = torch.tensor(div, device="cuda:0")
    mt = torch.tensor(m, device="cuda:0")
    want = p0
    for color in (0, 1, 0, 1):
        kernels.k3_gauss_seidel_rb(p, dt, mt, color)
        want = fo.gsrb_half_sweep(want, div, m, color)
    np.testing.assert_array_equal(p.cpu().numpy(), want)


@pytest.fixture(params=[1, 2], ids=["two_kernel", "one_kernel"])
def pcg_variant(request):
    """Both PCG iteration forms (1: direction+SpMV and update kernels,
    2: one Chronopoulos-Gear kernel), forced through the tunable."""
    from gpufluid_amd import autotune
    autotune.set_pcg_variant(request.param)
    yield request.param
    autotune.set_pcg_variant(0)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_pcg_reaches_1e8_and_matches_direct_solve(precision, pcg_variant):
    from gpufluid_amd import _lib
    print(f"[pcg] precision {precision} variant {pcg_variant}", flush=True)
    s = _prepared(shape=(32, 32, 32), pcg_precision=precision)
    if precision == "fp64":
        s._sc.ensure_fp64()
    div = s.div.numpy().astype(np.float64)
    m = s.marker.numpy()
    d = s._descriptor()
    it, res, nrep = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    prec = _lib.GF_PCG_MIXED if precision == "mixed" else _lib.GF_PCG_FP64
    _lib.call("S2.6.3", "gf_scene_pcg_sync", ctypes.byref(d), s.n_active_tiles, prec, 1e-8, 500,
              4, -1, ctypes.byref(it), ctypes.byref(res), ctypes.byref(nrep), None, s._stream())
    s._cache.clear()
    assert 0 < it.value < 500
    assert res.value < 1e-8
    exact = fo.solve_exact(div, m)
    got = s.p.numpy()
    assert gu.rel_err(got, exact) < TOL
    # the FP32-rounded result still satisfies the system to FP32 accuracy
    assert fo.residual_norm(got, div, m) < 1e-5


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_pcg_x_deferral_is_bit_identical(precision):
    """Two-kernel PCG: deferring x += alpha p over up to three update kernels
    (the next flushing one adds the pending terms in iteration order) gives
    the same bits as adding it in every update: pressure, iteration count,
    residual and replacement count, for chunks of several lengths and for
    solves that stop (by tolerance or by the iteration cap) right after an
    update that left one, two or three terms pending, where the direction
    kernel adds them."""
    from gpufluid_amd import _lib, autotune
    s = _prepared(shape=(32, 32, 32), pcg_precision=precision)
    if precision == "fp64":
        s._sc.ensure_fp64()
    d = s._descriptor()
    prec = _lib.GF_PCG_MIXED if precision == "mixed" else _lib.GF_PCG_FP64
    pending_at_stop = set()
    autotune.set_pcg_variant(1)
    try:
        for check_every, tol, cap in [(10, 1e-8, 500), (7, 1e-8, 500), (4, 1e-5, 500),
                                      (10, 1e-6, 500), (3, 1e-7, 500), (10, 1e-8, 37),
                                      (6, 1e-8, 41), (10, 1e-8, 22), (9, 1e-8, 30)]:
            got = {}
            for xe in (1, 2, 3, 4):
                autotune.set_pcg_x_every(xe)
                it, res, nrep = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
                _lib.call("S2.6.3", "gf_scene_pcg_sync", ctypes.byref(d), s.n_active_tiles, prec,
                          tol, cap, check_every, -1, ctypes.byref(it), ctypes.byref(res),
                          ctypes.byref(nrep), None, s._stream())
                s._cache.clear()
                got[xe] = (s.p.numpy().copy(), it.value, res.value, nrep.value)
            print(f"[pcg x] {precision} chunk {check_every} tol {tol} cap {cap}: "
                  + " vs ".join(str(got[xe][1:]) for xe in got), flush=True)
            for xe in (2, 3, 4):
                np.testing.assert_array_equal(got[xe][0], got[1][0])
                assert got[xe][1:] == got[1][1:]
            # terms pending (with x added every 4th update) after the update at
            # which the solve stopped: its position j in its chunk
            j = (got[1][1] - 1) % check_every
            if j % 4 != 3 and j + 1 < check_every:
                pending_at_stop.add(j % 4 + 1)
    finally:
        autotune.set_pcg_x_every(0)
        autotune.set_pcg_variant(0)
    assert pending_at_stop == {1, 2, 3}


def test_pcg_reference_sign_is_negated(pcg_variant):
    from gpufluid_amd import _lib
    s = _prepared(shape=(16, 16, 16))
    d = s._descriptor()
    it, res, nrep = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    _lib.call("S2.6.3", "gf_scene_pcg_sync", ctypes.byref(d), s.n_active_tiles, 0, 1e-8, 300,
              4, 1, ctypes.byref(it), ctypes.byref(res), ctypes.byref(nrep), None, s._stream())
    s._cache.clear()
    x_ref, _ = fo.pcg_reference(s.div.numpy().astype(np.float64), s.marker.numpy(), 300,
                                tol=1e-10, dtype=np.float64)
    assert gu.rel_err(s.p.numpy(), x_ref) < TOL


@pytest.mark.parametrize("transfer,solver", [("flip", "jacobi"), ("flip", "pcg"),
                                             ("apic", "pcg"), ("pic", "jacobi"),
                                             ("flip", "gsrb")])
def test_full_substep_matches_oracle(transfer, solver):
    shape, dx, pos, vel = gu.dam_scene(shape=(20, 24, 16), vel_sigma=0.4, seed=11)
    fb = 0.0 if transfer == "pic" else 0.95
    s = gu.make_solver(shape, dx, pos, vel, transfer_mode=transfer, flip_blend=fb)
    dt = 1.0 / 240
    if solver in ("jacobi", "gsrb"):
        s.step(dt, pressure_iters=30, pressure_solver=solver)
        st, grid = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure=solver,
                              iters=30, transfer=transfer, flip_blend=fb)
    else:
        s.step(dt, pressure_iters=400, pressure_solver="pcg", pressure_tol=1e-8)
        st, grid = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure="exact",
                              transfer=transfer, flip_blend=fb)
    for name in ("u", "v", "w"):
        assert gu.rel_err(getattr(s, name).numpy(), grid[name]) < TOL, name
    assert gu.rel_err(s.p.numpy(), grid["p"]) < TOL
    pp, vv = s.get_particles()
    assert gu.rel_err(vv, st["vel"]) < TOL
    assert gu.rel_err(pp, st["pos"]) < TOL
    if transfer == "apic":
        assert gu.rel_err(s.affine_C.numpy(), st["affine"]) < TOL


def test_multi_step_matches_oracle_flip_pcg():
    """Five coupled substeps stay within tolerance of the FP64 oracle."""
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=4)
    s = gu.make_solver(shape, dx, pos, vel)
    st = gu.oracle_state(shape, dx, pos, vel)
    dt = 1.0 / 240
    for _ in range(5):
        s.step(dt, pressure_iters=300, pressure_solver="pcg", pressure_tol=1e-9)
        st, grid = fo.substep(st, dt, pressure="exact")
    pp, vv = s.get_particles()
    assert gu.rel_err(vv, st["vel"]) < 1e-4
    assert gu.rel_err(pp, st["pos"]) < 1e-5


def test_empty_scene_steps():
    from gpufluid_amd import FlipSolver3D
    s = FlipSolver3D(16, 16, 16)
    s.step(0.005, pressure_solver="pcg")
    s.step(0.005)
    v = s.v.numpy()
    assert np.isfinite(v).all()
    assert s.last_pressure_iters == 80  # Jacobi reports its sweep count


def test_has_diverged_detects_nan():
    shape, dx, pos, vel = gu.dam_scene(n=16)
    pos[3, 1] = np.nan
    s = gu.make_solver(shape, dx, pos, vel)
    assert s.has_diverged()
