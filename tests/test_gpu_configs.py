"""The benchmark configurations run through the HIP path at their full sizes.

* C1 (32^3 FLIP dam break, seed 0, dt 1/240, FLIP 0.95, Jacobi-PCG to 1e-6,
  100 substeps): the first substep against the FP64 oracle, then every
  substep of the 100 against one FP64 oracle substep started from the GPU's
  own state, plus the reference's self-consistency checks of a long run
  (tests/test_f3_solver.py:26-51: finite, inside the domain, the centroid
  falls).
* C2 (128^3 APIC dam break, 3.83 M particles, FP32 PCG + FP64 refinement to
  1e-8) and C3 (256^3 pool + sphere, 40.2 M particles, mixed PCG to 1e-8 with
  a 500-iteration cap): on the second substep (the first starts at rest or
  with two uniform velocities, so its P2G is nearly trivial),
    - the sort permutation and the marker bit for bit against
      np.argsort(kind="stable") and set algebra on the host,
    - P2G (us, vs, ws) against the FP64 oracle: the whole grid at C2, a
      32^3 sub-box at C3 (particles of the box and its one-cell halo,
      shifted by a multiple of 8 cells, which is exact in FP32),
    - the FP64 relative residual of the solver's FP64 iterate recomputed on
      the host from the assembled operator, against the reported one and
      the tolerance,
    - the projected velocity's divergence on fluid cells,
    - G2P + advection of a particle sample against the FP64 oracle given the
      GPU's own grid fields,
    - particle count and the clamp box.
* C5 (256 scenes of edge 24..96): a stratified sample of 16 scenes, every edge
  size including 96, packed against the same scenes run alone: the first
  substep's marker and sort order bit for bit, the particles bit for bit where
  the first solve did no iteration, then agreement to the PCG's rounding.

Tolerances: integer work bit-exact; FP32 fields 1e-5 relative to the FP64
oracle (max-abs convention).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import gfutil as gu
from oracle import flip_oracle as fo
from test_gpu_batch import check_against_alone, packed_first_substep

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _log(msg):
    print(f"[configs] {msg}", flush=True)


# --------------------------------------------------------------------- C1
def test_c1_first_substep_matches_oracle():
    from gpufluid_amd import FlipSolver3D, seeding
    sc = seeding.config_c1()
    p = sc.params
    _log(f"C1 {sc.shape}, {sc.n_particles} particles, 1 substep vs the FP64 oracle")
    s = FlipSolver3D(*sc.shape, dx=sc.dx, flip_blend=p["flip_blend"], device="cuda:0")
    s.set_particles(sc.pos, sc.vel)
    s.step(p["dt"], pressure_iters=1000, pressure_solver="pcg", pressure_tol=p["tol"])
    assert s.last_pressure_residual < p["tol"]
    st, grid = fo.substep(gu.oracle_state(sc.shape, sc.dx, sc.pos, sc.vel), p["dt"],
                          pressure="exact", flip_blend=p["flip_blend"])
    # one velocity scale for the three components (w is zero up to rounding
    # in this scene, so its own maximum is no scale)
    scale = max(float(np.abs(grid[k]).max()) for k in ("u", "v", "w"))
    for name in ("u", "v", "w"):
        e = float(np.abs(getattr(s, name).numpy() - grid[name]).max()) / scale
        assert e < TOL, (name, e)
    pp, vv = s.get_particles()
    ev, ep = gu.rel_err(vv, st["vel"]), gu.rel_err(pp, st["pos"])
    _log(f"C1 substep 1: PCG {s.last_pressure_iters} it, residual "
         f"{s.last_pressure_residual:.2e}; rel err vel {ev:.2e} pos {ep:.2e}")
    assert ev < TOL and ep < TOL


@pytest.mark.slow
def test_c1_hundred_substeps_each_matches_an_oracle_substep():
    from gpufluid_amd import FlipSolver3D, seeding
    sc = seeding.config_c1()
    p = sc.params
    _log(f"C1 {sc.n_particles} particles, {p['substeps']} substeps, each vs an oracle substep")
    s = FlipSolver3D(*sc.shape, dx=sc.dx, flip_blend=p["flip_blend"], device="cuda:0")
    s.set_particles(sc.pos, sc.vel)
    walls = fo.wall_marker(*sc.shape)
    dom = np.array([fo.domain(n, sc.dx) for n in sc.shape])
    eps = np.float32(sc.dx) * np.float32(1.001)
    pos, vel = sc.pos, sc.vel
    y0 = float(pos[:, 1].mean())
    worst = (0.0, 0.0)
    for k in range(p["substeps"]):
        s.step(p["dt"], pressure_iters=1000, pressure_solver="pcg", pressure_tol=p["tol"])
        assert s.last_pressure_residual < p["tol"], (k, s.last_pressure_residual)
        npos, nvel = s.get_particles()
        st, _ = fo.substep({"pos": pos, "vel": vel, "marker": walls, "dx": sc.dx}, p["dt"],
                           pressure="exact", flip_blend=p["flip_blend"])
        ev, ep = gu.rel_err(nvel, st["vel"]), gu.rel_err(npos, st["pos"])
        worst = (max(worst[0], ev), max(worst[1], ep))
        assert ev < TOL and ep < TOL, (k, ev, ep)
        assert np.isfinite(npos).all() and np.isfinite(nvel).all()
        assert (npos >= eps).all() and (npos <= dom - eps).all()
        pos, vel = npos, nvel
        if k % 20 == 19:
            _log(f"C1 substep {k + 1}: PCG {s.last_pressure_iters} it, worst rel err "
                 f"vel {worst[0]:.2e} pos {worst[1]:.2e}")
    assert len(pos) == sc.n_particles
    assert float(pos[:, 1].mean()) < y0          # the column collapses under gravity
    assert not s.has_diverged()


# ------------------------------------------------------------- C2 and C3
def _second_substep(sc, transfer, tol, max_iter, dt, flip_blend=0.95):
    """Run two substeps; record the second's sort input/output and the
    particle state its P2G read.  Returns the solver and the record."""
    from gpufluid_amd import FlipSolver3D
    s = FlipSolver3D(*sc.shape, dx=sc.dx, flip_blend=flip_blend, transfer_mode=transfer,
                     device="cuda:0")
    s.set_particles(sc.pos, sc.vel)
    s.step(dt, pressure_iters=max_iter, pressure_solver="pcg", pressure_tol=tol)
    n = s.n_particles
    live = s._sc.live(n)
    rec = {"in_pos": live["pos"].cpu().numpy(), "in_pid": live["pid"].cpu().numpy()}

    def hook(stage):
        if stage == "sorted":
            lv = s._sc.live(n)
            rec["pid"] = lv["pid"].cpu().numpy()
            rec["pos"] = lv["pos"].cpu().numpy()
            rec["vel"] = lv["vel"].cpu().numpy()
            if transfer == "apic":
                rec["affine"] = lv["affine"].cpu().numpy().reshape(n, 3, 3)
            rec["marker"] = s.marker.numpy()
            rec["n_active"] = s._sc.n_active_host

    s._step_impl(dt, max_iter, "pcg", tol, hook=hook)
    torch.cuda.synchronize()
    return s, rec


def _check_sort(s, rec):
    key = fo.flat_key(*fo.cell_index(rec["in_pos"], s.dx, (s.nx, s.ny, s.nz)),
                      (s.nx, s.ny, s.nz))
    order = fo.stable_order(key)
    np.testing.assert_array_equal(rec["pid"], rec["in_pid"][order])
    np.testing.assert_array_equal(rec["pos"], rec["in_pos"][order])
    m_want = fo.update_marker(fo.wall_marker(s.nx, s.ny, s.nz), rec["in_pos"], s.dx)
    np.testing.assert_array_equal(rec["marker"], m_want)
    tiles = np.zeros(tuple(-(-d // 8) for d in (s.nx, s.ny, s.nz)), bool)
    ii, jj, kk = np.nonzero(m_want == 1)
    tiles[ii // 8, jj // 8, kk // 8] = True
    assert rec["n_active"] == int(tiles.sum())
    return m_want


def _check_pressure(s, marker, tol, max_iter):
    """Host FP64 residual of the solver's FP64 iterate x64 + x32 against the
    reported one and the tolerance; the FP32 pressure handed to G2P must be
    that iterate rounded.  Returns the right-hand side and the residual of the
    rounded FP32 pressure (the floor any FP32 projection sits on)."""
    sc = s._sc
    x = gu.compact_to_dense_host(s, sc.x64) + gu.compact_to_dense_host(s, sc.x32).astype(np.float64)
    div = gu.compact_to_dense_host(s, sc.b)
    res = fo.residual_norm(x, div, marker)
    res32 = fo.residual_norm(x.astype(np.float32), div, marker)
    rep = s.last_pressure_residual
    _log(f"PCG {s.last_pressure_iters} it: reported residual {rep:.3e}, host FP64 {res:.3e}, "
         f"FP32-rounded pressure {res32:.3e}")
    assert abs(res - rep) <= 0.02 * rep + 1e-12
    if s.last_pressure_iters < max_iter:       # stopped on the tolerance
        assert res < tol * 1.02
    np.testing.assert_array_equal(s.p.numpy(), x.astype(np.float32))
    return div, res32


def _check_projection(s, marker, div_before, dt, rho, res32):
    """After the gradient subtract, the face sums of a fluid cell are
    dt/(rho dx) (b + A p): the relative divergence left equals the residual
    of the FP32 pressure used, up to FP32 rounding of the faces."""
    d = fo.divergence(s.u.numpy().astype(np.float64), s.v.numpy().astype(np.float64),
                      s.w.numpy().astype(np.float64), marker, 1.0, 1.0, 1.0)
    fl = marker == 1
    before = div_before[fl].astype(np.float64) * dt / (rho * float(np.float32(s.dx)))
    r = float(np.linalg.norm(d[fl]) / np.linalg.norm(before))
    _log(f"projected face-sum norm / initial: {r:.3e} (FP32 pressure residual {res32:.3e})")
    assert r < 2.0 * res32 + 1e-5


def _check_g2p(s, rec, transfer, dt, flip_blend, nsample, seed):
    n = s.n_particles
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, size=min(n, nsample), replace=False))
    shape = (s.nx, s.ny, s.nz)
    dom = [fo.domain(k, s.dx) for k in shape]
    f = {k: getattr(s, k).numpy().astype(np.float64) for k in ("u", "v", "w", "us", "vs", "ws")}
    pos, vel = rec["pos"][idx], rec["vel"][idx]
    if transfer == "apic":
        npos, nvel, C = fo.g2p_apic_advect(pos, f["u"], f["v"], f["w"], s.dx, dt, dom)
    else:
        npos, nvel = fo.g2p_advect(pos, vel, f["u"], f["v"], f["w"], f["us"], f["vs"], f["ws"],
                                   s.dx, dt, flip_blend, dom)
    live = s._sc.live(n)
    gpos = live["pos"].cpu().numpy()[idx]
    gvel = live["vel"].cpu().numpy()[idx]
    ev, ep = gu.rel_err(gvel, nvel), gu.rel_err(gpos, npos)
    _log(f"G2P sample of {len(idx)}: rel err vel {ev:.2e} pos {ep:.2e}")
    assert ev < TOL and ep < TOL
    if transfer == "apic":
        gC = live["affine"].cpu().numpy().reshape(n, 3, 3)[idx]
        assert gu.rel_err(gC, C) < TOL
    eps = np.float32(s.dx) * np.float32(1.001)
    allp = live["pos"].cpu().numpy()
    assert (allp >= eps).all() and (allp <= np.array(dom) - eps).all()
    assert np.isfinite(live["vel"].cpu().numpy()).all()


def _check_p2g_box(s, rec, lo, B):
    """Oracle P2G + normalize of the particles in cells [lo-1, lo+B+1) per
    axis, shifted by (lo-8) cells, against us/vs/ws on the faces that only
    those particles reach."""
    dx = s.dx
    shape = np.array((s.nx, s.ny, s.nz))
    lo = np.asarray(lo)
    o = lo - 8                                   # multiple of 8 cells: exact shift
    c = np.stack(fo.cell_index(rec["pos"], dx, tuple(shape)), axis=1)
    sel = np.all((c >= lo - 1) & (c < lo + B + 1), axis=1)
    pos = (rec["pos"][sel] - (o * np.float32(dx)).astype(np.float32)).astype(np.float32)
    assert np.array_equal(pos + (o * np.float32(dx)).astype(np.float32), rec["pos"][sel])
    aff = rec.get("affine")
    sub = (B + 16,) * 3
    m = np.zeros(sub, np.int32)
    u, v, w, uw, vw, ww = fo.p2g(pos, rec["vel"][sel], m, dx,
                                 affine=None if aff is None else aff[sel])
    errs = []
    for comp, (f, wt, key) in enumerate(((u, uw, "us"), (v, vw, "vs"), (w, ww, "ws"))):
        want, _ = fo.normalize(f, wt)
        sl_loc = [slice(8, 8 + B)] * 3
        sl_loc[comp] = slice(8, 8 + B + 1)
        sl_glob = [slice(lo[a], lo[a] + B) for a in range(3)]
        sl_glob[comp] = slice(lo[comp], lo[comp] + B + 1)
        got = getattr(s, key).numpy()[tuple(sl_glob)]
        errs.append(gu.rel_err(got, want[tuple(sl_loc)]))
    _log(f"P2G sub-box {tuple(lo)}+{B}: {int(sel.sum())} particles, rel err {errs}")
    assert max(errs) < TOL


@pytest.mark.slow
def test_c2_full_size_second_substep():
    from gpufluid_amd import seeding
    sc = seeding.config_c2()
    p = sc.params
    _log(f"C2 {sc.shape}, {sc.n_particles} particles, APIC, mixed PCG to {p['tol']}")
    s, rec = _second_substep(sc, "apic", p["tol"], p["max_iter"], p["dt"])
    assert s.n_particles == sc.n_particles == 3830400
    marker = _check_sort(s, rec)
    # P2G of the whole grid against the FP64 oracle (APIC, affine from substep 1)
    u, v, w, uw, vw, ww = fo.p2g(rec["pos"], rec["vel"], marker, s.dx, affine=rec["affine"])
    for f, wt, key in ((u, uw, "us"), (v, vw, "vs"), (w, ww, "ws")):
        want, _ = fo.normalize(f, wt)
        e = gu.rel_err(getattr(s, key).numpy(), want)
        assert e < TOL, (key, e)
    div, res32 = _check_pressure(s, marker, p["tol"], p["max_iter"])
    _check_projection(s, marker, div, p["dt"], 1.0, res32)
    _check_g2p(s, rec, "apic", p["dt"], 0.0, 400_000, seed=2)


@pytest.mark.slow
def test_c3_full_size_second_substep():
    from gpufluid_amd import seeding
    sc = seeding.config_c3()
    p = sc.params
    _log(f"C3 {sc.shape}, {sc.n_particles} particles, mixed PCG to {p['tol']} cap {p['max_iter']}")
    s, rec = _second_substep(sc, "flip", p["tol"], p["max_iter"], p["dt"], p["flip_blend"])
    assert s.n_particles == sc.n_particles == 40196736
    marker = _check_sort(s, rec)
    # the pool, the sphere's lower edge, and a box straddling the pool surface
    for lo in ((112, 144, 112), (40, 64, 184), (200, 8, 8)):
        _check_p2g_box(s, rec, lo, 32)
    div, res32 = _check_pressure(s, marker, p["tol"], p["max_iter"])
    assert s.last_pressure_residual < 1e-5
    _check_projection(s, marker, div, p["dt"], 1.0, res32)
    _check_g2p(s, rec, "flip", p["dt"], p["flip_blend"], 500_000, seed=3)


# --------------------------------------------------------------------- C5
@pytest.mark.slow
def test_c5_stratified_sample_packed_matches_alone():
    from gpufluid_amd import seeding
    from gpufluid_amd.batch import BatchedScenes
    scenes = seeding.config_c5()
    edges = np.array([sc.shape[0] for sc in scenes])
    pick = []
    for e in (24, 32, 48, 64, 96):          # stratified: every edge size, 16 scenes
        idx = np.nonzero(edges == e)[0]
        pick += list(idx[: 4 if e == 96 else 3])
    sample = [scenes[i] for i in pick]
    assert len(sample) == 16 and max(sc.shape[0] for sc in sample) == 96
    _log(f"C5 sample {pick}: edges {[sc.shape[0] for sc in sample]}, "
         f"{sum(sc.n_particles for sc in sample)} particles")
    dt = sample[0].params["dt"]
    b = BatchedScenes([(sc.shape[0], sc.pos, sc.vel) for sc in sample], dt=dt, device="cuda:0")
    substeps = 3
    its = []
    first = None
    for k in range(substeps):
        b.step(pressure_iters=500, pressure_tol=1e-8)
        its.append(b.last_iters.copy())
        assert (b.last_residual < 1e-8).all()
        if k == 0:
            first = packed_first_substep(b)
    check_against_alone(b, [(sc.shape[0], sc.pos, sc.vel) for sc in sample], its, first,
                        substeps, "configs C5")
