"""Per-particle colour (S2.15) and scalar (S2.18) transfer on the HIP path
against the oracle (solver3d.py:562-755, drivers 3304-3359), seed_box's
append-or-replace semantics (2827-2910) and attributes in checkpoints."""
from __future__ import annotations

import numpy as np
import pytest

import gfutil as gu
from oracle import flip_oracle as fo

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _attrs(n, seed, t_lo=250.0, t_hi=400.0):
    rng = np.random.default_rng(seed)
    col = rng.random((n, 3)).astype(np.float32)
    temp = (t_lo + (t_hi - t_lo) * rng.random(n)).astype(np.float32)
    return col, temp


@pytest.mark.parametrize("transfer", ["flip", "apic"])
def test_transfer_matches_oracle_at_advected_positions(transfer):
    """The transfer runs after G2P at the advected positions: the oracle's
    transfer at the HIP path's own positions must give the same attributes."""
    shape, dx, pos, vel = gu.dam_scene(shape=(20, 24, 16), vel_sigma=0.4, seed=21)
    col, temp = _attrs(len(pos), 1)
    print(f"[attr] {transfer} {shape} n={len(pos)}", flush=True)
    s = gu.make_solver(shape, dx, pos, vel, transfer_mode=transfer)
    s.attr_color = col
    s.attr_temperature = temp
    s.step(1.0 / 240, pressure_iters=30)
    p_after, _ = s.get_particles()
    ref_c = fo.attr_transfer(p_after, col, dx, shape, fo.ATTR_COLOR)
    ref_t = fo.attr_transfer(p_after, temp, dx, shape, fo.ATTR_SCALAR)
    got_c, got_t = s.attr_color.numpy(), s.attr_temperature.numpy()
    assert got_c.shape == (len(pos), 3) and got_t.shape == (len(pos),)
    assert gu.rel_err(got_c, ref_c) < TOL
    assert gu.rel_err(got_t, ref_t) < TOL
    assert not np.array_equal(got_c, col)          # the transfer did run


def test_transfer_wide_value_range_and_zero():
    """Fixed-point scaling follows the largest |value|: values spanning
    1e-3..1e4 with both signs, and an all-zero attribute."""
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=22)
    rng = np.random.default_rng(2)
    temp = (rng.choice([-1.0, 1.0], len(pos)) * 10.0 ** rng.uniform(-3, 4, len(pos)))
    temp = temp.astype(np.float32)
    s = gu.make_solver(shape, dx, pos, vel)
    s.attr_temperature = temp
    s.attr_color = np.zeros((len(pos), 3), np.float32)
    s.step(1.0 / 240, pressure_iters=20)
    p_after, _ = s.get_particles()
    ref_t = fo.attr_transfer(p_after, temp, dx, shape, fo.ATTR_SCALAR)
    assert gu.rel_err(s.attr_temperature.numpy(), ref_t) < TOL
    assert not s.attr_color.numpy().any()


def _group_boxes(s, group=256):
    """Cells in the base-cell box of each 256-slot workgroup of the scatter,
    from the slot-ordered positions after the step."""
    q = np.floor(s._pos.cpu().numpy() / np.float32(s.dx) - np.float32(0.5)).astype(np.int64)
    out = []
    for g in range(0, len(q), group):
        b = q[g:g + group]
        out.append(int(np.prod(b.max(0) - b.min(0) + 2)))
    return np.array(out)


def test_transfer_dense_block_and_sparse_spray():
    """A dense block (workgroups sum in LDS) plus a sparse spray across the
    domain (workgroups whose box exceeds the LDS budget add to the grid
    directly): both paths match the oracle."""
    shape, dx, pos, vel = gu.dam_scene(n=32, seed=27, x_max=0.3, y_max=0.4)
    rng = np.random.default_rng(27)
    spray = (dx * 1.5 + rng.random((1500, 3)) * (1 - 3 * dx)).astype(np.float32)
    pos = np.concatenate([pos, spray])
    vel = np.concatenate([vel, np.zeros_like(spray)])
    col, temp = _attrs(len(pos), 7)
    s = gu.make_solver(shape, dx, pos, vel)
    s.set_particles(pos, vel, color=col, temperature=temp)
    s.step(1.0 / 240, pressure_iters=20)
    boxes = _group_boxes(s)
    cap_color = 48 * 1024 // (8 * 4)
    assert (boxes <= cap_color).any() and (boxes > cap_color).any(), boxes
    p_after, _ = s.get_particles()
    assert gu.rel_err(s.attr_color.numpy(),
                      fo.attr_transfer(p_after, col, dx, shape, fo.ATTR_COLOR)) < TOL
    assert gu.rel_err(s.attr_temperature.numpy(),
                      fo.attr_transfer(p_after, temp, dx, shape, fo.ATTR_SCALAR)) < TOL


def test_transfer_is_deterministic():
    shape, dx, pos, vel = gu.dam_scene(shape=(24, 20, 20), vel_sigma=0.3, seed=23)
    col, temp = _attrs(len(pos), 3)
    out = []
    for _ in range(2):
        s = gu.make_solver(shape, dx, pos, vel)
        s.set_particles(pos, vel, color=col, temperature=temp)
        for _ in range(3):
            s.step(1.0 / 240, pressure_iters=20)
        out.append((s.attr_color.numpy(), s.attr_temperature.numpy()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_full_substep_with_attributes_matches_oracle():
    shape, dx, pos, vel = gu.dam_scene(shape=(20, 24, 16), vel_sigma=0.4, seed=24)
    col, temp = _attrs(len(pos), 4)
    s = gu.make_solver(shape, dx, pos, vel)
    s.set_particles(pos, vel, color=col, temperature=temp)
    dt = 1.0 / 240
    s.step(dt, pressure_iters=400, pressure_solver="pcg", pressure_tol=1e-8)
    st0 = gu.oracle_state(shape, dx, pos, vel)
    st0["color"], st0["temperature"] = col, temp
    st, _ = fo.substep(st0, dt, pressure="exact")
    assert gu.rel_err(s.attr_color.numpy(), st["color"]) < 1e-4
    assert gu.rel_err(s.attr_temperature.numpy(), st["temperature"]) < 1e-4


def test_seed_box_append_or_replace():
    from gpufluid_amd import FlipSolver3D
    from gpufluid_amd.seeding import seed_box_positions
    n = 16
    s = FlipSolver3D(n, n, n, dx=1 / n, device="cuda:0")
    box_a = ((0.2, 0.2, 0.2), (0.4, 0.5, 0.4))
    box_b = ((0.6, 0.2, 0.6), (0.8, 0.4, 0.8))
    pa = seed_box_positions(*box_a, 8, s.dx)
    pb = seed_box_positions(*box_b, 8, s.dx)
    # no attributes anywhere: the second call replaces
    s.seed_box(*box_a, ppc=8)
    s.seed_box(*box_b, ppc=8)
    assert s.n_particles == len(pb) and s.attr_color is None
    # coloured seeds append
    s = FlipSolver3D(n, n, n, dx=1 / n, device="cuda:0")
    s.seed_box(*box_a, ppc=8, color=(1.0, 0.0, 0.0))
    s.seed_box(*box_b, ppc=8, color=(0.0, 0.0, 1.0))
    assert s.n_particles == len(pa) + len(pb)
    np.testing.assert_array_equal(s.pos.numpy(), np.concatenate([pa, pb]))
    col = s.attr_color.numpy()
    np.testing.assert_array_equal(col[: len(pa)], np.tile([1, 0, 0], (len(pa), 1)))
    np.testing.assert_array_equal(col[len(pa):], np.tile([0, 0, 1], (len(pb), 1)))
    assert s.attr_temperature is None
    # an uncoloured seed into a coloured scene is padded white; a temperature
    # seed pads the earlier particles with 0
    s.seed_box(*box_a, ppc=8, temperature=300.0)
    m = len(pa) + len(pb)
    assert s.n_particles == m + len(pa)
    np.testing.assert_array_equal(s.attr_color.numpy()[m:], np.ones((len(pa), 3)))
    t = s.attr_temperature.numpy()
    assert not t[:m].any() and (t[m:] == 300.0).all()
    # previous particles keep position and velocity after stepping
    s.step(0.005, pressure_iters=10)
    v_before = s.vel.numpy()
    s.seed_box(*box_b, ppc=8, color=(0.0, 1.0, 0.0))
    np.testing.assert_array_equal(s.vel.numpy()[: len(v_before)], v_before)
    assert not s.vel.numpy()[len(v_before):].any()


def test_attribute_length_is_checked():
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=25)
    s = gu.make_solver(shape, dx, pos, vel)
    with pytest.raises(ValueError):
        s.attr_color = np.zeros((len(pos) + 1, 3), np.float32)
    with pytest.raises(ValueError):
        s.set_particles(pos, vel, temperature=np.zeros(len(pos) - 1, np.float32))


def test_checkpoint_carries_attributes(tmp_path):
    from gpufluid_amd import FlipSolver3D
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=26)
    col, temp = _attrs(len(pos), 6)
    a = gu.make_solver(shape, dx, pos, vel)
    a.set_particles(pos, vel, color=col, temperature=temp)
    for _ in range(3):
        a.step(0.005, pressure_iters=20)
    ck = tmp_path / "ck.npz"
    a.save_checkpoint(ck)
    d = np.load(ck, allow_pickle=False)
    np.testing.assert_array_equal(d["attr_color"], a.attr_color.numpy())
    np.testing.assert_array_equal(d["attr_temperature"], a.attr_temperature.numpy())
    b = FlipSolver3D(*shape, dx=dx, device="cuda:0")
    b.load_checkpoint(ck)
    np.testing.assert_array_equal(b.attr_color.numpy(), a.attr_color.numpy())
    np.testing.assert_array_equal(b.attr_temperature.numpy(), a.attr_temperature.numpy())
    for s in (a, b):
        for _ in range(3):
            s.step(0.005, pressure_iters=20)
    assert gu.rel_err(b.attr_color.numpy(), a.attr_color.numpy()) < 1e-5
    assert gu.rel_err(b.attr_temperature.numpy(), a.attr_temperature.numpy()) < 1e-5


@pytest.mark.parametrize("mode", ["color", "scalar"])
def test_block_callables_match_oracle_pieces(mode):
    """S2.15.1-3 / S2.18.1-3 block callables on the reference's separate FP32
    grids, particles in caller order: P2G adds the exact sums rounded to FP32
    to grids that already hold values; normalize and G2P on the same grids
    are bit-exact with the oracle's."""
    import torch
    from gpufluid_amd import kernels as K
    shape, dx = (14, 12, 16), 1.0 / 16
    rng = np.random.default_rng(31)
    pos = (rng.random((2500, 3)) * (np.array(shape) + 1) * dx - 0.5 * dx).astype(np.float32)
    ch = 3 if mode == "color" else 1
    attr = (rng.random((len(pos), ch)) * 4 - 1).astype(np.float32).reshape(len(pos), -1)
    if ch == 1:
        attr = attr[:, 0].copy()
    dev = "cuda:0"
    T = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    base = [rng.random(shape).astype(np.float32) * 0.125 for _ in range(ch + 1)]
    grids = [T(b) for b in base]
    pt, at = T(pos), T(attr)
    print(f"[attr] block callables {mode} {shape} n={len(pos)}", flush=True)
    if ch == 3:
        K.k3_p2g_color(pt, at, *grids, dx, *shape)
    else:
        K.k3_p2g_scalar(pt, at, *grids, dx, *shape)
    vsum, wsum = fo.attr_p2g(pos, attr, dx, shape)
    exp = [b + vsum[c].astype(np.float32) for c, b in enumerate(base[:ch])]
    exp.append(base[ch] + wsum.astype(np.float32))
    for c in range(ch + 1):
        got = grids[c].cpu().numpy()
        assert gu.rel_err(got, exp[c]) < 1e-6
    # normalize and gather on the GPU's own grids
    g_host = [g.cpu().numpy() for g in grids]
    if ch == 3:
        K.k3_normalize_color(*grids)
    else:
        K.k3_normalize_scalar(*grids)
    want_g, want_w = fo.attr_normalize(np.stack(g_host[:ch]), g_host[ch],
                                       fo.ATTR_COLOR if ch == 3 else fo.ATTR_SCALAR)
    for c in range(ch):
        np.testing.assert_array_equal(grids[c].cpu().numpy(), want_g[c])
    if ch == 3:
        K.k3_g2p_color(pt, at, *grids, dx, *shape)
    else:
        K.k3_g2p_scalar(pt, at, *grids, dx, *shape)
    want = fo.attr_g2p(pos, attr, want_g, want_w, dx)
    np.testing.assert_array_equal(at.cpu().numpy(), want)
