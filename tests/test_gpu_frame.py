"""Between-frame passes on the HIP path: GPU reseed (S2.11.GPU), the host-
array reseed (S2.11), outflow compaction (D4.7.GPU) and prepare_frame's
inflow/outflow handling (F3.6), against the oracle (oracle/frame_oracle.py)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import gfutil as gu
from oracle import frame_oracle as fr

pytestmark = pytest.mark.gpu


def _scene(seed, n=20000, shape=(24, 20, 16)):
    rng = np.random.default_rng(seed)
    dx = 1.0 / max(shape)
    ext = np.array(shape) * dx
    pos = (dx + rng.random((n, 3)) * (ext - 2 * dx) * np.array([0.6, 1, 1])).astype(np.float32)
    pos[: n // 4] = pos[: n // 4] * np.float32(0.3) + np.float32(2 * dx)   # a dense clump
    marker = np.zeros(shape, np.int32)
    marker[0] = marker[-1] = 2
    marker[:, 0] = marker[:, -1] = 2
    marker[:, :, 0] = marker[:, :, -1] = 2
    marker[3:5, 3:5, 3:5] = 2
    marker[15:19, 5:9, 5:9] = 1
    vel = rng.normal(size=pos.shape).astype(np.float32)
    return pos, vel, marker, dx


def test_reseed_gpu_matches_oracle_bit_for_bit():
    from gpufluid_amd.reseed import ReseedConfig, reseed_particles_gpu
    pos, vel, marker, dx = _scene(1)
    col = np.random.default_rng(2).random((len(pos), 3)).astype(np.float32)
    temp = np.random.default_rng(3).random(len(pos)).astype(np.float32)
    cfg = ReseedConfig(min_per_cell=4, max_per_cell=12, jitter=0.3)
    dev = torch.device("cuda:0")
    print(f"[frame] reseed gpu n={len(pos)}", flush=True)
    out = reseed_particles_gpu(torch.tensor(pos, device=dev), torch.tensor(vel, device=dev),
                               torch.tensor(marker, device=dev), dx, cfg,
                               np.random.default_rng(7), attr_color_wp=torch.tensor(col, device=dev),
                               attr_temperature_wp=torch.tensor(temp, device=dev))
    p, v, c, t, n_emit, n_cull = out
    rp, rv, (rc, rt), r_emit, r_drop = fr.reseed(pos, vel, marker, dx, 4, 12, 0.3,
                                                 np.random.default_rng(7), True,
                                                 attrs=(col, temp), emit_fill=(1.0, 0.0))
    assert (n_emit, n_cull) == (r_emit, r_drop) and n_emit > 0 and n_cull > 0
    np.testing.assert_array_equal(p.numpy(), rp)
    np.testing.assert_array_equal(v.numpy(), rv)
    np.testing.assert_array_equal(c.numpy(), rc)
    np.testing.assert_array_equal(t.numpy(), rt)


def test_host_reseed_matches_oracle_and_return_forms():
    from gpufluid_amd.reseed import ReseedConfig, reseed_particles
    pos, vel, marker, dx = _scene(4)
    cfg = ReseedConfig(min_per_cell=3, max_per_cell=10)
    out4 = reseed_particles(pos, vel, marker, dx, cfg, np.random.default_rng(9))
    assert len(out4) == 4
    rp, rv, _, r_emit, r_drop = fr.reseed(pos, vel, marker, dx, 3, 10, 0.3,
                                          np.random.default_rng(9), False)
    np.testing.assert_array_equal(out4[0], rp)
    np.testing.assert_array_equal(out4[1], rv)
    assert out4[2:] == (r_emit, r_drop)
    temp = np.arange(len(pos), dtype=np.float32)
    out6 = reseed_particles(pos, vel, marker, dx, cfg, np.random.default_rng(9),
                            attr_temperature=temp)
    assert len(out6) == 6 and out6[4] is None
    _, _, (_, rt), _, _ = fr.reseed(pos, vel, marker, dx, 3, 10, 0.3, np.random.default_rng(9),
                                    False, attrs=(None, temp), emit_fill=(0.0, 20.0))
    np.testing.assert_array_equal(out6[5], rt)


def test_reseed_empty_particle_set_emits_into_fluid_marker():
    from gpufluid_amd.reseed import ReseedConfig, reseed_particles_gpu
    _, _, marker, dx = _scene(5)
    dev = torch.device("cuda:0")
    z = torch.zeros((0, 3), device=dev)
    p, v, c, t, n_emit, n_cull = reseed_particles_gpu(z, z, torch.tensor(marker, device=dev), dx,
                                                      ReseedConfig(), np.random.default_rng(1))
    assert n_cull == 0 and n_emit == 4 * int((marker == 1).sum()) == p.shape[0]
    assert c is None and t is None


def test_solver_reassignment_like_the_frame_loop():
    """The CLI's pattern: reseed the solver's arrays, assign them back, step."""
    from gpufluid_amd.reseed import ReseedConfig, reseed_particles_gpu
    shape, dx, pos, vel = gu.dam_scene(n=20, seed=6)
    s = gu.make_solver(shape, dx, pos, vel, transfer_mode="apic")
    s.attr_color = np.random.default_rng(1).random((len(pos), 3)).astype(np.float32)
    s.step(1 / 240, pressure_iters=20)
    s.step(1 / 240, pressure_iters=20)
    out = reseed_particles_gpu(s.pos, s.vel, s.marker, s.dx, ReseedConfig(max_per_cell=6),
                               np.random.default_rng(2), attr_color_wp=s.attr_color)
    p, v, c, _, n_emit, n_cull = out
    assert n_emit + n_cull > 0
    s.pos = p
    s.vel = v
    s.attr_color = c
    s.affine_C = None
    s.n_particles = int(p.shape[0])
    np.testing.assert_array_equal(s.pos.numpy(), p.numpy())
    np.testing.assert_array_equal(s.vel.numpy(), v.numpy())
    assert not s.affine_C.numpy().any()
    s.step(1 / 240, pressure_iters=20)
    assert not s.has_diverged() and s.attr_color.numpy().shape == (s.n_particles, 3)


def test_prepare_frame_outflow_then_inflow():
    from gpufluid_amd import FlipSolver3D
    from gpufluid_amd.regions import InflowBox, OutflowBox
    n = 20
    s = FlipSolver3D(n, n, n, dx=1 / n, device="cuda:0")
    s.seed_box((0.1, 0.1, 0.1), (0.6, 0.5, 0.6), ppc=8, temperature=50.0)
    for _ in range(3):
        s.step(1 / 240, pressure_iters=20)
    s.add_solid_from_sdf(np.where(np.indices((n, n, n))[0] >= 17, -1.0, 1.0))
    ob = OutflowBox(lo=(0.0, 0.0, 0.0), hi=(0.3, 0.3, 1.0))
    ib = InflowBox(lo=(0.5, 0.6, 0.4), hi=(0.95, 0.8, 0.6), velocity=(0.0, -1.0, 0.0),
                   rate_per_sec=6000.0, temperature=80.0)
    s.add_outflow(ob)
    s.add_inflow(ib)
    pos0, vel0 = s.get_particles()
    t0 = s.attr_temperature.numpy()
    alive = fr.outflow_alive(pos0, [(ob.lo, ob.hi)])
    emit = ib.emit(0, 1 / 30, np.random.default_rng(0))          # the solver's own rng seed
    ix = np.clip((emit / np.float32(s.dx)).astype(np.int32), 0, n - 1)
    keep = s._marker_host[ix[:, 0], ix[:, 1], ix[:, 2]] != 2
    assert (~alive).any() and (~keep).any()
    s.prepare_frame(0, 1 / 30)
    p1, v1 = s.get_particles()
    m = int(alive.sum())
    np.testing.assert_array_equal(p1[:m], pos0[alive])
    np.testing.assert_array_equal(v1[:m], vel0[alive])
    np.testing.assert_array_equal(p1[m:], emit[keep])
    assert (v1[m:] == np.float32([0, -1, 0])).all()
    t1 = s.attr_temperature.numpy()
    np.testing.assert_array_equal(t1[:m], t0[alive])
    assert (t1[m:] == 80.0).all()
    s.step(1 / 240, pressure_iters=20)
    assert not s.has_diverged()


def test_animated_obstacles_match_oracle():
    from gpufluid_amd import FlipSolver3D
    from gpufluid_amd.animation import KeyframeMotion, LinearMotion
    shape = (24, 20, 22)
    dx = 1.0 / 24
    s = FlipSolver3D(*shape, dx=dx, device="cuda:0")
    idx = np.indices(shape)
    static = np.where((idx[0] >= 19) & (idx[1] <= 3), -1.0, 1.0).astype(np.float32)
    s.add_solid_from_sdf(static)
    s.add_animated_obstacle("sphere", (0.3, 0.4, 0.4), LinearMotion(velocity=(0.24, 0.0, 0.0)),
                            radius=0.12)
    s.add_animated_obstacle("box", (0.6, 0.3, 0.5),
                            KeyframeMotion(keyframes=[(0, (0.6, 0.3, 0.5)), (10, (0.6, 0.5, 0.5))]),
                            half_size=(0.08, 0.05, 0.1))
    s.add_animated_obstacle("cylinder_y", (0.45, 0.5, 0.7), None, radius=0.09, half_height=0.2)
    wall = np.zeros(shape, np.int32)
    wall[0] = wall[-1] = 2
    wall[:, 0] = wall[:, -1] = 2
    wall[:, :, 0] = wall[:, :, -1] = 2
    wall[static <= 0] = 2
    for frame in (0, 5, 12):
        print(f"[frame] animated obstacles frame {frame}", flush=True)
        s.prepare_frame(frame, 1 / 24)
        c_sph = np.float32([0.3, 0.4, 0.4]) + np.float32(frame / 24.0) * np.float32([0.24, 0, 0])
        t = min(frame, 10) / 10.0
        c_box = (np.float32([0.6, 0.3, 0.5]) * np.float32(1 - t) + np.float32([0.6, 0.5, 0.5]) *
                 np.float32(t)) if frame < 10 else np.float32([0.6, 0.5, 0.5])
        # keyframes: centred difference over one frame at 24 fps (clamped ends)
        v_box = s._obstacle_velocity_at(s._anim_specs[1], frame)
        want = {0: 0.24, 5: 0.48, 12: 0.0}[frame]
        assert np.allclose(v_box, (0.0, want, 0.0), atol=1e-5), v_box
        obs = [("sphere", c_sph, {"radius": 0.12}, (0.24, 0.0, 0.0)),
               ("box", c_box, {"half_size": (0.08, 0.05, 0.1)}, v_box),
               ("cylinder_y", (0.45, 0.5, 0.7), {"radius": 0.09, "half_height": 0.2}, (0, 0, 0))]
        m, su, sv, sw = fr.stamp_obstacles(wall, dx, obs)
        np.testing.assert_array_equal(s.marker.numpy(), m)
        np.testing.assert_array_equal(s.solid_u.numpy(), su)
        np.testing.assert_array_equal(s.solid_v.numpy(), sv)
        np.testing.assert_array_equal(s.solid_w.numpy(), sw)
        np.testing.assert_array_equal(s._marker_host, m)
        assert (m != wall).any()


def test_moving_obstacle_pushes_the_fluid():
    from gpufluid_amd import FlipSolver3D
    from gpufluid_amd.animation import LinearMotion
    n = 24
    s = FlipSolver3D(n, n, n, dx=1 / n, device="cuda:0", gravity=0.0)
    s.seed_box((0.05, 0.05, 0.05), (0.95, 0.6, 0.95), ppc=8)
    s.add_animated_obstacle("sphere", (0.3, 0.3, 0.5), LinearMotion(velocity=(1.0, 0.0, 0.0)),
                            radius=0.1)
    for f in range(3):
        s.prepare_frame(f, 1 / 24)
        s.step(1 / 240, pressure_iters=200, pressure_solver="pcg")
    pos, vel = s.get_particles()
    ahead = (pos[:, 0] > 0.42) & (pos[:, 0] < 0.55) & (np.abs(pos[:, 1] - 0.3) < 0.08) & \
            (np.abs(pos[:, 2] - 0.5) < 0.08)
    assert ahead.any() and vel[ahead, 0].mean() > 0.05
    assert not s.has_diverged()
