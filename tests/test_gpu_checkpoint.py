"""Checkpoint save/load (F3.5) in the reference's .npz layout: resume
matches the uninterrupted run, topology and dx are validated, and the file
carries the reference's keys."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(n=16, **kw):
    from gpufluid_amd import FlipSolver3D
    s = FlipSolver3D(nx=n, ny=n, nz=n, dx=1 / n, device="cuda:0", **kw)
    s.seed_box(lo=(0.20, 0.30, 0.20), hi=(0.50, 0.70, 0.50), ppc=8)
    return s


@pytest.mark.parametrize("transfer", ["flip", "apic"])
def test_resume_matches_uninterrupted_run(tmp_path, transfer):
    """As the reference's tests/test_f3_5_checkpoint.py: 20 steps straight vs
    10 steps, checkpoint, fresh solver, load, 10 steps."""
    from gpufluid_amd import FlipSolver3D
    print(f"[checkpoint] resume, {transfer}", flush=True)
    a = _solver(transfer_mode=transfer)
    for _ in range(20):
        a.step(0.005, pressure_iters=20)
    pos_a, vel_a = a.get_particles()
    b = _solver(transfer_mode=transfer)
    for _ in range(10):
        b.step(0.005, pressure_iters=20)
    ck = tmp_path / "sub" / "ck.npz"
    b.save_checkpoint(ck)
    c = FlipSolver3D(nx=16, ny=16, nz=16, dx=1 / 16, device="cuda:0", transfer_mode=transfer)
    c.load_checkpoint(ck)
    assert c.n_particles == b.n_particles
    for _ in range(10):
        c.step(0.005, pressure_iters=20)
    pos_c, vel_c = c.get_particles()
    rms = float(np.sqrt(((pos_c - pos_a) ** 2).sum(1).mean()))
    print(f"[checkpoint] resume rms {rms:.3e}", flush=True)
    assert pos_c.shape == pos_a.shape
    assert rms < 1e-5


def test_checkpoint_layout_and_validation(tmp_path):
    from gpufluid_amd import FlipSolver3D
    s = _solver()
    s.step(0.005, pressure_iters=10)
    ck = tmp_path / "ck.npz"
    s.save_checkpoint(ck)
    d = np.load(ck, allow_pickle=False)
    for key in ("pos", "vel", "affine_C", "marker", "attr_color", "attr_temperature", "nx", "ny",
                "nz", "dx", "gravity", "flip_blend", "rho", "viscosity", "transfer_mode",
                "surface_tension"):
        assert key in d.files, key
    pos, vel = s.get_particles()
    np.testing.assert_array_equal(d["pos"], pos)
    np.testing.assert_array_equal(d["vel"], vel)
    np.testing.assert_array_equal(d["marker"], s.marker.numpy())
    assert d["marker"].dtype == np.int32 and str(d["transfer_mode"]) == "flip"
    with pytest.raises(RuntimeError, match="checkpoint grid"):
        FlipSolver3D(nx=8, ny=16, nz=16, dx=1 / 16, device="cuda:0").load_checkpoint(ck)
    with pytest.raises(RuntimeError, match="dx="):
        FlipSolver3D(nx=16, ny=16, nz=16, dx=1 / 32, device="cuda:0").load_checkpoint(ck)
    # solids survive the round trip
    t = FlipSolver3D(nx=16, ny=16, nz=16, dx=1 / 16, device="cuda:0")
    sdf = np.ones((16, 16, 16), np.float32)
    sdf[6:9, 2:5, 6:9] = -1.0
    t.add_solid_from_sdf(sdf)
    t.seed_box(lo=(0.2, 0.5, 0.2), hi=(0.5, 0.7, 0.5), ppc=8)
    t.save_checkpoint(ck)
    u = FlipSolver3D(nx=16, ny=16, nz=16, dx=1 / 16, device="cuda:0")
    u.load_checkpoint(ck)
    assert (u.marker.numpy()[6:9, 2:5, 6:9] == 2).all()
    u.step(0.005, pressure_iters=10)
    assert (u.marker.numpy()[6:9, 2:5, 6:9] == 2).all()
