"""Gradient of the differentiable rollout at the full C4 workload.

64^3 box (x in [0.2,0.5], y < 0.5, 292,144 particles, initial velocities
N(0, 0.5^2) from seed 3), 40 FLIP substeps at dt = 1/240, mixed PCG to 1e-8,
checkpoints every 10 substeps; loss = |centre of mass - (0.7,0.3,0.5)|^2.

* The FP64 reference gradient of the forward the GPU runs: vector-Jacobian
  products of the FP64 substep (oracle/torch_rollout.py) chained along the
  GPU's own substep states.  The GPU adjoint must agree to 1e-3 (it does to
  ~1e-6: every reverse kernel is the exact transpose of its forward stage).
* The gradient of an independent FP64 rollout.  Its trajectory ends about
  3.6e-6 from the FP32 one: FP32 rounding in every stage, amplified by the
  dynamics (an FP64 rollout does not change a single fluid mask relative to
  the FP32 one).  The gradient here moves with the trajectory: two FP64
  rollouts whose initial velocities differ by a relative 6e-7 have
  gradients 1.6e-4 apart with final positions 4.7e-7 apart, about 340x.  So
  the FP32 forward's gradient sits ~1.2e-3 from the FP64 rollout's by
  construction, and the bound here is 2e-3.
The adjoint's time against the forward's (<= 3x) is measured by bench.py.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _log(msg):
    print(f"[adjoint c4] {msg}", flush=True)


def test_c4_gradient_against_fp64_references():
    from gpufluid_amd import FlipSolver3D, seeding
    from gpufluid_amd.adjoint import FlipRollout
    from oracle import flip_oracle as fo
    from oracle import torch_rollout as tr
    torch.set_num_threads(16)
    sc = seeding.config_c4()
    n = sc.shape[0]
    p = sc.params
    assert sc.n_particles == 292144
    target = torch.tensor(p["target"], dtype=torch.float64)
    s = FlipSolver3D(n, n, n, dx=sc.dx, flip_blend=0.95, device="cuda:0")
    roll = FlipRollout(s, p["dt"], p["substeps"], checkpoint_every=p["checkpoint_every"],
                       pressure_iters=500, pressure_tol=1e-8)
    p0 = torch.as_tensor(sc.pos, device="cuda:0")
    v0 = torch.as_tensor(sc.vel, device="cuda:0")
    _log("forward and reverse on the GPU")
    x = roll.run(p0, v0, keep_states=True)
    npart = x.shape[0]
    com = x.double().mean(0).cpu()
    gx = (2.0 * (com - target) / npart).float().to("cuda:0").expand(npart, 3).contiguous()
    _, gvel = roll.backward(gx)
    g = gvel.double().cpu().numpy()
    states = [(a.cpu().numpy(), b.cpu().numpy()) for a, b in roll.states]
    assert len(states) == p["substeps"]
    walls = fo.wall_marker(n, n, n)

    _log("FP64 vector-Jacobian products along the GPU trajectory")
    _, g_same = tr.vjp_along(states, walls, sc.dx, p["dt"], gx.double().cpu().numpy())
    e_same = float(np.linalg.norm(g - g_same) / np.linalg.norm(g_same))
    _log(f"GPU adjoint vs FP64 reference along the same trajectory: {e_same:.3e}")
    assert e_same < 1e-3

    _log("FP64 rollout and its autograd gradient")
    xr0 = torch.from_numpy(sc.pos.astype(np.float64))
    vr0 = torch.from_numpy(sc.vel.astype(np.float64)).requires_grad_()
    xr, _ = tr.rollout(xr0, vr0, walls, sc.dx, p["dt"], p["substeps"])
    tr.com_loss(xr, p["target"]).backward()
    g64 = vr0.grad.numpy()
    e_free = float(np.linalg.norm(g - g64) / np.linalg.norm(g64))
    drift = float(np.abs(xr.detach().numpy() - x.double().cpu().numpy()).max())
    _log(f"GPU adjoint vs the FP64 rollout's gradient: {e_free:.3e} (trajectories end "
         f"{drift:.2e} apart)")
    assert drift < 1e-5
    assert e_free < 2e-3
