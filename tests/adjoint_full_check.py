"""Gradient checks of the differentiable rollout at the full C4 workload (run
as a script on a GPU host; about two minutes, mostly FP64 on the CPU).

Workload: 64^3 box (x in [0.2,0.5], y < 0.5, 292,144 particles, velocities
N(0, 0.5^2) from seed 3), 40 FLIP substeps at dt = 1/240, mixed PCG to 1e-8,
checkpoints every 10 substeps; loss = |centre of mass - (0.7,0.3,0.5)|^2.

Checks printed (one JSON line):
  fp64_vjp_same_trajectory_rel_err: the GPU adjoint against FP64
    vector-Jacobian products evaluated at the GPU's own substep states (the
    adjoint arithmetic, independent of forward drift);
  fp64_rollout_rel_err: against autograd of an independent FP64 rollout
    (adds the FP32/FP64 drift of the forward trajectory);
  fp64_fd_eps1e-5_unit_dirs: FP64 central differences on unit-norm
    directions next to the FP64 and GPU directional derivatives;
  fp32_fd_frozen_masks: central differences of the GPU forward with masks
    frozen.  At this size they cross discontinuities that frozen masks do not
    remove (a face whose only particle leaves its support jumps to 0), so they
    bound nothing here; tests/test_gpu_adjoint.py runs the same check at 16^3.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from gpufluid_amd import FlipSolver3D, seeding  # noqa: E402
from gpufluid_amd.adjoint import FlipRollout  # noqa: E402
from oracle import flip_oracle as fo  # noqa: E402
from oracle import torch_rollout as tr  # noqa: E402


def main(fd_dirs=16, eps=1e-3):
    sc = seeding.config_c4()
    n = sc.shape[0]
    p = sc.params
    target = torch.tensor(p["target"], dtype=torch.float64)
    s = FlipSolver3D(n, n, n, dx=sc.dx, flip_blend=0.95, device="cuda:0")
    roll = FlipRollout(s, p["dt"], p["substeps"], checkpoint_every=p["checkpoint_every"],
                       pressure_iters=500, pressure_tol=1e-8)
    p0 = torch.as_tensor(sc.pos, device="cuda:0")
    v0 = torch.as_tensor(sc.vel, device="cuda:0")

    def loss_of(x):
        return float(((x.double().cpu().mean(0) - target) ** 2).sum())

    print("[adjoint-check] forward + reverse on the GPU", file=sys.stderr, flush=True)
    roll.trace_masks = True
    x = roll.run(p0, v0, structure="record")
    frozen, base_trace = roll.structure, list(roll.mask_trace)
    npart = x.shape[0]
    com = x.double().mean(0).cpu()
    gx = (2.0 * (com - target) / npart).float().to("cuda:0").expand(npart, 3).contiguous()
    _, gvel = roll.backward(gx)
    g = gvel.double().cpu().numpy()
    out = {"workload": "rollout_box_64 (292,144 particles, 40 FLIP substeps, checkpoint every 10)"}
    walls = fo.wall_marker(n, n, n)

    print("[adjoint-check] FP64 VJP along the GPU trajectory", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    roll.run(p0, v0, keep_states=True)
    states = [(a.cpu().numpy(), b.cpu().numpy()) for a, b in roll.states]
    _, gv_same = tr.vjp_along(states, walls, sc.dx, p["dt"], gx.double().cpu().numpy())
    out["fp64_vjp_same_trajectory_rel_err"] = float(np.linalg.norm(g - gv_same)
                                                    / np.linalg.norm(gv_same))
    out["fp64_vjp_seconds"] = round(time.perf_counter() - t0, 1)

    print("[adjoint-check] FP64 rollout autograd", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    xr0 = torch.from_numpy(sc.pos.astype(np.float64))
    vr0 = torch.from_numpy(sc.vel.astype(np.float64)).requires_grad_()
    xr, _ = tr.rollout(xr0, vr0, walls, sc.dx, p["dt"], p["substeps"])
    tr.com_loss(xr, p["target"]).backward()
    gr = vr0.grad.numpy()
    out["fp64_rollout_rel_err"] = float(np.linalg.norm(g - gr) / np.linalg.norm(gr))

    # FP64 central differences on unit-norm directions: at eps 1e-5 no particle
    # crosses a mask or face-support change and FP64 round-off stays far below
    # the signal, so this pins the FP64 gradient at full size
    print("[adjoint-check] FP64 central differences", file=sys.stderr, flush=True)
    fd64 = []
    rng64 = np.random.default_rng(777)
    for _ in range(2):
        d = rng64.normal(size=sc.vel.shape)
        d /= np.linalg.norm(d)
        lv = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                xs, _ = tr.rollout(xr0, torch.from_numpy(sc.vel.astype(np.float64) + sgn * 1e-5 * d),
                                   walls, sc.dx, p["dt"], p["substeps"], checkpoint=False)
            lv.append(float(tr.com_loss(xs, p["target"])))
        fd = (lv[0] - lv[1]) / 2e-5
        fd64.append({"fd": fd, "fp64_grad": float((gr * d).sum()), "gpu_grad": float((g * d).sum())})
    out["fp64_fd_eps1e-5_unit_dirs"] = fd64
    out["fp64_reference_seconds"] = round(time.perf_counter() - t0, 1)

    print("[adjoint-check] FP32 central differences of the GPU forward", file=sys.stderr, flush=True)
    rng = np.random.default_rng(12345)
    errs, crossed = [], 0
    for _ in range(fd_dirs):
        d = rng.normal(size=sc.vel.shape).astype(np.float32)
        dd = torch.as_tensor(d, device="cuda:0")
        lp = loss_of(roll.run(p0, v0 + eps * dd, structure=frozen))
        crossed += roll.mask_trace != base_trace
        lm = loss_of(roll.run(p0, v0 - eps * dd, structure=frozen))
        crossed += roll.mask_trace != base_trace
        an = float((g * d.astype(np.float64)).sum())
        errs.append(abs((lp - lm) / (2 * eps) - an) / max(abs(an), 1e-30))
    out["fp32_fd_frozen_masks"] = {"eps": eps, "directions": fd_dirs,
                                   "runs_that_would_change_a_mask": f"{crossed} of {2 * fd_dirs}",
                                   "rel_err_max": float(np.max(errs)),
                                   "rel_err_median": float(np.median(errs))}
    print(json.dumps(out))
    return out


if __name__ == "__main__":
    main()
