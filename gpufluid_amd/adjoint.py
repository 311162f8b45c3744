"""Differentiable multi-substep FLIP/PIC rollouts.

Forward: the fused substep of ``FlipSolver3D`` (sort, P2G, PCG, G2P),
unchanged.  Reverse: the hand-written adjoint kernels of
``gpufluid_amd/csrc/gf_adjoint.hip`` — G2P's gather becomes a scatter, P2G's
scatter a gather, and the pressure solve's adjoint is a second PCG on the same
SPD matrix — with checkpoint-every-k recomputation: the forward keeps the
particle state every ``checkpoint_every`` substeps, and the reverse pass
re-runs each segment (bit-identically: every forward kernel is deterministic)
while recording what the adjoint needs.

The reference solver is not differentiable; this module has no counterpart
there.  Masks (fluid/solid cells) and the cell binning are piecewise constant
in the particle positions and are held fixed, as usual for FLIP adjoints.

    roll = FlipRollout(solver, dt=1/240, substeps=40, checkpoint_every=10)
    x = rollout_positions(roll, pos0, vel0)        # vel0.requires_grad_()
    loss = ((x.double().mean(0) - target) ** 2).sum()
    loss.backward()                                  # vel0.grad
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .blocks import block


class FlipRollout:
    """``substeps`` FLIP/PIC substeps of ``solver`` with the PCG pressure solve."""

    def __init__(self, solver, dt: float, substeps: int, checkpoint_every: int = 10,
                 pressure_iters: int = 500, pressure_tol: float = 1e-8):
        if solver.transfer_mode not in ("flip", "pic"):
            raise NotImplementedError("the substep adjoint covers FLIP/PIC transfers")
        if solver.viscosity > 0.0 or solver.surface_tension > 0.0:
            raise NotImplementedError("the substep adjoint covers the substep without viscosity "
                                      "or surface tension")
        if substeps < 1 or checkpoint_every < 1:
            raise ValueError("substeps and checkpoint_every must be >= 1")
        self.s = solver
        self.dt = float(dt)
        self.substeps = int(substeps)
        self.k = int(checkpoint_every)
        self.iters = int(pressure_iters)
        self.tol = float(pressure_tol)
        self.flip_blend = solver.flip_blend if solver.transfer_mode == "flip" else 0.0
        f = solver._f
        self._work = {name: torch.zeros_like(f[base]) for base in ("u", "v", "w")
                      for name in ("m" + base, "a" + base, "a" + base + "s")}
        self._checkpoints = []
        self.forward_iters = []
        # per-substep checksums of the fluid mask of the last run(); two runs
        # with equal traces saw the same discrete structure (finite-difference
        # checks use this to recognise perturbations that cross a mask change)
        self.trace_masks = False
        self.mask_trace = []
        self._mask_w = None

    # ---------------------------------------------------------------- state
    def _snap_particles(self):
        live = self.s._sc.live(self.s.n_particles)
        return {k: live[k].clone() for k in ("pos", "vel", "pid")}

    def _restore_particles(self, snap):
        s = self.s
        live = s._sc.live(s.n_particles)
        for k in ("pos", "vel", "pid"):
            live[k].copy_(snap[k])
        s._cache.clear()

    def _structure(self):
        s, sc = self.s, self.s._sc
        out = {"marker": s._f["marker"].clone(), "tile_map": sc.tile_map.clone(),
               "flags": sc.flags.clone(), "nbr": sc.nbr.clone()}
        out["n_active"] = sc.n_active_host
        return out

    def _set_structure(self, st):
        s, sc = self.s, self.s._sc
        s._f["marker"].copy_(st["marker"])
        sc.tile_map.copy_(st["tile_map"])
        sc.flags.copy_(st["flags"])
        sc.nbr.copy_(st["nbr"])
        sc.n_active_host = st["n_active"]

    def _substep(self, record=None, frozen=None, keep=None):
        s = self.s
        sc = s._sc

        def hook(stage):
            if stage == "sorted":
                if self.trace_masks and record is None:
                    # the mask the particles produced, before any freezing
                    m = s._f["marker"]
                    if self._mask_w is None:
                        self._mask_w = (torch.arange(m.numel(), device=m.device, dtype=torch.int64)
                                        % 1000003 + 1).view(m.shape)
                    self.mask_trace.append((m.to(torch.int64) * self._mask_w).sum())
                if keep is not None:
                    keep.append(self._structure())
                if frozen is not None:
                    self._set_structure(frozen)
                if record is not None:
                    record.update(self._snap_particles())
            elif record is not None:
                f = s._f
                for k in ("u", "v", "w", "us", "vs", "ws", "marker"):
                    record[k] = f[k].clone()
                record["tile_map"] = sc.tile_map.clone()
                record["flags"] = sc.flags.clone()
                record["nbr"] = sc.nbr.clone()
                record["n_active"] = sc.n_active_host

        s._step_impl(self.dt, self.iters, "pcg", self.tol, hook=hook)
        return s.last_pressure_iters

    # -------------------------------------------------------------- forward
    @block("X1.1", "Differentiable rollout: forward with checkpoints every k substeps")
    def run(self, pos0, vel0, structure=None, keep_states=False):
        """Run the rollout from (pos0, vel0) (caller order); returns the final
        positions in caller order as an (n,3) FP32 tensor.

        ``structure``: None (each substep derives its fluid mask and tile
        tables from the particles), "record" (also keep them, per substep, in
        ``self.structure``), or a list kept by an earlier run, which is reused:
        the discrete structure is frozen, as the adjoint assumes, so finite
        differences of such runs measure the same derivative.
        ``keep_states``: keep (pos, vel) in caller order at the start of every
        substep in ``self.states`` (for reference checks along this trajectory)."""
        s = self.s
        s.set_particles(pos0, vel0)
        self._checkpoints = []
        self.forward_iters = []
        self.mask_trace = []
        keep = None
        if isinstance(structure, str):
            if structure != "record":
                raise ValueError(f"unknown structure mode {structure!r}")
            self.structure = keep = []
        elif structure is not None and len(structure) != self.substeps:
            raise ValueError("structure must hold one entry per substep")
        self.states = []
        for k in range(self.substeps):
            if k % self.k == 0:
                self._checkpoints.append(self._snap_particles())
            if keep_states:
                s._cache.clear()
                self.states.append((s.pos.t.clone(), s.vel.t.clone()))
            frozen = structure[k] if isinstance(structure, list) else None
            self.forward_iters.append(self._substep(frozen=frozen, keep=keep))
        if self.mask_trace:
            self.mask_trace = [int(t) for t in self.mask_trace]
        return s.pos.t.clone()

    # -------------------------------------------------------------- reverse
    def _adjoint_desc(self, rec, apos, avel):
        s = self.s
        a = _lib.GfAdjoint()
        a.nx, a.ny, a.nz = s.nx, s.ny, s.nz
        a.dx = s.dx
        a.dom[0], a.dom[1], a.dom[2] = s.dom
        a.n = s.n_particles
        P = _lib.ptr
        a.pos, a.vel, a.pid = P(rec["pos"]), P(rec["vel"]), P(rec["pid"])
        a.marker, a.tile_map = P(rec["marker"]), P(rec["tile_map"])
        a.u, a.v, a.w = P(rec["u"]), P(rec["v"]), P(rec["w"])
        a.us, a.vs, a.ws = P(rec["us"]), P(rec["vs"]), P(rec["ws"])
        w = self._work
        a.mu, a.mv, a.mw = P(w["mu"]), P(w["mv"]), P(w["mw"])
        a.au, a.av, a.aw = P(w["au"]), P(w["av"]), P(w["aw"])
        a.aus, a.avs, a.aws = P(w["aus"]), P(w["avs"]), P(w["aws"])
        a.apos, a.avel = P(apos), P(avel)
        return a

    def _adjoint_substep(self, rec, apos, avel):
        s = self.s
        sc = s._sc
        st = s._stream()
        a = self._adjoint_desc(rec, apos, avel)
        ref = ctypes.byref(a)
        _lib.call("X1.2", "gf_adj_clear", ref, st)
        _lib.call("X1.2", "gf_adj_weights", ref, st)
        _lib.call("X1.2", "gf_adj_g2p", ref, self.flip_blend, self.dt, st)
        # the pressure adjoint runs on this substep's compact tiles
        sc.tile_map.copy_(rec["tile_map"])
        sc.flags.copy_(rec["flags"])
        sc.nbr.copy_(rec["nbr"])
        sc.n_active_host = rec["n_active"]
        _lib.call("X1.2", "gf_adj_pressure_rhs", ref, self.dt, s.rho, _lib.ptr(sc.b), st)
        d = s._descriptor()
        prec = _lib.GF_PCG_MIXED if s.pcg_precision == "mixed" else _lib.GF_PCG_FP64
        if prec == _lib.GF_PCG_FP64:
            sc.ensure_fp64()
            d = s._descriptor()
        sign = -1 if s.pcg_sign == "physical" else 1
        it = ctypes.c_int(0)
        _lib.call("X1.2", "gf_scene_pcg_sync", ctypes.byref(d), rec["n_active"], prec, self.tol,
                  self.iters, s.pcg_check_every, sign, ctypes.byref(it), None, None, None, st)
        _lib.call("X1.2", "gf_adj_faces", ref, _lib.ptr(sc.pc), self.dt, s.rho, st)
        _lib.call("X1.2", "gf_adj_p2g", ref, st)
        return it.value

    @block("X1.2", "Differentiable rollout: reverse pass with segment recomputation")
    def backward(self, grad_pos, grad_vel=None):
        """Adjoint of the last ``run``: given dL/d(final positions) (and
        optionally dL/d(final velocities)), caller order, returns
        (dL/d pos0, dL/d vel0) in caller order."""
        if not self._checkpoints:
            raise RuntimeError("backward() needs a preceding run()")
        s = self.s
        n = s.n_particles
        dev = s.device
        apos = grad_pos.detach().to(device=dev, dtype=torch.float32).reshape(n, 3).contiguous().clone()
        avel = (torch.zeros((n, 3), dtype=torch.float32, device=dev) if grad_vel is None else
                grad_vel.detach().to(device=dev, dtype=torch.float32).reshape(n, 3).contiguous().clone())
        self.adjoint_iters = []
        for seg in reversed(range(len(self._checkpoints))):
            self._restore_particles(self._checkpoints[seg])
            recs = []
            for _ in range(seg * self.k, min((seg + 1) * self.k, self.substeps)):
                rec = {}
                self._substep(rec)
                recs.append(rec)
            for rec in reversed(recs):
                self.adjoint_iters.append(self._adjoint_substep(rec, apos, avel))
        return apos, avel


class _RolloutFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel0, pos0, rollout):
        ctx.rollout = rollout
        return rollout.run(pos0.detach(), vel0.detach())

    @staticmethod
    def backward(ctx, grad_out):
        gpos0, gvel0 = ctx.rollout.backward(grad_out)
        return gvel0, gpos0, None


def rollout_positions(rollout: FlipRollout, pos0: torch.Tensor, vel0: torch.Tensor) -> torch.Tensor:
    """Final positions of ``rollout`` from (pos0, vel0), differentiable with
    respect to vel0 and pos0 through torch.autograd (caller order, (n,3) FP32)."""
    return _RolloutFunction.apply(vel0, pos0, rollout)
