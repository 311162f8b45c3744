"""FlipSolver3D: the reference solver's Python surface over the gfx950 kernels.

Drop-in for src/gpufluid/solvers/solver3d.py:2346-4222 (``FlipSolver3D``) on
the substep hot path: constructor arguments, ``seed_box``, ``step``,
``step_cfl``, ``get_particles``, ``has_diverged``, ``add_solid_from_sdf``,
``cell_centers_np`` and the grid/particle attributes keep the reference's
names, defaults and meaning.  Each substep runs four stages of hand-written
HIP kernels (libgpufluid_hip.so):

  clear       gf_scene_sort_sync   bin + stable counting sort of particles by
                                   cell, marker, active 8^3 tiles
  p2g         gf_scene_p2g         P2G + normalize + gravity + BC + divergence
  pressure    gf_scene_jacobi | gf_scene_gsrb | gf_scene_pcg_sync   on compacted
                                   active tiles
  g2p_advect  gf_scene_g2p         gradient subtract + BC + G2P + advection

Differences from the reference that callers can observe (all deliberate):
  * particles are stored in cell-sorted order; ``pos``/``vel``/``affine_C``
    and ``get_particles()`` return them in the original particle order;
  * PCG solves the Jacobi-consistent system A p = -div by default
    (``pcg_sign="physical"``); the reference PCG returns -p because it starts
    from r = +div (solver3d.py:2962-2963) — ``pcg_sign="reference"``
    reproduces that literally;
  * PCG checks convergence every iteration on the device (the host reads the
    flag every ``pcg_check_every`` iterations) and reports the exact count;
    the tolerance is exposed (``pcg_tol``, ``step(pressure_tol=...)``);
  * dot products are FP64 and deterministic; ``pcg_precision="mixed"`` adds
    FP64 residual replacement so tolerances down to 1e-8 are reachable.
"""
from __future__ import annotations

import ctypes
import logging
import math

import numpy as np
import torch

from . import _lib
from .blocks import BlockError, block
from .device import DeviceArray, as_tensor, default_device
from .profiling import StepProfiler
from .scene import SceneBuffers
from .regions import FrameEventQueue
from .seeding import seed_box_positions
from .subdense import SubDenseBox


class FlipDivergenceError(RuntimeError):
    """Raised by frame loops when ``has_diverged()`` is true (solver3d.py:2317-2343)."""

    def __init__(self, frame: int, last_dumped_frame: int, requested_frames: int,
                 detail: str = "") -> None:
        super().__init__(detail or f"FLIP solver diverged at frame {frame} "
                         f"(last dumped: {last_dumped_frame}/{requested_frames})")
        self.frame = int(frame)
        self.last_dumped_frame = int(last_dumped_frame)
        self.requested_frames = int(requested_frames)


@block("S2.10", "Host CFL substep count from a particle velocity array")
def cfl_substep_count(vel_np: np.ndarray, dx: float, target_dt: float, cfl: float = 0.5,
                      max_substeps: int = 32) -> int:
    """ceil(target_dt / (cfl*dx/|v|max)) clamped to [1, max_substeps]
    (solver3d.py:2218-2238)."""
    if vel_np.size == 0:
        return 1
    vmax = float(np.linalg.norm(vel_np, axis=1).max())
    if vmax < 1e-6:
        return 1
    n = max(1, int(np.ceil(target_dt / (cfl * dx / vmax))))
    return min(n, max_substeps)


def _rows(value, width, device):
    """(n, width) float32 contiguous device tensor of an array-like."""
    t = value.t if isinstance(value, DeviceArray) else value
    return as_tensor(t, torch.float32, device).reshape(-1, width).contiguous()


_TRANSFER = {"flip": _lib.GF_TRANSFER_FLIP, "pic": _lib.GF_TRANSFER_FLIP,
             "apic": _lib.GF_TRANSFER_APIC}


@block("F3.2", "FLIP/PIC/APIC solver on MI355X: MAC grid state and substep pipeline")
class FlipSolver3D:
    """3D FLIP/PIC/APIC solver holding all state on the HIP device.

    Parameters follow solver3d.py:2358-2371; viscosity (S2.13) and surface
    tension (S2.14) run between P2G and the pressure solve as in the
    reference.  enable_sub_dense (F3.7) keeps a box around the fluid: the grid
    attributes show the box and the viscosity sweep runs over it, as the
    reference's box-shaped arrays do (storage stays full-size; the pressure
    vectors are compacted to active tiles anyway).  Extra keyword-only
    parameters:
    pcg_precision ("mixed" | "fp64"), pcg_sign ("physical" | "reference"),
    pcg_tol (reference default 1e-4), pcg_check_every (host readback cadence).
    """

    def __init__(self, nx: int = 48, ny: int = 48, nz: int = 48, dx: float | None = None,
                 gravity: float = -9.81, flip_blend: float = 0.95, rho: float = 1.0,
                 viscosity: float = 0.0, viscosity_iters: int = 12,
                 surface_tension: float = 0.0, csf_smoothing_passes: int = 2,
                 transfer_mode: str = "flip", device=None, enable_timing: bool = False,
                 enable_cuda_graphs: bool = False, enable_sub_dense: bool = False,
                 sub_rebuild_every: int = 8, sub_dilation: int = 4, *,
                 pcg_precision: str = "mixed", pcg_sign: str = "physical",
                 pcg_tol: float = 1e-4, pcg_check_every: int = 10):
        if transfer_mode not in _TRANSFER:
            raise ValueError(f"unknown transfer_mode: {transfer_mode!r}")
        if pcg_precision not in ("mixed", "fp64"):
            raise ValueError(f"unknown pcg_precision: {pcg_precision!r}")
        if pcg_sign not in ("physical", "reference"):
            raise ValueError(f"unknown pcg_sign: {pcg_sign!r}")
        self.device = torch.device(device) if device is not None else default_device()
        if self.device.type != "cuda":
            raise RuntimeError("FlipSolver3D runs on the HIP device only")
        _lib.lib()  # fail now if the HIP library is missing
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx = dx if dx is not None else 1.0 / max(nx, ny, nz)
        self.dom = tuple(float(np.float32(n * self.dx)) for n in (self.nx, self.ny, self.nz))
        self.gravity = gravity
        self.flip_blend = flip_blend
        self.rho = rho
        self.viscosity = viscosity
        self.viscosity_iters = viscosity_iters
        self.surface_tension = surface_tension
        self.csf_smoothing_passes = max(1, int(csf_smoothing_passes))
        self.transfer_mode = transfer_mode
        # Sub-dense box (F3.7, solver3d.py:2441-2473): the box itself and its
        # rebuild policy live in SubDenseBox; the fields stay full-size.
        self._sub = SubDenseBox(self.nx, self.ny, self.nz, enable_sub_dense, sub_rebuild_every,
                                sub_dilation)
        if enable_sub_dense and int(sub_dilation) < 1:
            logging.getLogger("gpufluid.solver").warning(
                "FlipSolver3D: enable_sub_dense=True with sub_dilation=%d: fluid can reach the "
                "box edge, where the viscosity stencil is cut off; use it for bookkeeping only",
                int(sub_dilation))
        self.pcg_precision = pcg_precision
        self.pcg_sign = pcg_sign
        self.pcg_tol = float(pcg_tol)
        self.pcg_check_every = int(pcg_check_every)
        self._prof = StepProfiler(enabled=enable_timing, device=self.device)
        # B5.2 graph cache: captured substeps keyed by everything their
        # launches bake in (see _graph_key); ping-ponged buffers need two
        self._enable_cuda_graphs = bool(enable_cuda_graphs)
        self._graphs = {}
        self._cuda_graph = None
        self._cuda_graph_key = None
        self._cuda_graph_hits = 0
        self._cuda_graph_misses = 0
        self._graph_stats = None      # (solver, iters) of a replay whose stats are unread

        dev = self.device
        f32 = dict(dtype=torch.float32, device=dev)
        shp_u, shp_v, shp_w = (nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)
        self._f = {
            "u": torch.zeros(shp_u, **f32), "v": torch.zeros(shp_v, **f32),
            "w": torch.zeros(shp_w, **f32),
            "us": torch.zeros(shp_u, **f32), "vs": torch.zeros(shp_v, **f32),
            "ws": torch.zeros(shp_w, **f32),
            "solid_u": torch.zeros(shp_u, **f32), "solid_v": torch.zeros(shp_v, **f32),
            "solid_w": torch.zeros(shp_w, **f32),
        }
        m = np.zeros((nx, ny, nz), dtype=np.int32)
        m[0, :, :] = 2; m[-1, :, :] = 2
        m[:, 0, :] = 2; m[:, -1, :] = 2
        m[:, :, 0] = 2; m[:, :, -1] = 2
        self._marker_host = m
        self._wall_marker = m.copy()
        self._f["marker"] = torch.as_tensor(m, device=dev).contiguous()
        self._uvw_weights = None
        self._sc = SceneBuffers(self.nx, self.ny, self.nz, self.dx, self.dom, dev)
        self.n_particles = 0
        self._fluid_cells = 0
        self._pressure_stats = (0, None, 0)   # iterations, residual, replacements
        self.pcg_kernel_timing = False      # per-kernel HIP-event timing of PCG launches
        # (dirspmv, update: ms summed over sampled iterations; sampled iterations;
        #  replacements: ms summed), see gf_scene_pcg_sync
        self.last_pcg_kernel_ms = None
        self._cache = {}
        self._pcg_done = DeviceArray(torch.zeros(1, dtype=torch.int32, device=dev))
        self._pcg_tol_sq = DeviceArray(torch.zeros(1, dtype=torch.float32, device=dev))
        self._static_obstacle_sdfs: list = []
        # per-particle attributes in particle-id (caller) order, None = absent
        self._attr = {"color": None, "temperature": None}
        self._attr_ws = None
        self._csf_ws = None
        # P2G form: True = each tile scatters its own particles and the tiles'
        # partial sums are merged in a second pass (gf_scene_p2g_staged);
        # False = one pass in which every tile also reads its halo's particles
        self.p2g_staged = True
        self._p2g_ws = None
        self._na_pinned = None
        # inflow/outflow regions drained by prepare_frame (D4.7 / F3.6)
        self.inflows: list = []
        self.outflows: list = []
        self._rng = np.random.default_rng(0)
        self._frame_events = FrameEventQueue()
        self._anim_specs: list = []

    # ------------------------------------------------------------ attributes
    def _arr(self, name):
        return DeviceArray(self._f[name])

    def _boxed(self, t, axis):
        """t, or its sub-dense box part (a view) while the box is smaller
        than the grid: the reference's box-shaped arrays (solver3d.py:3858-3879)."""
        return DeviceArray(t[self._sub.face_slices(axis)] if self._sub.active else t)

    u = property(lambda self: self._boxed(self._f["u"], 0))
    v = property(lambda self: self._boxed(self._f["v"], 1))
    w = property(lambda self: self._boxed(self._f["w"], 2))
    us = property(lambda self: self._boxed(self._f["us"], 0))
    vs = property(lambda self: self._boxed(self._f["vs"], 1))
    ws = property(lambda self: self._boxed(self._f["ws"], 2))
    solid_u = property(lambda self: self._arr("solid_u"))
    solid_v = property(lambda self: self._arr("solid_v"))
    solid_w = property(lambda self: self._arr("solid_w"))
    marker = property(lambda self: self._arr("marker"))

    def _weights(self):
        """uw/vw/ww: scratch of the block-level P2G (the fused substep keeps
        the weights in LDS and never writes them)."""
        if self._uvw_weights is None:
            f = self._f
            self._uvw_weights = tuple(torch.zeros_like(f[k]) for k in ("u", "v", "w"))
        return self._uvw_weights

    uw = property(lambda self: self._boxed(self._weights()[0], 0))
    vw = property(lambda self: self._boxed(self._weights()[1], 1))
    ww = property(lambda self: self._boxed(self._weights()[2], 2))

    def _dense(self, key, compact):
        t = self._cache.get(key)
        if t is None:
            t = torch.zeros((self.nx, self.ny, self.nz), dtype=torch.float32, device=self.device)
            s = self._descriptor()
            _lib.call("F3.3", "gf_scene_compact_to_dense", ctypes.byref(s), _lib.ptr(compact),
                      _lib.ptr(t), self._stream())
            self._cache[key] = t
        return self._boxed(t, None)

    @property
    def p(self):
        """Pressure of the last substep as a dense (nx,ny,nz) array."""
        return self._dense("p", self._sc.pc)

    @property
    def p_tmp(self):
        return self._dense("p_tmp", self._sc.ptmp)

    @property
    def div(self):
        """Divergence (Poisson right-hand side) of the last substep, dense."""
        return self._dense("div", self._sc.b)

    def _live(self, key):
        if self.n_particles == 0 and self._sc.bufs[self._sc.cur] is None:
            return None
        return self._sc.live(self.n_particles)[key]

    _pos = property(lambda self: self._live("pos"))
    _vel = property(lambda self: self._live("vel"))
    _pid = property(lambda self: self._live("pid"))
    _affine = property(lambda self: self._live("affine"))

    def _unsorted(self, key, t, width):
        if t is None:
            return None
        out = self._cache.get(key)
        if out is None:
            out = torch.empty((self.n_particles, width), dtype=torch.float32, device=self.device)
            s = self._descriptor()
            _lib.call("F3.3", "gf_scene_unsort", ctypes.byref(s), _lib.ptr(t), width,
                      _lib.ptr(out), self._stream())
            self._cache[key] = out
        return out

    @property
    def pos(self):
        t = self._unsorted("pos", self._pos, 3)
        return None if t is None else DeviceArray(t, vec=1)

    @pos.setter
    def pos(self, value):
        """Replace the particle positions, as the reference's callers assign
        ``solver.pos`` (cli/commands.py:646-655): velocities become zero and
        affine matrices zero until assigned; attributes are kept when their
        length still matches, else dropped until assigned."""
        t = None if value is None else _rows(value, 3, self.device)
        n = 0 if t is None else t.shape[0]
        keep = {k: (a if a is not None and a.shape[0] == n else None)
                for k, a in self._attr.items()}
        self.set_particles(t if t is not None else torch.zeros((0, 3), device=self.device),
                           color=keep["color"], temperature=keep["temperature"])

    @property
    def vel(self):
        t = self._unsorted("vel", self._vel, 3)
        return None if t is None else DeviceArray(t, vec=1)

    @vel.setter
    def vel(self, value):
        """Replace the velocities (caller order, same particle count)."""
        self._set_rows("vel", value, 3)

    @property
    def affine_C(self):
        t = self._unsorted("affine", self._affine, 9)
        return None if t is None else DeviceArray(t.view(-1, 3, 3), vec=2)

    @affine_C.setter
    def affine_C(self, value):
        """Replace the APIC matrices (caller order); None zeroes them, as the
        reference's step rebuilds a dropped affine_C as zeros."""
        if self.n_particles == 0:
            return
        if value is None:
            a = self._sc.live(self.n_particles)["affine"]
            if a is not None:
                a.zero_()
            self._cache.clear()
            return
        self._sc.ensure_particles(self.n_particles, True)
        self._set_rows("affine", value, 9)

    def _set_rows(self, key, value, width):
        n = self.n_particles
        t = _rows(value, width, self.device)
        if t.shape[0] != n:
            raise ValueError(f"{key} needs {n} rows, got {t.shape[0]}")
        if n == 0:
            return
        live = self._sc.live(n)
        live[key].copy_(t[live["pid"].long()])
        self._cache.clear()

    # ------------------------------------------------- per-particle attributes
    def _get_attr(self, key):
        t = self._attr[key]
        return None if t is None else DeviceArray(t, vec=1 if t.dim() == 2 else 0)

    def _set_attr(self, key, value, n=None):
        if value is None:
            self._attr[key] = None
            return
        n = self.n_particles if n is None else n
        t = as_tensor(value, torch.float32, self.device)
        shape = (n, 3) if key == "color" else (n,)
        if t.numel() != math.prod(shape):
            raise ValueError(f"attr_{key} needs shape {shape}, got {tuple(t.shape)}")
        t = t.reshape(shape)
        self._attr[key] = t.clone().contiguous()

    attr_color = property(lambda self: self._get_attr("color"),
                          lambda self, v: self._set_attr("color", v),
                          doc="Per-particle linear RGB (n,3), advected by the S2.15 transfer "
                              "at the end of each substep; None when absent.")
    attr_temperature = property(lambda self: self._get_attr("temperature"),
                                lambda self, v: self._set_attr("temperature", v),
                                doc="Per-particle scalar (n,), advected by the S2.18 transfer "
                                    "at the end of each substep; None when absent.")

    # ------------------------------------------------------------- particles
    def set_particles(self, pos, vel=None, affine=None, *, color=None, temperature=None) -> None:
        """Replace the particle set (positions (n,3), velocities (n,3) and,
        for APIC, affine matrices (n,3,3); missing ones are zero) together
        with its per-particle attributes (colour (n,3), temperature (n,);
        a missing one is absent afterwards)."""
        dev = self.device
        p = as_tensor(pos, torch.float32, dev).reshape(-1, 3)
        n = p.shape[0]
        self._set_attr("color", color, n)
        self._set_attr("temperature", temperature, n)
        sc = self._sc
        apic = self.transfer_mode == "apic" or affine is not None
        self._cuda_graph_invalidate()
        sc.ensure_particles(n, apic)
        self.n_particles = n
        self._cache.clear()
        if n == 0:
            return
        live = sc.live(n)
        live["pos"].copy_(p)
        if vel is None:
            live["vel"].zero_()
        else:
            live["vel"].copy_(as_tensor(vel, torch.float32, dev).reshape(n, 3))
        live["pid"].copy_(torch.arange(n, dtype=torch.int32, device=dev))
        if live["affine"] is not None:
            if affine is None:
                live["affine"].zero_()
            else:
                live["affine"].copy_(as_tensor(affine, torch.float32, dev).reshape(n, 9))

    @block("D4.5.1", "Box seeder (uniform jittered, ppc particles per cell)")
    def seed_box(self, lo, hi, ppc: int = 8, color=None, temperature=None):
        """Seed [lo, hi] as FlipSolver3D.seed_box (solver3d.py:2827-2910).

        Append-or-replace as there: when this call or the particles already
        present carry a colour or temperature, the new particles are appended
        (previous particles keep position and velocity; every affine matrix
        is reset to zero); otherwise the particle set is replaced.  New
        particles take ``color`` (an RGB triple) / ``temperature``; without
        one they are padded with white / 0.  Earlier particles without the
        attribute are padded the same way (the reference leaves that mix
        unsupported)."""
        new = seed_box_positions(lo, hi, ppc, self.dx)
        n_prev = self.n_particles
        have_attrs = (color is not None or temperature is not None
                      or self._attr["color"] is not None or self._attr["temperature"] is not None)
        if not (have_attrs and n_prev > 0):
            n_prev = 0
        if n_prev:
            pos = np.concatenate([self.pos.numpy(), new], axis=0)
            vel = np.concatenate([self.vel.numpy(), np.zeros_like(new)], axis=0)
        else:
            pos, vel = new, np.zeros_like(new)
        n_new = len(new)

        def merged(key, value, pad, width):
            prev = self._attr[key]
            if value is None and prev is None:
                return None
            fill = np.float32(pad) if value is None else np.asarray(value, np.float32)
            add = np.broadcast_to(fill, (n_new,) + width).astype(np.float32)
            if not n_prev:
                return add
            old = (prev.cpu().numpy() if prev is not None
                   else np.full((n_prev,) + width, pad, np.float32))
            return np.concatenate([old, add], axis=0)

        col = merged("color", color, 1.0, (3,))
        temp = merged("temperature", temperature, 0.0, ())
        self.set_particles(pos, vel, color=col, temperature=temp)

    def get_particles(self):
        if self._pos is None:
            return None, None
        return self.pos.numpy(), self.vel.numpy()

    # ------------------------------------------------------------ checkpoint
    @block("F3.5", "Save solver state to .npz (resumable)")
    def save_checkpoint(self, path) -> None:
        """Write the solver state in the reference's .npz layout
        (solver3d.py:4127-4158): particles in caller order, the marker grid
        (walls and obstacles included), the grid and physics parameters, and
        the per-particle colour/temperature (empty arrays when absent)."""
        from pathlib import Path
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        if self.n_particles:
            pos, vel = self.pos.numpy(), self.vel.numpy()
        else:
            pos = vel = np.zeros((0, 3), dtype=np.float32)
        aff = self.affine_C
        affine = (aff.numpy().reshape(-1, 3, 3) if aff is not None and self.n_particles
                  else np.zeros((0, 3, 3), dtype=np.float32))
        np.savez_compressed(
            str(path),
            pos=np.asarray(pos, np.float32), vel=np.asarray(vel, np.float32),
            affine_C=np.asarray(affine, np.float32), marker=self.marker.numpy().astype(np.int32),
            attr_color=(self.attr_color.numpy() if self._attr["color"] is not None
                        else np.zeros((0, 3), dtype=np.float32)),
            attr_temperature=(self.attr_temperature.numpy()
                              if self._attr["temperature"] is not None
                              else np.zeros((0,), dtype=np.float32)),
            nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx,
            gravity=self.gravity, flip_blend=self.flip_blend, rho=self.rho,
            viscosity=self.viscosity, transfer_mode=str(self.transfer_mode),
            surface_tension=self.surface_tension,
        )

    @block("F3.5", "Load checkpoint into a fresh solver (state-only; topology must match)")
    def load_checkpoint(self, path) -> None:
        """Restore a checkpoint written by either implementation
        (solver3d.py:4160-4222): the grid shape and dx must match (else
        RuntimeError); drift of gravity/rho/viscosity/flip_blend/
        surface_tension is logged and the live values are kept.  Empty
        colour/temperature arrays mean the attribute was absent; a length
        that does not match the particle count raises RuntimeError."""
        import logging
        data = np.load(str(path), allow_pickle=False)
        nx, ny, nz = int(data["nx"]), int(data["ny"]), int(data["nz"])
        if (nx, ny, nz) != (self.nx, self.ny, self.nz):
            raise RuntimeError(f"checkpoint grid {(nx, ny, nz)} != solver {(self.nx, self.ny, self.nz)}")
        ckpt_dx = float(data["dx"])
        if abs(ckpt_dx - self.dx) > 1e-9:
            raise RuntimeError(
                f"checkpoint dx={ckpt_dx} != solver dx={self.dx} "
                f"(same grid resolution but different domain size — "
                f"particle world-coords would land in wrong cells)")
        for key, live in (("gravity", self.gravity), ("rho", self.rho),
                          ("viscosity", self.viscosity), ("flip_blend", self.flip_blend),
                          ("surface_tension", self.surface_tension)):
            if key in data.files:
                saved = float(data[key])
                if abs(saved - float(live)) > 1e-9:
                    logging.getLogger("gpufluid.checkpoint").warning(
                        "resume: %s drift saved=%s live=%s — running with LIVE value "
                        "(resume is no longer bit-exact).", key, saved, live)
        pos, vel, affine = data["pos"], data["vel"], data["affine_C"]
        attrs = {}
        for key in ("attr_color", "attr_temperature"):
            if key not in data.files:
                # an older checkpoint without the key: the reference keeps the
                # live attribute (solver3d.py:4200-4214); here only while it
                # still has one entry per restored particle
                live = self._attr[key[5:]]
                attrs[key] = live if live is not None and len(live) == len(pos) else None
                continue
            a = data[key]
            if len(a) == 0:
                a = None
            if a is not None and len(a) != len(pos):
                raise RuntimeError(f"checkpoint {key} holds {len(a)} entries for {len(pos)} "
                                   "particles")
            attrs[key] = a
        marker = data["marker"].astype(np.int32)
        self.set_particles(pos, vel, affine if len(affine) else None,
                           color=attrs["attr_color"], temperature=attrs["attr_temperature"])
        self._marker_host = marker.copy()
        self._f["marker"].copy_(torch.as_tensor(marker, device=self.device))
        self._cache.clear()
        self._cuda_graph_invalidate()

    # ------------------------------------------------- frame hook (F3.6, D4.7)
    @block("D4.7", "Register an inflow region (continuous particle emission)")
    def add_inflow(self, region):
        self.inflows.append(region)

    @block("D4.7", "Register an outflow region (particle removal)")
    def add_outflow(self, region):
        self.outflows.append(region)

    @block("D4.6", "Register an analytic obstacle that moves per frame")
    def add_animated_obstacle(self, kind: str, base_center, motion, **params):
        """kind: 'sphere' (radius) | 'box' (half_size) | 'cylinder_y' (radius,
        half_height) | 'mesh' (mesh_path to an OBJ or STL file, optional scale
        and translate; or the triangles themselves as ``triangles`` (n,3,3)),
        moving by a LinearMotion / KeyframeMotion (solver3d.py:2516-2519).  A
        mesh moves by its centre's displacement from ``base_center``
        (solver3d.py:2604-2612); its file is read once, here."""
        if kind not in ("sphere", "box", "cylinder_y", "mesh"):
            raise ValueError(f"unsupported animated obstacle kind: {kind!r}")
        spec = {"kind": kind, "base_center": tuple(base_center), "motion": motion,
                "params": dict(params)}
        if kind == "mesh":
            from .meshes import read_mesh
            if "triangles" in params:
                tri = np.asarray(params["triangles"], dtype=np.float64).reshape(-1, 3, 3)
            elif "mesh_path" in params:
                tri = read_mesh(params["mesh_path"])
            else:
                raise ValueError("a mesh obstacle needs mesh_path or triangles")
            # FP64 vertices times scale; the frame's translation is added in
            # stamping, then the triangles are rounded to FP32
            spec["mesh_base"] = tri * float(params.get("scale", 1.0))
        self._anim_specs.append(spec)

    def _obstacle_velocity_at(self, spec, frame_idx: int):
        """Linear motion: its velocity; keyframes: centred difference of the
        centre over one frame at 24 fps (solver3d.py:2521-2535)."""
        from .animation import evaluate_center
        m = spec.get("motion")
        if m is None:
            return (0.0, 0.0, 0.0)
        if getattr(m, "kind", None) == "linear":
            return tuple(float(c) for c in m.velocity)
        if getattr(m, "kind", None) == "keyframes":
            c1 = evaluate_center(spec["base_center"], m, frame_idx + 1)
            c0 = evaluate_center(spec["base_center"], m, frame_idx - 1)
            fps = m.fps if hasattr(m, "fps") else 24
            return tuple(float(c) for c in (c1 - c0) * fps / 2.0)
        return (0.0, 0.0, 0.0)

    def _stamp_animated_obstacles(self, frame_idx: int):
        """Marker = walls + static obstacles + this frame's animated
        obstacles, and their velocities on the solid faces, on the device
        (solver3d.py:3918-3940)."""
        from .animation import evaluate_center
        base = self._wall_marker.copy()
        for sdf in self._static_obstacle_sdfs:
            base[sdf <= 0.0] = 2
        rows = []
        meshes = []
        kinds = {"sphere": 0, "box": 1, "cylinder_y": 2}
        for spec in self._anim_specs:
            c = evaluate_center(spec["base_center"], spec["motion"], frame_idx)
            p = spec["params"]
            if spec["kind"] == "mesh":
                # translate + (centre - base centre), in FP32 (solver3d.py:2609-2612)
                delta = np.asarray(c, np.float32) - np.asarray(spec["base_center"], np.float32)
                t_total = np.asarray(p.get("translate", (0.0, 0.0, 0.0)), np.float32) + delta
                meshes.append((spec, t_total, self._obstacle_velocity_at(spec, frame_idx)))
                continue
            if spec["kind"] == "sphere":
                par = (p["radius"], 0.0, 0.0)
            elif spec["kind"] == "box":
                par = tuple(np.broadcast_to(np.asarray(p["half_size"], np.float32), (3,)))
            else:
                par = (p["radius"], p["half_height"], 0.0)
            vel = self._obstacle_velocity_at(spec, frame_idx)
            rows.append([kinds[spec["kind"]], *[float(x) for x in c], *map(float, par), *vel])
        dev = self.device
        obs = torch.tensor(np.asarray(rows, np.float32).reshape(-1, 10), device=dev)
        base_t = torch.as_tensor(base, device=dev)
        f = self._f
        _lib.call("D4.6", "gf_anim_obstacles", _lib.ptr(obs), len(rows), _lib.ptr(base_t),
                  _lib.ptr(f["marker"]), _lib.ptr(f["solid_u"]), _lib.ptr(f["solid_v"]),
                  _lib.ptr(f["solid_w"]), self.nx, self.ny, self.nz, float(self.dx),
                  self._stream())
        if meshes:
            # after the analytic ones, in list order: each mesh's inside cells
            # not yet solid, and its velocity on their faces (2598-2631)
            from .meshes import mark_mesh, place
            faces = (f["solid_u"], f["solid_v"], f["solid_w"])
            for spec, t_total, vel in meshes:
                mark_mesh(f["marker"], place(spec["mesh_base"], t_total), self.dx, faces, vel,
                          self._stream())
        self._marker_host = f["marker"].cpu().numpy()

    @block("F3.6", "Per-frame hook: rebuild marker for anim obstacles, emit inflow particles, "
                   "drop outflow particles")
    def prepare_frame(self, frame_idx: int, frame_dt: float):
        """prepare_frame (solver3d.py:3911-4072): animated obstacles are
        stamped into the marker and the solid face velocities; then regions
        publish this frame's events; outflow boxes are compacted away on the
        device first, then the emitted particles (minus those landing in a
        solid cell of the host marker) are appended with their event's colour
        (else 0) and temperature (else 20) when the attribute is tracked.
        Last, with enable_sub_dense, the box is rebuilt when due
        (solver3d.py:4050-4072)."""
        # the frame may change the particle set and the solids: drop the
        # captured substeps, as the reference does (solver3d.py:3912-3915)
        self._cuda_graph_invalidate()
        if self._anim_specs:
            self._stamp_animated_obstacles(frame_idx)
        q = self._frame_events
        q.clear()
        for inflow in self.inflows:
            inflow.publish_for_frame(q, frame_idx, frame_dt, self._rng)
        for outflow in self.outflows:
            outflow.publish_for_frame(q, frame_idx)
        out_ev = q.drain_outflows()
        if out_ev:
            self._apply_outflows_gpu(out_ev)
        emits = q.drain_emits()
        if emits:
            self._emit_particles(emits)
        if self._sub.enabled and self._should_rebuild_sub_dense(frame_idx):
            # the marker of the last substep's P2G: before any step it holds
            # only walls and obstacles and gives no box
            bbox = self._compute_active_bbox(self._f["marker"].cpu().numpy())
            if bbox is not None:
                self._rebuild_sub_dense(*bbox)
                self._last_sub_rebuild_frame = frame_idx
                self._check_particles_in_sub_bbox(frame_idx)

    def _emit_particles(self, emits):
        """Append the frame's emitted particles (see prepare_frame)."""
        epos = np.concatenate([e.positions for e in emits]).astype(np.float32)
        evel = np.concatenate([e.velocities for e in emits]).astype(np.float32)
        col = temp = None
        if self._attr["color"] is not None:
            col = np.concatenate([
                np.broadcast_to(np.asarray(e.color if e.color is not None else (0.0, 0.0, 0.0),
                                           np.float32), (len(e.positions), 3))
                for e in emits])
        if self._attr["temperature"] is not None:
            temp = np.concatenate([
                np.full(len(e.positions), 20.0 if e.temperature is None else e.temperature,
                        np.float32) for e in emits])
        if len(epos):
            mh = self._marker_host
            idx = [np.clip((epos[:, a] / self.dx).astype(np.int32), 0, mh.shape[a] - 1)
                   for a in range(3)]
            keep = mh[idx[0], idx[1], idx[2]] != 2
            if not keep.all():
                epos, evel = epos[keep], evel[keep]
                col = None if col is None else col[keep]
                temp = None if temp is None else temp[keep]
        if not len(epos):
            return
        dev = self.device
        cur_p = self.pos.t if self.n_particles else torch.zeros((0, 3), device=dev)
        cur_v = self.vel.t if self.n_particles else torch.zeros((0, 3), device=dev)
        cat = lambda a, b: torch.cat([a, torch.as_tensor(np.ascontiguousarray(b),  # noqa: E731
                                                         device=dev)])
        new_col = None if col is None else cat(self._attr["color"], col)
        new_temp = None if temp is None else cat(self._attr["temperature"], temp)
        self.set_particles(cat(cur_p, epos), cat(cur_v, evel), color=new_col,
                           temperature=new_temp)

    # ------------------------------------------------ F3.7 sub-dense box
    _sub_offset = property(lambda self: self._sub.offset)
    _sub_shape = property(lambda self: self._sub.shape)
    _enable_sub_dense = property(lambda self: self._sub.enabled)
    _sub_dilation = property(lambda self: self._sub.dilation,
                             lambda self, d: setattr(self._sub, "dilation", max(0, int(d))))
    _sub_rebuild_every = property(lambda self: self._sub.rebuild_every,
                                  lambda self, n: setattr(self._sub, "rebuild_every",
                                                          max(1, int(n))))
    _last_sub_rebuild_frame = property(
        lambda self: self._sub.last_rebuild_frame,
        lambda self, f: setattr(self._sub, "last_rebuild_frame", int(f)))
    sub_dilation = _sub_dilation
    sub_rebuild_every = _sub_rebuild_every

    @block("F3.7", "Compute the 8^3-tile bbox of active fluid cells, dilated by sub_dilation "
                   "and clamped to the domain.")
    def _compute_active_bbox(self, marker_host: np.ndarray):
        """((lo), (hi)) cells of the dilated fluid-tile box, or None
        (solver3d.py:3670-3702)."""
        return self._sub.bbox(marker_host)

    @block("F3.7", "Map a world-space particle position to local sub-dense cell coords + flag "
                   "whether it lies inside the bbox.")
    def _pos_to_sub_cell(self, pos):
        """((li, lj, lk), inside) of one position (solver3d.py:3704-3726)."""
        return self._sub.pos_to_cell(pos, self.dx)

    def _check_particles_in_sub_bbox(self, frame_idx: int) -> int:
        """Count of particles outside the box, with a one-shot warning
        (solver3d.py:3728-3768)."""
        if not self._sub.enabled or self.n_particles == 0:
            return 0
        return self._sub.count_outside(self.pos.numpy(), self.dx, frame_idx)

    def _should_rebuild_sub_dense(self, frame_idx: int) -> bool:
        """Periodic / first-call / edge-proximity trigger (solver3d.py:3770-3819).
        The host marker is refreshed from the device when the device one holds
        fluid (a marker set by hand on the host is kept otherwise)."""
        if not self._sub.enabled:
            return False
        if self._sub.last_rebuild_frame >= 0 and \
                frame_idx - self._sub.last_rebuild_frame < self._sub.rebuild_every:
            dev_marker = self._f["marker"].cpu().numpy()
            if np.any(dev_marker == 1):
                self._marker_host = dev_marker
        return self._sub.should_rebuild(frame_idx, self._marker_host)

    @block("F3.7", "Move the sub-dense box to [lo, hi); grid attributes show the box and the "
                   "viscosity sweep runs over it.")
    def _rebuild_sub_dense(self, new_lo, new_hi):
        """Box bookkeeping of solver3d.py:3821-3906.  Storage is full-size, so
        nothing is copied: the box's cells keep their global values."""
        self._sub.rebuild(new_lo, new_hi)
        self._cuda_graph_invalidate()

    @block("D4.7.GPU", "GPU outflow compaction (mark + scan + scatter)")
    def _apply_outflows_gpu(self, outflow_events) -> int:
        """Remove the particles inside any active outflow box, keeping the
        survivors' order (solver3d.py:2659-2730): mark, scan and scatter on
        the device, attributes in step; the APIC matrices restart at zero
        when particles were removed, as the reference drops them."""
        n = self.n_particles
        if n == 0:
            return 0
        active = [ev for ev in outflow_events if ev.inside]
        if not active:
            return n
        dev = self.device
        boxes = torch.tensor([list(ev.lo) + list(ev.hi) for ev in active], dtype=torch.float32,
                             device=dev)
        pos = self.pos.t
        alive = torch.empty(n, dtype=torch.int32, device=dev)
        prefix = torch.empty(n + 1, dtype=torch.int32, device=dev)
        part = torch.empty((n + 4095) // 4096 + 1, dtype=torch.int32, device=dev)
        n_alive = ctypes.c_int()
        st = self._stream()
        _lib.call("D4.7.GPU", "gf_outflow_mark_sync", _lib.ptr(pos), n, _lib.ptr(boxes),
                  len(active), _lib.ptr(alive), _lib.ptr(prefix), _lib.ptr(part),
                  ctypes.byref(n_alive), st)
        m = n_alive.value
        if m == n:
            return n

        def compact(t, width):
            out = torch.empty((m, width) if width > 1 else (m,), dtype=torch.float32, device=dev)
            _lib.call("D4.7.GPU", "gf_compact_rows", _lib.ptr(t.contiguous()), width, n,
                      _lib.ptr(alive), _lib.ptr(prefix), _lib.ptr(out), st)
            return out

        vel = self.vel.t
        col = None if self._attr["color"] is None else compact(self._attr["color"], 3)
        temp = None if self._attr["temperature"] is None else compact(self._attr["temperature"], 1)
        self.set_particles(compact(pos, 3), compact(vel, 3), color=col, temperature=temp)
        return m

    def has_diverged(self) -> bool:
        """NaN probe on particle positions (solver3d.py:2642-2656)."""
        if self._pos is None or self.n_particles == 0:
            return False
        return bool(torch.isnan(self._pos).any().item())

    @block("D4.4", "Inject SDF obstacle into marker grid")
    def add_solid_from_sdf(self, sdf, padding: float = 0.0):
        m = self._marker_host
        m[np.asarray(sdf) <= padding] = 2
        self._f["marker"].copy_(torch.as_tensor(m, device=self.device))
        self._static_obstacle_sdfs.append(np.asarray(sdf, dtype=np.float32))

    def cell_centers_np(self):
        i = (np.arange(self.nx) + 0.5) * self.dx
        j = (np.arange(self.ny) + 0.5) * self.dx
        k = (np.arange(self.nz) + 0.5) * self.dx
        X, Y, Z = np.meshgrid(i, j, k, indexing="ij")
        return np.stack([X, Y, Z], axis=-1).astype(np.float32)

    # --------------------------------------------------------------- substep
    def _stream(self):
        return _lib.stream_handle(self.device)

    def _descriptor(self):
        return self._sc.descriptor(self._f, self.n_particles, self.transfer_mode == "apic")

    def _viscosity(self, dt):
        """S2.13 as _step_impl runs it (solver3d.py:3524-3544): the post-BC
        faces are copied into the weight arrays as the right-hand side, then
        `viscosity_iters` Jacobi sweeps per component ping-pong between u and
        us (so us ends holding the last-but-one iterate, which FLIP then
        samples as the old velocity, as in the reference), then the solid
        BC."""
        f = self._f
        r = float(dt * self.viscosity / (self.dx * self.dx))
        rhs = self._weights()
        for t, key in zip(rhs, ("u", "v", "w")):
            t.copy_(f[key])
        box = self._sub
        if box.active and int(self.viscosity_iters) > 0:
            # The sweep covers the box only (its faces, the box as the stencil
            # bounds, as the reference's k3_jacobi_visc on its box-shaped
            # arrays); both ping-pong buffers hold u outside it.
            for key, sav in (("u", "us"), ("v", "vs"), ("w", "ws")):
                f[sav].copy_(f[key])
        for _ in range(int(self.viscosity_iters)):
            bufs = (*(_lib.ptr(t) for t in rhs),
                    *(_lib.ptr(f[k]) for k in ("u", "v", "w", "us", "vs", "ws")))
            if box.active:
                _lib.call("S2.13", "gf_k3_jacobi_visc3_box", *bufs, self.nx, self.ny, self.nz,
                          *box.offset, *box.shape, r, self._stream())
            else:
                _lib.call("S2.13", "gf_k3_jacobi_visc3", *bufs, self.nx, self.ny, self.nz, r,
                          self._stream())
            for key, sav in (("u", "us"), ("v", "vs"), ("w", "ws")):
                f[key], f[sav] = f[sav], f[key]
        self._solid_bc()

    def _p2g_stage(self):
        if self._p2g_ws is None:
            nf, nt = ctypes.c_int64(), ctypes.c_int64()
            _lib.call("S2.1", "gf_scene_p2g_stage_size", self.nx, self.ny, self.nz,
                      ctypes.byref(nf), ctypes.byref(nt))
            self._p2g_ws = (torch.empty(nf.value, dtype=torch.float32, device=self.device),
                            torch.empty(nt.value, dtype=torch.int32, device=self.device))
        return self._p2g_ws

    def _solid_bc(self):
        f = self._f
        _lib.call("S2.4", "gf_k3_enforce_solid_bc", *(_lib.ptr(f[k]) for k in
                  ("u", "v", "w", "marker", "solid_u", "solid_v", "solid_w")),
                  self.nx, self.ny, self.nz, self._stream())

    def _surface_tension(self, dt):
        """S2.14 as _step_impl runs it (solver3d.py:3548-3555, 3230-3300): the
        CSF impulse with force balance on u, v, w, then the solid BC."""
        work, red = self._csf_workspace()
        coeff = float(self.surface_tension * dt / max(self.rho, 1.0e-8))
        f = self._f
        _lib.call("S2.14", "gf_csf_apply", *(_lib.ptr(f[k]) for k in ("u", "v", "w", "marker")),
                  self.nx, self.ny, self.nz, float(self.dx), coeff,
                  int(self.csf_smoothing_passes), _lib.ptr(work), _lib.ptr(red), self._stream())
        self._solid_bc()

    def _csf_workspace(self):
        if self._csf_ws is None:
            wf, rd = ctypes.c_int64(), ctypes.c_int64()
            _lib.call("S2.14", "gf_csf_workspace", self.nx, self.ny, self.nz, ctypes.byref(wf),
                      ctypes.byref(rd))
            self._csf_ws = (torch.empty(wf.value, dtype=torch.float32, device=self.device),
                            torch.empty(rd.value, dtype=torch.float64, device=self.device))
        return self._csf_ws

    def _attr_workspace(self):
        if self._attr_ws is None:
            cells = self.nx * self.ny * self.nz
            dev = self.device
            self._attr_ws = (torch.empty(4 * cells, dtype=torch.int64, device=dev),
                             torch.empty(4 * cells, dtype=torch.float32, device=dev),
                             torch.empty(4, dtype=torch.int32, device=dev))
        return self._attr_ws

    def _attr_transfer(self, s, attr, channels, mode, stream):
        """S2.15 / S2.18: scatter, normalise and gather one attribute at the
        advected positions (solver3d.py:3304-3359)."""
        if attr.shape[0] != self.n_particles:
            raise RuntimeError(f"attribute holds {attr.shape[0]} particles, the solver "
                               f"{self.n_particles}")
        acc, grid, scratch = self._attr_workspace()
        _lib.call("S2.15" if mode == _lib.GF_ATTR_COLOR else "S2.18", "gf_scene_attr_transfer",
                  ctypes.byref(s), _lib.ptr(attr), channels, mode, _lib.ptr(acc),
                  _lib.ptr(grid), _lib.ptr(scratch), stream)

    @block("F3.3", "Per-step pipeline: sort, fused P2G, pressure, fused G2P")
    def step(self, dt: float, pressure_iters: int = 80, pressure_solver: str = "jacobi",
             pressure_block_sparse: bool = False, *, pressure_tol: float | None = None):
        """Advance one substep of length ``dt`` (solver3d.py:3422-3667).
        ``pressure_iters`` is the Jacobi / GS-RB sweep count or the PCG
        iteration cap.  ``pressure_block_sparse`` is accepted for
        compatibility: the pressure solve always runs on the active 8^3 tiles
        only, which is what the reference's per-tile variants compute."""
        if pressure_solver not in ("jacobi", "pcg", "gsrb"):
            if self._sub.active:  # the reference's sub-dense guard (solver3d.py:3429-3440)
                raise NotImplementedError(
                    "FlipSolver3D.step() with a sub-dense box rejects: "
                    f"pressure_solver={pressure_solver!r}; disable enable_sub_dense or use "
                    "'jacobi', 'gsrb' or 'pcg'")
            raise ValueError(f"unknown pressure_solver: {pressure_solver!r}")
        tol = self.pcg_tol if pressure_tol is None else float(pressure_tol)
        if self._cuda_graph_eligible(pressure_solver, pressure_block_sparse):
            self._step_graphed(float(dt), int(pressure_iters), pressure_solver, tol)
        else:
            self._step_impl(float(dt), int(pressure_iters), pressure_solver, tol)

    # ------------------------------------------------------- B5.2 graph cache
    _GRAPH_CACHE = 4   # captured substeps kept (a configuration takes one per ping-pong state)

    def _cuda_graph_eligible(self, pressure_solver: str, pressure_block_sparse: bool = False) -> bool:
        """Substeps replayed from a hipGraph (solver3d.py:3362-3391): with
        ``enable_cuda_graphs``, every solver and stage combination once
        particles exist, except while per-kernel PCG timing needs host
        events.  The stage profiler's sections record nothing while a
        substep is captured or replayed, as the reference's do
        (profiling.py:51-56)."""
        return (self._enable_cuda_graphs and self.n_particles > 0
                and pressure_solver in ("jacobi", "gsrb", "pcg")
                and not self.pcg_kernel_timing)

    def _cuda_graph_invalidate(self):
        """Drop every captured substep (solver3d.py:3414-3419)."""
        self._graphs.clear()
        self._cuda_graph = None
        self._cuda_graph_key = None

    def _ensure_workspaces(self, pressure_solver):
        """Allocate every buffer a substep may use, so that none is
        allocated while a substep is being captured."""
        self._sc.ensure_particles(self.n_particles, self.transfer_mode == "apic")
        if self.p2g_staged:
            self._p2g_stage()
        if self.viscosity > 0.0 or self.surface_tension > 0.0:
            self._weights()
        if self.surface_tension > 0.0:
            self._csf_workspace()
        if self._attr["color"] is not None or self._attr["temperature"] is not None:
            self._attr_workspace()
        if pressure_solver == "pcg" and self.pcg_precision == "fp64":
            self._sc.ensure_fp64()

    def _graph_key(self, dt, pressure_iters, pressure_solver, tol):
        """Everything a captured substep bakes into its launches: the
        descriptor (buffer addresses, sizes, particle count), the addresses
        of the buffers held here, the stage switches and constants, and the
        PCG tunables.  Equal keys give identical launch sequences (the
        reference keys on a subset, solver3d.py:3393-3412)."""
        from . import autotune
        d = self._descriptor()
        raw = ctypes.string_at(ctypes.addressof(d), ctypes.sizeof(d))
        held = [self._f[k] for k in sorted(self._f)]
        held += list(self._uvw_weights or ())
        held += [self._attr["color"], self._attr["temperature"]]
        for ws in (self._p2g_ws, self._attr_ws, self._csf_ws):
            held += list(ws or ())
        ptrs = tuple(0 if t is None else t.data_ptr() for t in held)
        return (raw, ptrs, autotune.get_pcg_workgroups(), autotune.get_pcg_variant(),
                autotune.get_pcg_x_every(),
                float(dt), int(pressure_iters), pressure_solver, float(tol), self.transfer_mode,
                self.pcg_precision, self.pcg_sign, int(self.pcg_check_every),
                float(self.gravity), float(self.rho), float(self.flip_blend),
                float(self.viscosity), int(self.viscosity_iters), float(self.surface_tension),
                int(self.csf_smoothing_passes), bool(self.p2g_staged), self._sub.offset,
                self._sub.shape)

    def _step_graphed(self, dt, pressure_iters, pressure_solver, tol):
        """Replay the captured substep for this key, capturing it first on a
        miss (solver3d.py:3436-3469).  Unlike the reference, the call that
        captures also replays, so every call advances the simulation.  The
        pressure statistics and tile counts are read when first asked for."""
        sc = self._sc
        self._ensure_workspaces(pressure_solver)
        key = self._graph_key(dt, pressure_iters, pressure_solver, tol)
        entry = self._graphs.get(key)
        if entry is None:
            entry = self._capture_substep(dt, pressure_iters, pressure_solver, tol)
            self._store_graph(key, entry)
            # The particle buffers (and, with an odd sweep count, the face
            # buffers) ping-pong, so the next substep starts from the other
            # state: capture it now as well, so that one configuration is one
            # miss, as in the reference's single-graph cache.
            self._capture_follower(entry, dt, pressure_iters, pressure_solver, tol)
            self._cuda_graph_misses += 1
        else:
            self._cuda_graph_hits += 1
        graph, fields, cur = entry
        self._cache.clear()
        graph.replay()
        # the host-side state the substep leaves: ping-ponged particle and
        # viscosity buffers
        self._f.update(fields)
        sc.cur = cur
        self._cuda_graph, self._cuda_graph_key = graph, key
        self._graph_stats = (pressure_solver, int(pressure_iters))

    def _store_graph(self, key, entry):
        if len(self._graphs) >= self._GRAPH_CACHE:
            self._graphs.pop(next(iter(self._graphs)))
        self._graphs[key] = entry

    def _capture_follower(self, entry, dt, pressure_iters, pressure_solver, tol):
        """Capture the substep that starts from the state `entry` leaves."""
        start = (dict(self._f), self._sc.cur)
        self._f.update(entry[1])
        self._sc.cur = entry[2]
        try:
            key = self._graph_key(dt, pressure_iters, pressure_solver, tol)
            if key not in self._graphs:
                self._store_graph(key, self._capture_substep(dt, pressure_iters,
                                                             pressure_solver, tol))
        finally:
            self._f.update(start[0])
            self._sc.cur = start[1]

    def _capture_substep(self, dt, pressure_iters, pressure_solver, tol):
        """Record one substep into a hipGraph (nothing runs); returns the
        graph and the host-side buffer state it ends in, and restores the
        state it started from."""
        dev = self.device
        start = (dict(self._f), self._sc.cur)
        g = torch.cuda.CUDAGraph()
        stream = torch.cuda.Stream(device=dev)
        stream.wait_stream(torch.cuda.current_stream(dev))
        try:
            with torch.cuda.stream(stream):
                with torch.cuda.graph(g, stream=stream):
                    self._step_impl(dt, pressure_iters, pressure_solver, tol, capture=True)
            end = (dict(self._f), self._sc.cur)
        finally:
            torch.cuda.current_stream(dev).wait_stream(stream)
            self._f.update(start[0])
            self._sc.cur = start[1]
        return g, end[0], end[1]

    def _stats(self):
        if self._graph_stats is not None:
            solver, iters = self._graph_stats
            self._graph_stats = None
            if solver == "pcg":
                it, res, nrep = (ctypes.c_int * 1)(), (ctypes.c_double * 1)(), (ctypes.c_int * 1)()
                d = self._descriptor()
                _lib.call("F3.3", "gf_scene_pcg_stats", ctypes.byref(d), it, res, nrep,
                          self._stream())
                self._pressure_stats = (int(it[0]), float(res[0]), int(nrep[0]))
            else:
                self._pressure_stats = (iters, None, 0)
            na = self._sc.n_active[:2].cpu()
            self._sc.n_active_host = int(na[0])
            self._fluid_cells = int(na[1])
        return self._pressure_stats

    last_pressure_iters = property(lambda self: self._stats()[0],
                                   doc="PCG iterations (Jacobi / GS-RB: sweeps) of the last substep")
    last_pressure_residual = property(lambda self: self._stats()[1],
                                      doc="PCG true relative residual of the last substep")
    last_pressure_replacements = property(lambda self: self._stats()[2],
                                          doc="FP64 residual replacements of the last PCG solve")

    @property
    def _sc_fluid_cells(self) -> int:
        self._stats()
        return self._fluid_cells

    def _step_impl(self, dt, pressure_iters, pressure_solver, tol, hook=None, capture=False):
        """The substep pipeline.  ``hook(stage)``, if given, is called after
        the sort ("sorted": particles in this substep's cell order) and at the
        end ("done"); differentiable rollouts record forward state there.
        ``capture``: enqueue every launch with no host synchronisation (the
        pressure solvers read the active tile count on the device), for
        recording into a hipGraph; the statistics are then read afterwards."""
        prof = self._prof
        sc = self._sc
        apic = self.transfer_mode == "apic"
        mode = _TRANSFER[self.transfer_mode]
        flip_blend = self.flip_blend
        n = self.n_particles
        stream = self._stream()
        self._cache.clear()
        self._graph_stats = None
        sc.ensure_particles(n, apic)
        # Without a hook the tile counts are read back after P2G is enqueued
        # (an async copy into pinned memory and an event), so the device is
        # not idle while the host reads them.
        deferred = hook is None and not capture
        with prof.section("clear"):
            s = self._descriptor()
            na = (ctypes.c_int * 2)()
            try:
                _lib.call("F3.3", "gf_scene_sort_sync", ctypes.byref(s), int(apic),
                          None if deferred or capture else na, stream)
            except BlockError:
                sc.reset_counts()
                raise
            if deferred:
                if self._na_pinned is None:
                    self._na_pinned = torch.empty(2, dtype=torch.int32, pin_memory=True)
                self._na_pinned.copy_(sc.n_active[:2], non_blocking=True)
                na_ready = torch.cuda.Event()
                na_ready.record()
            if n > 0:
                sc.swap()
            if not deferred and not capture:
                sc.n_active_host = na[0]
                self._fluid_cells = na[1]
        if hook is not None:
            hook("sorted")       # may substitute the masks and tile tables (frozen structure)
        s = self._descriptor()
        with prof.section("p2g"):
            if self.p2g_staged:
                stage, has = self._p2g_stage()
                _lib.call("S2.1", "gf_scene_p2g_staged", ctypes.byref(s), mode, self.gravity, dt,
                          self.rho, _lib.ptr(stage), _lib.ptr(has), stream)
            else:
                _lib.call("S2.1", "gf_scene_p2g", ctypes.byref(s), mode, self.gravity, dt,
                          self.rho, stream)
        if deferred:
            na_ready.synchronize()
            sc.n_active_host = int(self._na_pinned[0])
            self._fluid_cells = int(self._na_pinned[1])
        nact = _lib.GF_NACT_DEVICE if capture else sc.n_active_host
        prof.fused("normalize")
        prof.fused("gravity_bc")
        pre = None
        if self.viscosity > 0.0 or self.surface_tension > 0.0:
            if self.viscosity > 0.0:
                with prof.section("viscosity"):
                    self._viscosity(dt)
            if self.surface_tension > 0.0:
                with prof.section("surface_tension"):
                    self._surface_tension(dt)
            pre = self._weights()
            for t, key in zip(pre, ("u", "v", "w")):
                t.copy_(self._f[key])
            s = self._descriptor()
            with prof.section("divergence"):
                _lib.call("S2.5", "gf_scene_divergence", ctypes.byref(s), dt, self.rho, stream)
        else:
            prof.fused("divergence")
        with prof.section("pressure"):
            if pressure_solver == "jacobi":
                _lib.call("S2.6.1", "gf_scene_jacobi", ctypes.byref(s), nact,
                          pressure_iters, stream)
                self._pressure_stats = (pressure_iters, None, 0)
            elif pressure_solver == "gsrb":
                _lib.call("S2.6.2", "gf_scene_gsrb", ctypes.byref(s), nact,
                          pressure_iters, stream)
                self._pressure_stats = (pressure_iters, None, 0)
            else:
                prec = _lib.GF_PCG_MIXED if self.pcg_precision == "mixed" else _lib.GF_PCG_FP64
                if prec == _lib.GF_PCG_FP64:
                    sc.ensure_fp64()
                    s = self._descriptor()
                sign = -1 if self.pcg_sign == "physical" else 1
                ce = self.pcg_check_every
                if capture:
                    # the whole iteration cap is enqueued; the chunks after
                    # the stop pass through (on-device stop flag)
                    _lib.call("S2.6.3", "gf_scene_pcg_enqueue", ctypes.byref(s), nact, prec, tol,
                              pressure_iters, ce, sign, 0, -(-pressure_iters // ce), stream)
                else:
                    it = ctypes.c_int(0)
                    res = ctypes.c_double(0.0)
                    nrep = ctypes.c_int(0)
                    kms = (ctypes.c_double * 4)() if self.pcg_kernel_timing else None
                    _lib.call("S2.6.3", "gf_scene_pcg_sync", ctypes.byref(s), nact, prec, tol,
                              pressure_iters, ce, sign, ctypes.byref(it),
                              ctypes.byref(res), ctypes.byref(nrep), kms, stream)
                    if kms is not None:
                        self.last_pcg_kernel_ms = tuple(kms)
                    self._pressure_stats = (it.value, res.value, nrep.value)
                if capture or sc.n_active_host > 0:
                    _lib.call("S2.6.3", "gf_scene_pcg_flags", ctypes.byref(s), tol,
                              _lib.ptr(self._pcg_done.t), _lib.ptr(self._pcg_tol_sq.t), stream)
                else:   # nothing to solve: the reference's early return
                    self._pcg_done.t.zero_()
                    self._pcg_tol_sq.t.zero_()
        with prof.section("g2p_advect"):
            if pre is None:
                _lib.call("S2.8", "gf_scene_g2p", ctypes.byref(s), _lib.ptr(sc.pc), mode,
                          self.gravity, dt, self.rho, flip_blend, stream)
            else:
                _lib.call("S2.8", "gf_scene_g2p_pre", ctypes.byref(s), _lib.ptr(sc.pc),
                          *(_lib.ptr(t) for t in pre), mode, dt, self.rho, flip_blend, stream)
        prof.fused("grad_subtract_bc")
        # the colour section opens every step, as the reference's does
        # (solver3d.py:3656-3658); the transfer runs when the attribute exists
        with prof.section("color"):
            if n > 0 and self._attr["color"] is not None:
                self._attr_transfer(s, self._attr["color"], 3, _lib.GF_ATTR_COLOR, stream)
        if n > 0 and self._attr["temperature"] is not None:
            with prof.section("temperature"):
                self._attr_transfer(s, self._attr["temperature"], 1, _lib.GF_ATTR_SCALAR, stream)
        if hook is not None:
            hook("done")

    @block("F3.4", "CFL-adaptive substepping")
    def step_cfl(self, target_dt: float, pressure_iters: int = 80, cfl: float = 0.5,
                 max_substeps: int = 16, pressure_solver: str = "jacobi") -> int:
        """Substep count from the device |v|max (solver3d.py:4077-4121)."""
        n = cfl_substep_count_gpu(self, self.dx, target_dt, cfl=cfl, max_substeps=max_substeps)
        sub_dt = target_dt / n
        for _ in range(n):
            self.step(sub_dt, pressure_iters=pressure_iters, pressure_solver=pressure_solver)
        return n

    # -------------------------------------------------------------- queries
    @property
    def n_active_tiles(self) -> int:
        self._stats()
        return self._sc.n_active_host

    def fluid_cell_count(self) -> int:
        return int((self._f["marker"] == 1).sum().item())


@block("S2.10.GPU", "Device |v|max reduction for the CFL substep count")
def cfl_substep_count_gpu(solver_or_vel, dx: float, target_dt: float, cfl: float = 0.5,
                          max_substeps: int = 32, scratch=None) -> int:
    """ceil(target_dt / (cfl*dx/|v|max)) with |v|max reduced on the device and
    one float read back (solver3d.py:2255-2280)."""
    if isinstance(solver_or_vel, FlipSolver3D):
        vel = solver_or_vel._vel
    else:
        vel = solver_or_vel.t if isinstance(solver_or_vel, DeviceArray) else solver_or_vel
    if vel is None or vel.shape[0] == 0:
        return 1
    out = scratch if scratch is not None else torch.zeros(1, dtype=torch.float32, device=vel.device)
    out.zero_()
    _lib.call("S2.10.GPU", "gf_k3_vmax_sq_reduce", _lib.ptr(vel.contiguous()), int(vel.shape[0]),
              _lib.ptr(out), _lib.stream_handle(vel.device))
    vmax = math.sqrt(float(out.item()))
    if vmax < 1e-6:
        return 1
    n = max(1, int(np.ceil(target_dt / (cfl * dx / vmax))))
    return min(n, max_substeps)
