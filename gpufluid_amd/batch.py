"""Many independent FLIP scenes of uneven grid sizes in one set of launches.

The scenes are packed into one grid: scene s with n_s cells per axis and cell
size dx_s sits at a tile-aligned cell offset o_s, inside its own one-cell
solid shell (the reference walls, solver3d.py:2475-2482), so no stencil,
scatter or gather ever couples two scenes.  Particles keep their scene's world
units (its particle ids form one contiguous range), and the kernels take each
scene's dx, offset, size, g*dt and clamp box from a per-scene table (through
the tile's scene id, or the particle's id range when binning): a particle's
cell, its P2G weights, its G2P samples and its clamp are computed exactly as
for the scene run alone, and only integer indices are shifted by o_s.

Every stage is then one launch for all scenes: the sort, P2G and G2P kernels
cover all tiles, and the PCG runs segmented — compact tiles are numbered
scene-major, each scene gets its own workgroups, dot products, alpha/beta and
stop flag — so each scene is solved exactly as it would be alone (same
tolerance on its own ||b||), with finished scenes passing through.  A substep
pair can be captured in a hipGraph (the particle buffers ping-pong, so two
substeps bring the pointers back) with a predicted PCG budget and a synchronous
continuation for scenes that need more iterations.  No reference counterpart:
the reference runs one scene per solver.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .blocks import block

TILE = 8
BIN_EDGE = 128


def _ceil8(n: int) -> int:
    return (n + TILE - 1) // TILE * TILE


def _decode(entries: torch.Tensor, n: int):
    """(done, iters) of n 16-double solver log entries (ints from byte 72:
    iters, done, ...)."""
    raw = entries.detach().cpu().numpy().view(np.int32).reshape(n, 32)
    return raw[:, 19] != 0, raw[:, 18].astype(np.int64)


@dataclass
class Packing:
    offsets: np.ndarray   # (S,3) cell offset of each scene
    edges: np.ndarray     # (S,) cells per axis of each scene
    shape: tuple          # packed grid (nx, ny, nz)


@block("X2.1", "Pack cubic scenes into one grid at tile-aligned offsets")
def pack_scenes(edges) -> Packing:
    """Deterministic packing: scenes of one slot edge e (edge rounded up to 8)
    fill bins of k^3 slots, k = max(1, 128 // e); bins of one edge form a
    near-cubic block; blocks are placed one after another along x.  Every
    offset is a multiple of 8, and coordinates stay below a few hundred cells
    per axis for typical batches, which keeps FP32 positions close to the
    per-scene precision."""
    edges = np.asarray([int(e) for e in edges], dtype=np.int64)
    if edges.size == 0 or (edges < 3).any():
        raise ValueError("every scene needs at least 3 cells per axis")
    slots = np.array([_ceil8(e) for e in edges])
    offsets = np.zeros((len(edges), 3), dtype=np.int64)
    bins = []  # (bin edge, [(scene, local offset)])
    for e in sorted(set(slots.tolist()), reverse=True):
        members = [i for i in range(len(edges)) if slots[i] == e]
        k = max(1, BIN_EDGE // e)
        per = k ** 3
        for b0 in range(0, len(members), per):
            chunk = members[b0:b0 + per]
            local = []
            for q, sc in enumerate(chunk):
                a, r = divmod(q, k * k)
                bb, c = divmod(r, k)
                local.append((sc, (a * e, bb * e, c * e)))
            bins.append((k * e if len(chunk) > 1 else e, local))
    x0, ny, nz = 0, 0, 0
    for be in sorted(set(b[0] for b in bins), reverse=True):
        group = [b for b in bins if b[0] == be]
        m = max(1, math.ceil(round(len(group) ** (1.0 / 3.0), 9)))
        while m * m * max(1, math.ceil(len(group) / (m * m))) < len(group):
            m += 1
        depth = math.ceil(len(group) / (m * m))
        for q, (_, local) in enumerate(group):
            a, r = divmod(q, m * m)
            bb, c = divmod(r, m)
            base = np.array([x0 + a * be, bb * be, c * be])
            for sc, off in local:
                offsets[sc] = base + np.array(off)
        x0 += depth * be
        ny = max(ny, m * be)
        nz = max(nz, m * be)
    return Packing(offsets, edges, (int(x0), int(ny), int(nz)))


class BatchedScenes:
    """FLIP substeps of many independent cubic scenes (unit boxes of n^3 cells,
    walls on all faces, as ``FlipSolver3D``) in one packed grid.

    ``scenes``: sequence of (n, pos (m,3), vel (m,3)) in world units.
    """

    def __init__(self, scenes, gravity: float = -9.81, flip_blend: float = 0.95,
                 rho: float = 1.0, dt: float = 1.0 / 240.0, device=None,
                 pcg_precision: str = "mixed", pcg_check_every: int = 10):
        from .solver3d import FlipSolver3D
        self.edges = [int(sc[0]) for sc in scenes]
        self.pack = pack_scenes(self.edges)
        nx, ny, nz = self.pack.shape
        self.dt = float(dt)
        self.gravity = float(gravity)
        self.s = FlipSolver3D(nx, ny, nz, dx=1.0, gravity=gravity, flip_blend=flip_blend, rho=rho,
                              device=device, pcg_precision=pcg_precision,
                              pcg_check_every=pcg_check_every)
        S = len(self.edges)
        marker = np.full((nx, ny, nz), 2, dtype=np.int32)
        tnx, tny, tnz = nx // TILE, ny // TILE, nz // TILE
        tile_scene = np.full((tnx, tny, tnz), -1, dtype=np.int32)
        par = np.zeros((S, _lib.SCENE_PAR), dtype=np.float32)
        self.dx = np.array([1.0 / n for n in self.edges])
        f32 = np.float32
        for i, (n, o) in enumerate(zip(self.edges, self.pack.offsets)):
            marker[o[0] + 1:o[0] + n - 1, o[1] + 1:o[1] + n - 1, o[2] + 1:o[2] + n - 1] = 0
            t0, t1 = o // TILE, (o + _ceil8(n)) // TILE
            tile_scene[t0[0]:t1[0], t0[1]:t1[1], t0[2]:t1[2]] = i
            # the FP32 values FlipSolver3D(n, n, n, dx=1/n) hands its kernels:
            # g*dt, the clamp box [1.001 dx, n dx - 1.001 dx], dx
            dx32 = f32(1.0 / n)
            eps = f32(dx32 * f32(1.001))
            dom = f32(n * (1.0 / n))
            par[i, 0] = f32(f32(gravity) * f32(dt))
            par[i, 1:4] = eps
            par[i, 4:7] = f32(dom - eps)
            par[i, 7] = dx32
            par[i, 8:11] = o
            par[i, 11] = n
        flat = tile_scene.reshape(-1)
        order = np.concatenate([np.nonzero(flat == i)[0] for i in range(S)]
                               + [np.nonzero(flat < 0)[0]]).astype(np.int32)
        s = self.s
        s._marker_host = marker
        s._wall_marker = marker.copy()
        s._f["marker"].copy_(torch.as_tensor(marker, device=s.device))
        counts = [len(sc[1]) for sc in scenes]
        self.start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        s._sc.set_scenes(flat, order, par, self.start)
        self.scene_par = par
        # particles stay in their scene's world units (scene i: ids start[i]..)
        pos = np.concatenate([np.asarray(sc[1], dtype=np.float32).reshape(-1, 3) for sc in scenes])
        vel = np.concatenate([np.asarray(sc[2], dtype=np.float32).reshape(-1, 3) for sc in scenes])
        s.set_particles(pos, vel)
        self.n_particles = s.n_particles
        self.last_iters = None
        self.last_residual = None
        self._graph = None
        self._graph_budget = 0
        self._budget_floor = 0      # chunks; raised when a replay needed a larger budget

    @property
    def n_scenes(self) -> int:
        return len(self.edges)

    # ------------------------------------------------------------- substeps
    @block("X2.2", "One substep of every packed scene (sort, P2G, segmented PCG, G2P)")
    def step(self, pressure_iters: int = 500, pressure_tol: float = 1e-8):
        s = self.s
        s._step_impl(self.dt, int(pressure_iters), "pcg", float(pressure_tol))
        self._read_stats()

    def _read_stats(self):
        S = self.n_scenes
        it = (ctypes.c_int * S)()
        res = (ctypes.c_double * S)()
        d = self.s._descriptor()
        _lib.call("X2.2", "gf_scene_pcg_stats", ctypes.byref(d), it, res, None, self.s._stream())
        self.last_iters = np.array(it[:], dtype=np.int64)
        self.last_residual = np.array(res[:])

    def _enqueue_substep(self, pressure_iters, pressure_tol, budget):
        """Launches of one substep with no host synchronisation."""
        s = self.s
        sc = s._sc
        st = s._stream()
        d = s._descriptor()
        _lib.call("X2.3", "gf_scene_sort_sync", ctypes.byref(d), 0, None, st)
        sc.swap()
        d = s._descriptor()
        if s.p2g_staged:
            stage, has = s._p2g_stage()
            _lib.call("X2.3", "gf_scene_p2g_staged", ctypes.byref(d), 0, s.gravity, self.dt,
                      s.rho, _lib.ptr(stage), _lib.ptr(has), st)
        else:
            _lib.call("X2.3", "gf_scene_p2g", ctypes.byref(d), 0, s.gravity, self.dt, s.rho, st)
        prec = _lib.GF_PCG_MIXED if s.pcg_precision == "mixed" else _lib.GF_PCG_FP64
        _lib.call("X2.3", "gf_scene_pcg_enqueue", ctypes.byref(d), 0, prec, pressure_tol,
                  pressure_iters, s.pcg_check_every, -1, 0, budget, st)
        _lib.call("X2.3", "gf_scene_g2p", ctypes.byref(d), _lib.ptr(sc.pc), 0, s.gravity, self.dt,
                  s.rho, s.flip_blend, st)
        return d

    @block("X2.3", "Two substeps of every packed scene replayed from a captured hipGraph")
    def step2_graph(self, pressure_iters: int = 500, pressure_tol: float = 1e-8,
                    budget_iters: int | None = None):
        """Two substeps from one hipGraph replay.  The PCG of each substep gets
        ``budget_iters`` iterations inside the graph (default: the cap).  If a
        scene has not stopped within the budget in either substep, the pair is
        rerun from a copy of the particle state with the budget raised by half (the
        graph is recaptured), so results never depend on the budget.  A budget
        that had to grow stays as the floor for later calls, and a pair that
        comes within 10% of its budget raises that floor by a quarter, so a
        solve whose iteration count drifts upward is recaptured ahead of time
        instead of rerun."""
        s = self.s
        ce = s.pcg_check_every
        cap_chunks = math.ceil(pressure_iters / ce)
        budget = max(1, min(math.ceil((budget_iters or pressure_iters) / ce), cap_chunks))
        budget = max(budget, min(self._budget_floor, cap_chunks))
        s._cache.clear()
        grew = False
        while True:
            if self._graph is None or self._graph_budget != budget:
                self._capture(pressure_iters, pressure_tol, budget)
            snap = {k: v.clone() for k, v in s._sc.live(s.n_particles).items() if v is not None}
            self._graph.replay()
            torch.cuda.current_stream(s.device).synchronize()
            done_a, it_a = _decode(self._res_a, self.n_scenes)
            done_b, it_b = _decode(self._res_b(), self.n_scenes)
            self.last_iters_pair = (it_a, it_b)
            if (done_a.all() and done_b.all()) or budget >= cap_chunks:
                self._read_stats()
                s._cache.clear()
                if grew:
                    self._budget_floor = max(self._budget_floor, budget)
                # iteration counts drift as the scenes evolve: a pair that came
                # within 10% of the budget raises the floor ahead of time (a
                # recapture costs far less than rerunning a pair)
                need = int(max(it_a.max(), it_b.max()))
                if 10 * need > 9 * budget * ce:
                    self._budget_floor = max(self._budget_floor,
                                             min(cap_chunks, math.ceil(1.25 * need / ce)))
                return
            grew = True
            live = s._sc.live(s.n_particles)
            for k, v in snap.items():
                live[k].copy_(v)
            budget = min(cap_chunks, budget + (budget + 1) // 2)

    def _res_b(self):
        S = self.n_scenes
        return self.s._sc.scalars[32 * S:48 * S]

    def _capture(self, pressure_iters, pressure_tol, budget):
        s = self.s
        S = self.n_scenes
        self._res_a = torch.zeros(16 * S, dtype=torch.float64, device=s.device)
        cur = s._sc.cur
        g = torch.cuda.CUDAGraph()
        stream = torch.cuda.Stream(device=s.device)
        stream.wait_stream(torch.cuda.current_stream(s.device))
        with torch.cuda.stream(stream):
            with torch.cuda.graph(g, stream=stream):
                self._enqueue_substep(pressure_iters, pressure_tol, budget)
                self._res_a.copy_(self._res_b())
                self._enqueue_substep(pressure_iters, pressure_tol, budget)
        torch.cuda.current_stream(s.device).wait_stream(stream)
        if s._sc.cur != cur:
            raise RuntimeError("captured substep pair must leave the particle buffers in place")
        self._graph = g
        self._graph_budget = budget

    # -------------------------------------------------------------- results
    def scene_particles(self, i: int):
        """(pos, vel) of scene i in its world units as FP64 numpy arrays."""
        s = self.s
        s._cache.clear()
        pos = s.pos.t[self.start[i]:self.start[i + 1]].double().cpu().numpy()
        vel = s.vel.t[self.start[i]:self.start[i + 1]].double().cpu().numpy()
        return pos, vel

    def scene_marker(self, i: int):
        """Scene i's block of the packed marker (its n^3 cells), numpy int32."""
        n, o = self.edges[i], self.pack.offsets[i]
        m = self.s._f["marker"][o[0]:o[0] + n, o[1]:o[1] + n, o[2]:o[2] + n]
        return m.cpu().numpy()
