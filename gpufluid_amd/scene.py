"""Device workspace of one scene and its gf_scene_t descriptor.

All buffers are torch tensors allocated once per grid (and grown for the
particle count); the HIP library never allocates.  Sizes come from
gf_scene_sizes(): dense cell arrays (C), per-tile tables (ntiles) and the
compacted pressure vectors (ntiles*512 entries, only the first
n_active*512 are touched each substep).  At 256^3 the whole workspace,
FP64 PCG vectors included, is about 1.5 GB plus ~80 B per particle slot.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

SCENE_PAR = _lib.SCENE_PAR


class SceneBuffers:
    def __init__(self, nx: int, ny: int, nz: int, dx: float, dom, device: torch.device):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx = float(dx)
        self.dom = tuple(float(d) for d in dom)
        self.device = device
        sz = _lib.sizes(nx, ny, nz)
        self.cells, self.ntiles, self.compact = sz.cells, sz.ntiles, sz.compact
        i32 = dict(dtype=torch.int32, device=device)
        f32 = dict(dtype=torch.float32, device=device)
        f64 = dict(dtype=torch.float64, device=device)
        self.counts = torch.zeros(sz.cells, **i32)        # zero between substeps
        self.cell_start = torch.zeros(sz.cells + 1, **i32)
        self.scan_part = torch.zeros(sz.scan_parts, **i32)
        self.tile_occ = torch.zeros(sz.ntiles, **i32)
        self.tile_map = torch.full((sz.ntiles,), -1, **i32)
        self.active = torch.zeros(sz.ntiles, **i32)
        self.n_active = torch.zeros(4, **i32)
        self.nbr = torch.full((sz.ntiles * 6,), -1, **i32)
        self.flags = torch.zeros(sz.compact, dtype=torch.uint8, device=device)
        self.b = torch.zeros(sz.compact, **f32)
        self.x32 = torch.zeros(sz.compact, **f32)
        self.r32 = torch.zeros(sz.compact, **f32)
        self.p32 = torch.zeros(sz.compact, **f32)
        self.p32b = torch.zeros(sz.compact, **f32)
        self.cg32 = torch.zeros(3 * sz.compact, **f32)
        self.ap32 = torch.zeros(sz.compact, **f32)
        self.pc = torch.zeros(sz.compact, **f32)
        self.ptmp = torch.zeros(sz.compact, **f32)
        self.x64 = torch.zeros(sz.compact, **f64)
        self.r64 = self.p64 = self.p64b = self.ap64 = self.cg64 = None  # FP64 PCG, on first use
        self.partials = torch.zeros(sz.partials, **f64)
        self.scalars = torch.zeros(sz.scalars, **f64)
        self.pcg_part = torch.zeros(sz.pcg_part, **i32)
        self.pcg_part_base = sz.pcg_part
        # packed scenes (gpufluid_amd.batch): per-tile scene ids and scene tables
        self.n_scenes = 0
        self.tile_scene = self.tile_order = self.scene_par = self.scene_slot = None
        self.scene_pstart = None
        self.cap = 0
        self.bufs = [None, None]     # ping-pong particle storage, cap rows each
        self.cur = 0
        self.keys = self.tmp = None
        self.n_active_host = 0

    def ensure_fp64(self):
        if self.r64 is None:
            kw = dict(dtype=torch.float64, device=self.device)
            self.r64 = torch.zeros(self.compact, **kw)
            self.p64 = torch.zeros(self.compact, **kw)
            self.p64b = torch.zeros(self.compact, **kw)
            self.cg64 = torch.zeros(3 * self.compact, **kw)
            self.ap64 = torch.zeros(self.compact, **kw)

    def _alloc(self, cap, affine):
        f32 = dict(dtype=torch.float32, device=self.device)
        return {"pos": torch.zeros((cap, 3), **f32), "vel": torch.zeros((cap, 3), **f32),
                "pid": torch.zeros(cap, dtype=torch.int32, device=self.device),
                "affine": torch.zeros((cap, 9), **f32) if affine else None}

    def ensure_particles(self, n: int, affine: bool):
        """Grow both particle buffers to hold ``n`` (copying the live ``n``
        rows of the current one) and add affine storage when needed."""
        if n > self.cap or self.bufs[self.cur] is None:
            cap = max(n, self.cap, 1)
            old = self.bufs[self.cur]
            new = [self._alloc(cap, affine), self._alloc(cap, affine)]
            if old is not None:
                m = min(self.cap, n)
                for k in ("pos", "vel", "pid", "affine"):
                    if old[k] is not None and new[0][k] is not None:
                        new[0][k][:m].copy_(old[k][:m])
            self.bufs, self.cur, self.cap = new, 0, cap
            i32 = dict(dtype=torch.int32, device=self.device)
            self.keys = torch.zeros(cap, **i32)
            self.tmp = torch.zeros(cap, **i32)
        if affine:
            for b in self.bufs:
                if b is not None and b["affine"] is None:
                    b["affine"] = torch.zeros((self.cap, 9), dtype=torch.float32,
                                              device=self.device)

    def live(self, n: int) -> dict:
        """Views of the current particle state (n rows)."""
        b = self.bufs[self.cur]
        if b is None:
            return {"pos": None, "vel": None, "pid": None, "affine": None}
        return {k: (None if t is None else t[:n]) for k, t in b.items()}

    def swap(self):
        self.cur = 1 - self.cur

    def reset_counts(self):
        """Restore the zero-between-calls invariant of ``counts`` after a failure."""
        self.counts.zero_()

    def descriptor(self, fields: dict, n: int, affine: bool) -> _lib.GfScene:
        s = _lib.GfScene()
        s.nx, s.ny, s.nz = self.nx, self.ny, self.nz
        s.tnx, s.tny, s.tnz = (self.nx + 7) // 8, (self.ny + 7) // 8, (self.nz + 7) // 8
        s.dx = self.dx
        s.dom[0], s.dom[1], s.dom[2] = self.dom
        s.n = n
        s.cap = self.cap
        P = _lib.ptr
        cur = self.bufs[self.cur] or {}
        alt = self.bufs[1 - self.cur] or {}
        s.pos, s.vel, s.pid = P(cur.get("pos")), P(cur.get("vel")), P(cur.get("pid"))
        s.pos_b, s.vel_b, s.pid_b = P(alt.get("pos")), P(alt.get("vel")), P(alt.get("pid"))
        if affine:
            s.affine_c, s.affine_c_b = P(cur.get("affine")), P(alt.get("affine"))
        s.keys, s.tmp, s.counts, s.cell_start = P(self.keys), P(self.tmp), P(self.counts), P(self.cell_start)
        s.scan_part = P(self.scan_part)
        s.tile_occ, s.tile_map, s.active = P(self.tile_occ), P(self.tile_map), P(self.active)
        s.n_active, s.nbr = P(self.n_active), P(self.nbr)
        s.marker = P(fields["marker"])
        s.u, s.v, s.w = P(fields["u"]), P(fields["v"]), P(fields["w"])
        s.us, s.vs, s.ws = P(fields["us"]), P(fields["vs"]), P(fields["ws"])
        s.solid_u, s.solid_v, s.solid_w = (P(fields.get("solid_u")), P(fields.get("solid_v")),
                                           P(fields.get("solid_w")))
        s.flags, s.b = P(self.flags), P(self.b)
        s.x32, s.r32, s.p32, s.ap32 = P(self.x32), P(self.r32), P(self.p32), P(self.ap32)
        s.x64, s.r64, s.p64, s.ap64 = P(self.x64), P(self.r64), P(self.p64), P(self.ap64)
        s.pc, s.ptmp = P(self.pc), P(self.ptmp)
        s.partials, s.scalars = P(self.partials), P(self.scalars)
        s.p32b, s.p64b = P(self.p32b), P(self.p64b)
        s.cg32, s.cg64 = P(self.cg32), P(self.cg64)
        s.pcg_part = P(self.pcg_part)
        if self.n_scenes > 1:
            s.n_scenes = self.n_scenes
            s.tile_scene, s.tile_order = P(self.tile_scene), P(self.tile_order)
            s.scene_par, s.scene_slot = P(self.scene_par), P(self.scene_slot)
            s.scene_pstart = P(self.scene_pstart)
        return s

    def set_scenes(self, tile_scene, tile_order, scene_par, scene_pstart):
        """Describe packed scenes: per-tile scene id, scene-major tile order,
        the per-scene table (SCENE_PAR floats per scene: g*dt, clamp lo
        x,y,z, clamp hi x,y,z, dx, cell offset x,y,z, cells per axis) and the
        first particle id of each scene (n_scenes + 1 entries)."""
        dev = self.device
        n = int(scene_par.shape[0])
        self.n_scenes = n
        self.tile_scene = torch.as_tensor(tile_scene, dtype=torch.int32, device=dev).contiguous()
        self.tile_order = torch.as_tensor(tile_order, dtype=torch.int32, device=dev).contiguous()
        self.scene_par = torch.as_tensor(scene_par, dtype=torch.float32,
                                         device=dev).reshape(n, SCENE_PAR).contiguous()
        self.scene_pstart = torch.as_tensor(scene_pstart, dtype=torch.int32,
                                            device=dev).reshape(n + 1).contiguous()
        self.scene_slot = torch.zeros(n + 1, dtype=torch.int32, device=dev)
        self.pcg_part = torch.zeros(self.pcg_part_base + n, dtype=torch.int32, device=dev)
        self.scalars = torch.zeros(48 * n + 64, dtype=torch.float64, device=dev)


def byref(s):
    return ctypes.byref(s)
