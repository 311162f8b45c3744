"""S2 block interface: the reference's kernels as Python callables.

Each function replaces the Warp kernel of the same name
(src/gpufluid/solvers/solver3d.py) and is registered under the same block id.
Arguments keep the reference order and meaning with ``wp.launch``'s ``dim``
and ``device`` dropped: arrays are DeviceArray or torch tensors on the HIP
device, scalars are Python numbers (float -> FP32, int -> int32), and the
trailing sub-dense offsets ``(off_x, off_y, off_z)`` must be 0.  Work is
enqueued on torch's current stream; nothing is returned (outputs are written
in place), as with ``wp.launch``.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .blocks import BlockError, block
from .device import DeviceArray
from .scene import SceneBuffers


def _t(a):
    if a is None:
        return None
    t = a.t if isinstance(a, DeviceArray) else a
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError("block kernels take HIP device arrays (DeviceArray or torch tensor)")
    if not t.is_contiguous():
        raise ValueError("block kernels need C-contiguous arrays")
    return t


def _p(a):
    return _lib.ptr(_t(a))


def _offsets(block_id, *offs):
    if any(int(o) != 0 for o in offs):
        raise BlockError(block_id, "sub-dense offsets are not supported (dense storage only)")


def _st(a):
    return _lib.stream_handle(_t(a).device)


def _cells(marker):
    m = _t(marker)
    if m.dim() != 3 or m.dtype != torch.int32:
        raise ValueError("marker must be an int32 (nx, ny, nz) array")
    return tuple(int(x) for x in m.shape)


def k3_clear_grid(u, v, w, uw, vw, ww, marker, off_x=0, off_y=0, off_z=0):
    """k3_clear_grid (solver3d.py:41-65; not registered as a block there):
    zero the face and weight arrays and reset non-solid marker cells to air."""
    _offsets("S2.4", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.4", "gf_k3_clear_grid", _p(u), _p(v), _p(w), _p(uw), _p(vw), _p(ww),
              _p(marker), nx, ny, nz, _st(marker))


_P2G_SCENES: dict = {}


def _p2g_block(block_id, pos, vel, affine, u, v, w, uw, vw, ww, marker, dx, nx, ny, nz):
    m = _t(marker)
    shape = _cells(marker)
    if (nx, ny, nz) != shape:
        raise BlockError(block_id, f"grid ({nx},{ny},{nz}) does not match marker {shape}")
    pos_t = _t(pos).reshape(-1, 3)
    n = pos_t.shape[0]
    dev = pos_t.device
    stream = _st(pos)
    # Stable counting sort into a scratch scene, raw LDS accumulation, then the
    # reference's marker side effect on the caller's marker.
    key = (shape, float(dx), dev)
    sc = _P2G_SCENES.get(key)
    if sc is None:
        import numpy as np
        dom = tuple(float(np.float32(k * dx)) for k in shape)
        sc = SceneBuffers(nx, ny, nz, dx, dom, dev)
        sc.scratch_marker = torch.zeros(shape, dtype=torch.int32, device=dev)
        _P2G_SCENES.clear()
        _P2G_SCENES[key] = sc
    apic = affine is not None
    sc.ensure_particles(n, apic)
    live = sc.live(n)
    if n:
        live["pos"].copy_(pos_t)
        live["vel"].copy_(_t(vel).reshape(n, 3))
        live["pid"].copy_(torch.arange(n, dtype=torch.int32, device=dev))
        if apic:
            live["affine"].copy_(_t(affine).reshape(n, 9))
    sc.scratch_marker.copy_(torch.where(m == 2, m, torch.zeros_like(m)))
    dummy = torch.empty(0, device=dev)
    fields = {"marker": sc.scratch_marker, "u": _t(u), "v": _t(v), "w": _t(w),
              "us": dummy, "vs": dummy, "ws": dummy}
    s = sc.descriptor(fields, n, apic)
    _lib.call(block_id, "gf_scene_sort_sync", ctypes.byref(s), int(apic), None, stream)
    if n:
        sc.swap()
    s = sc.descriptor(fields, n, apic)
    mode = _lib.GF_TRANSFER_APIC if apic else _lib.GF_TRANSFER_FLIP
    _lib.call(block_id, "gf_scene_p2g_raw", ctypes.byref(s), mode, _p(uw), _p(vw), _p(ww), stream)
    _lib.call(block_id, "gf_k3_mark_fluid", _lib.ptr(pos_t.contiguous()), n, _p(marker),
              float(dx), nx, ny, nz, stream)


@block("S2.1", "P2G: scatter particle velocities -> MAC faces (trilinear)")
def k3_p2g(pos, vel, u, v, w, uw, vw, ww, marker, dx, nx, ny, nz, off_x=0, off_y=0, off_z=0):
    """Accumulates into u,v,w and uw,vw,ww (+=) and marks particle cells
    fluid (solver3d.py:72-114)."""
    _offsets("S2.1", off_x, off_y, off_z)
    _p2g_block("S2.1", pos, vel, None, u, v, w, uw, vw, ww, marker, dx, nx, ny, nz)


@block("S2.12", "APIC P2G: scatter (vel + affine*offset) per face")
def k3_p2g_apic(pos, vel, affine_C, u, v, w, uw, vw, ww, marker, dx, nx, ny, nz,
                off_x=0, off_y=0, off_z=0):
    """solver3d.py:1475-1588"""
    _offsets("S2.12", off_x, off_y, off_z)
    _p2g_block("S2.12", pos, vel, affine_C, u, v, w, uw, vw, ww, marker, dx, nx, ny, nz)


@block("S2.2", "Normalize MAC faces and save pre-pressure copy")
def k3_normalize(u, v, w, uw, vw, ww, us, vs, ws):
    nx, ny, nz = int(_t(v).shape[0]), int(_t(u).shape[1]), int(_t(u).shape[2])
    _lib.call("S2.2", "gf_k3_normalize", _p(u), _p(v), _p(w), _p(uw), _p(vw), _p(ww), _p(us),
              _p(vs), _p(ws), nx, ny, nz, _st(u))


@block("S2.3", "Add gravity g*dt to v faces")
def k3_add_gravity(v, g, dt):
    t = _t(v)
    nx, ny1, nz = (int(x) for x in t.shape)
    _lib.call("S2.3", "gf_k3_add_gravity", _p(v), nx, ny1 - 1, nz, float(g), float(dt), _st(v))


@block("S2.4", "Solid boundary condition: faces = obstacle velocity")
def k3_enforce_solid_bc(u, v, w, marker, solid_u, solid_v, solid_w, nx, ny, nz,
                        off_x=0, off_y=0, off_z=0):
    _offsets("S2.4", off_x, off_y, off_z)
    _lib.call("S2.4", "gf_k3_enforce_solid_bc", _p(u), _p(v), _p(w), _p(marker), _p(solid_u),
              _p(solid_v), _p(solid_w), int(nx), int(ny), int(nz), _st(u))


@block("S2.5", "Cell-wise divergence x rho*dx/dt (Poisson RHS)")
def k3_compute_divergence(u, v, w, div, marker, dx, dt, rho, off_x=0, off_y=0, off_z=0):
    _offsets("S2.5", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.5", "gf_k3_compute_divergence", _p(u), _p(v), _p(w), _p(div), _p(marker), nx,
              ny, nz, float(dx), float(dt), float(rho), _st(u))


@block("S2.6.1", "Jacobi pressure iteration (one sweep)")
def k3_jacobi_pressure(p_in, p_out, div, marker, off_x=0, off_y=0, off_z=0):
    _offsets("S2.6.1", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.6.1", "gf_k3_jacobi_pressure", _p(p_in), _p(p_out), _p(div), _p(marker), nx,
              ny, nz, _st(p_in))


@block("S2.13", "Implicit viscosity Jacobi sweep on one MAC component")
def k3_jacobi_visc(u_old, u_in, u_out, r):
    """k3_jacobi_visc (solver3d.py:178-197): u_out = (u_old + r*sum_nb(u_in)) /
    (1 + r*count_nb) over the in-range neighbours of each entry of one
    component array."""
    t = _t(u_in)
    if t.dim() != 3 or _t(u_old).shape != t.shape or _t(u_out).shape != t.shape:
        raise ValueError("k3_jacobi_visc: u_old, u_in, u_out must share one 3-D shape")
    sx, sy, sz = (int(x) for x in t.shape)
    _lib.call("S2.13", "gf_k3_jacobi_visc", _p(u_old), _p(u_in), _p(u_out), sx, sy, sz,
              float(r), _st(u_in))


@block("S2.6.2", "Gauss-Seidel red-black sweep (in-place)")
def k3_gauss_seidel_rb(p, div, marker, color, off_x=0, off_y=0, off_z=0):
    _offsets("S2.6.2", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.6.2", "gf_k3_gauss_seidel_rb", _p(p), _p(div), _p(marker), nx, ny, nz,
              int(color), _st(p))


@block("S2.6.3", "PCG: apply Laplacian operator A to x")
def k3_apply_A(x, y, marker, done, off_x=0, off_y=0, off_z=0):
    _offsets("S2.6.3", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.6.3", "gf_k3_apply_A", _p(x), _p(y), _p(marker), _p(done), nx, ny, nz, _st(x))


@block("S2.6.3", "PCG: Jacobi preconditioner diagonal")
def k3_compute_diag(marker, diag, off_x=0, off_y=0, off_z=0):
    _offsets("S2.6.3", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.6.3", "gf_k3_compute_diag", _p(marker), _p(diag), nx, ny, nz, _st(marker))


@block("S2.6.3", "PCG: z = r / diag on fluid cells")
def k3_apply_invM(r, diag, z, marker, done, off_x=0, off_y=0, off_z=0):
    _offsets("S2.6.3", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.6.3", "gf_k3_apply_invM", _p(r), _p(diag), _p(z), _p(marker), _p(done), nx,
              ny, nz, _st(r))


@block("S2.6.3", "PCG: fluid-cell dot product accumulated into out[0]")
def k3_dot_fluid(a, b, out, marker, done, off_x=0, off_y=0, off_z=0):
    _offsets("S2.6.3", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.6.3", "gf_k3_dot_fluid", _p(a), _p(b), _p(out), _p(marker), _p(done), nx, ny,
              nz, _st(a))


@block("S2.6.3", "GPU-resident PCG axpy with a device scalar")
def k3_axpy_devscalar(a_dev, sign, x, y_in, y_out, marker, done, off_x=0, off_y=0, off_z=0):
    _offsets("S2.6.3", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.6.3", "gf_k3_axpy_devscalar", _p(a_dev), float(sign), _p(x), _p(y_in),
              _p(y_out), _p(marker), _p(done), nx, ny, nz, _st(x))


def _scalar(op, a=None, b=None, c=None, tol=0.0, done=None):
    ref = next(x for x in (a, b, c, done) if x is not None)
    _lib.call("S2.6.3", "gf_k3_scalar_op", op, _p(a), _p(b), _p(c), float(tol), _p(done), _st(ref))


def k3_zero_scalar(s, done):
    _scalar(0, a=s, done=done)


def k3_div_scalar(numer, denom, out, done):
    _scalar(1, a=numer, b=denom, c=out, done=done)


def k3_copy_scalar(src, dst, done):
    _scalar(2, a=src, c=dst, done=done)


def k3_set_tol_sq(r0_norm2, tol, tol_sq):
    _scalar(3, a=r0_norm2, c=tol_sq, tol=tol)


@block("S2.6.3", "PCG on-device convergence check (stop flag)")
def k3_check_converged(r_norm2, tol_sq, done):
    _scalar(4, a=r_norm2, b=tol_sq, done=done)


def k3_zero_int_scalar(s):
    _scalar(5, done=s)


@block("S2.7", "Subtract pressure gradient from MAC faces")
def k3_subtract_pressure_grad(u, v, w, p, marker, dx, dt, rho, off_x=0, off_y=0, off_z=0):
    _offsets("S2.7", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.7", "gf_k3_subtract_pressure_grad", _p(u), _p(v), _p(w), _p(p), _p(marker),
              nx, ny, nz, float(dx), float(dt), float(rho), _st(u))


def _dom(dom):
    vals = [float(x) for x in (dom if not hasattr(dom, "numpy") else dom.numpy())]
    return (ctypes.c_float * 3)(*vals)


@block("S2.8", "G2P sample + FLIP/PIC blend (fused with S2.9)")
def k3_g2p_and_advect(pos, vel, u, v, w, us, vs, ws, dx, dt, nx, ny, nz, flip_blend, dom,
                      off_x=0, off_y=0, off_z=0):
    _offsets("S2.8", off_x, off_y, off_z)
    n = int(_t(pos).reshape(-1, 3).shape[0])
    _lib.call("S2.8", "gf_k3_g2p_and_advect", _p(pos), _p(vel), _p(u), _p(v), _p(w), _p(us),
              _p(vs), _p(ws), n, float(dx), float(dt), int(nx), int(ny), int(nz),
              float(flip_blend), _dom(dom), _st(pos))


@block("S2.12", "APIC G2P + advect (recovers C, no FLIP/PIC blend)")
def k3_g2p_apic_advect(pos, vel, affine_C, u, v, w, dx, dt, nx, ny, nz, dom,
                       off_x=0, off_y=0, off_z=0):
    _offsets("S2.12", off_x, off_y, off_z)
    n = int(_t(pos).reshape(-1, 3).shape[0])
    _lib.call("S2.12", "gf_k3_g2p_apic_advect", _p(pos), _p(vel), _p(affine_C), _p(u), _p(v),
              _p(w), n, float(dx), float(dt), int(nx), int(ny), int(nz), _dom(dom), _st(pos))


@block("S2.10.GPU", "GPU reduction: max |v|^2 (CFL helper)")
def k3_vmax_sq_reduce(vel, out_max_sq):
    n = int(_t(vel).reshape(-1, 3).shape[0])
    _lib.call("S2.10.GPU", "gf_k3_vmax_sq_reduce", _p(vel), n, _p(out_max_sq), _st(vel))


# ---------------------------------------------- block-sparse and per-tile forms
def _blocks3(block_id, block_active, bs):
    t = _t(block_active)
    if t.dim() != 3 or t.dtype != torch.int32:
        raise BlockError(block_id, "block_active must be an int32 (nbx, nby, nbz) array")
    return (int(bs), *(int(x) for x in t.shape))


def _tiles(block_id, active_coords, n_active_dev, bs):
    c = _t(active_coords)
    if c.dtype != torch.int32 or c.numel() % 3:
        raise BlockError(block_id, "active_coords must be int32 block coordinates, 3 per block")
    if _t(n_active_dev).dtype != torch.int32:
        raise BlockError(block_id, "n_active_dev must be an int32 device array")
    return _p(active_coords), c.numel() // 3, int(bs), _p(n_active_dev)


@block("S2.16", "Active-block bitmask builder (8^3 tiles with any fluid cell)")
def k_mark_active_blocks(marker, block_active, block_size):
    """k_mark_active_blocks (solver3d.py:952-969): set block_active[b] = 1 for
    every block holding a fluid cell (others keep their value)."""
    nx, ny, nz = _cells(marker)
    bs, nbx, nby, nbz = _blocks3("S2.16", block_active, block_size)
    _lib.call("S2.16", "gf_k3_mark_active_blocks", _p(marker), _p(block_active), nx, ny, nz, bs,
              nbx, nby, nbz, _st(marker))


@block("S2.6.4", "Block-sparse Jacobi pressure (skip inactive 8^3 blocks)")
def k3_jacobi_pressure_blocksparse(p_in, p_out, div, marker, block_active, block_size):
    """k3_jacobi_pressure_blocksparse (solver3d.py:975-1041)."""
    nx, ny, nz = _cells(marker)
    bs, nbx, nby, nbz = _blocks3("S2.6.4", block_active, block_size)
    _lib.call("S2.6.4", "gf_k3_jacobi_pressure_blocksparse", _p(p_in), _p(p_out), _p(div),
              _p(marker), _p(block_active), nx, ny, nz, bs, nbx, nby, nbz, _st(p_in))


@block("S2.6.4", "Per-tile Jacobi pressure: launch only on active 8^3 blocks")
def k3_jacobi_pressure_per_tile(p_in, p_out, div, marker, active_coords, block_size,
                                n_active_dev, off_x=0, off_y=0, off_z=0):
    """k3_jacobi_pressure_per_tile (solver3d.py:1116-1197): the Jacobi sweep on
    the cells of the first n_active_dev[0] listed blocks."""
    _offsets("S2.6.4", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    co, nb, bs, na = _tiles("S2.6.4", active_coords, n_active_dev, block_size)
    _lib.call("S2.6.4", "gf_k3_jacobi_pressure_per_tile", _p(p_in), _p(p_out), _p(div),
              _p(marker), co, nb, bs, na, nx, ny, nz, _st(p_in))


@block("S2.6.5", "Per-tile GS-RB: red-black sweep on active 8^3 blocks only")
def k3_gauss_seidel_rb_per_tile(p, div, marker, active_coords, block_size, color, n_active_dev,
                                off_x=0, off_y=0, off_z=0):
    """k3_gauss_seidel_rb_per_tile (solver3d.py:1270-1357)."""
    _offsets("S2.6.5", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    co, nb, bs, na = _tiles("S2.6.5", active_coords, n_active_dev, block_size)
    _lib.call("S2.6.5", "gf_k3_gauss_seidel_rb_per_tile", _p(p), _p(div), _p(marker), co, nb,
              bs, int(color), na, nx, ny, nz, _st(p))


@block("S2.6.6", "Per-tile PCG: apply A on active blocks")
def k3_apply_A_per_tile(x, y, marker, active_coords, block_size, n_active_dev, done,
                        off_x=0, off_y=0, off_z=0):
    """k3_apply_A_per_tile (solver3d.py:2038-2105)."""
    _offsets("S2.6.6", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    co, nb, bs, na = _tiles("S2.6.6", active_coords, n_active_dev, block_size)
    _lib.call("S2.6.6", "gf_k3_apply_A_per_tile", _p(x), _p(y), _p(marker), co, nb, bs, na,
              _p(done), nx, ny, nz, _st(x))


@block("S2.6.6", "Per-tile PCG: apply M^-1 on active blocks")
def k3_apply_invM_per_tile(r, diag, z, marker, active_coords, block_size, n_active_dev, done,
                           off_x=0, off_y=0, off_z=0):
    """k3_apply_invM_per_tile (solver3d.py:2108-2141)."""
    _offsets("S2.6.6", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    co, nb, bs, na = _tiles("S2.6.6", active_coords, n_active_dev, block_size)
    _lib.call("S2.6.6", "gf_k3_apply_invM_per_tile", _p(r), _p(diag), _p(z), _p(marker), co, nb,
              bs, na, _p(done), nx, ny, nz, _st(r))


@block("S2.6.6", "Per-tile PCG: fluid-cell dot product on active blocks")
def k3_dot_fluid_per_tile(a, b, out, marker, active_coords, block_size, n_active_dev, done,
                          off_x=0, off_y=0, off_z=0):
    """k3_dot_fluid_per_tile (solver3d.py:2144-2176): adds the fluid cells'
    a*b of the listed blocks into out[0]."""
    _offsets("S2.6.6", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    co, nb, bs, na = _tiles("S2.6.6", active_coords, n_active_dev, block_size)
    _lib.call("S2.6.6", "gf_k3_dot_fluid_per_tile", _p(a), _p(b), _p(out), _p(marker), co, nb,
              bs, na, _p(done), nx, ny, nz, _st(a))


@block("S2.6.6", "Per-tile PCG: AXPY with device-side scalar on active blocks")
def k3_axpy_devscalar_per_tile(a_dev, sign, x, y_in, y_out, marker, active_coords, block_size,
                               n_active_dev, done, off_x=0, off_y=0, off_z=0):
    """k3_axpy_devscalar_per_tile (solver3d.py:2179-2212)."""
    _offsets("S2.6.6", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    co, nb, bs, na = _tiles("S2.6.6", active_coords, n_active_dev, block_size)
    _lib.call("S2.6.6", "gf_k3_axpy_devscalar_per_tile", _p(a_dev), float(sign), _p(x),
              _p(y_in), _p(y_out), _p(marker), co, nb, bs, na, _p(done), nx, ny, nz, _st(x))


# ------------------------------------------------ surface tension sub-blocks
def _grid3(a):
    t = _t(a)
    if t.dim() != 3:
        raise ValueError("expected a 3-D cell array")
    return tuple(int(x) for x in t.shape)


@block("S2.14.1", "Build raw fluid indicator field from marker (pre-blur)")
def k3_indicator_from_marker(marker, chi, off_x=0, off_y=0, off_z=0):
    """k3_indicator_from_marker (solver3d.py:212-229): chi = 1 on fluid, else 0."""
    _offsets("S2.14.1", off_x, off_y, off_z)
    nx, ny, nz = _cells(marker)
    _lib.call("S2.14.1", "gf_k3_indicator_from_marker", _p(marker), _p(chi), nx, ny, nz,
              _st(marker))


@block("S2.14.1", "Box-blur smoothing pass for chi -> chi~")
def k3_box_blur_centred(src, dst):
    """k3_box_blur_centred (solver3d.py:235-256): mean of the in-range 3x3x3
    neighbourhood."""
    nx, ny, nz = _grid3(src)
    _lib.call("S2.14.1", "gf_k3_box_blur_centred", _p(src), _p(dst), nx, ny, nz, _st(src))


@block("S2.14.2", "Unit-normal field n = grad chi~ / |grad chi~| on cell centres")
def k3_csf_normal(chi, n_x, n_y, n_z, dx):
    """k3_csf_normal (solver3d.py:262-286)."""
    nx, ny, nz = _grid3(chi)
    _lib.call("S2.14.2", "gf_k3_csf_normal", _p(chi), _p(n_x), _p(n_y), _p(n_z), nx, ny, nz,
              float(dx), _st(chi))


@block("S2.14.3", "Curvature kappa = -div n")
def k3_csf_curvature(n_x, n_y, n_z, kappa, dx):
    """k3_csf_curvature (solver3d.py:292-313)."""
    nx, ny, nz = _grid3(kappa)
    _lib.call("S2.14.3", "gf_k3_csf_curvature", _p(n_x), _p(n_y), _p(n_z), _p(kappa), nx, ny,
              nz, float(dx), _st(kappa))


def _csf_apply(comp, vel, chi, kappa, marker, sigma_dt_over_rho, dx, nx, ny, nz, offs):
    _offsets("S2.14.4", *offs)
    shape = _cells(marker)
    if (int(nx), int(ny), int(nz)) != shape:
        raise BlockError("S2.14.4", f"grid ({nx},{ny},{nz}) does not match marker {shape}")
    _lib.call("S2.14.4", "gf_k3_csf_apply", comp, _p(vel), _p(chi), _p(kappa), _p(marker),
              float(sigma_dt_over_rho), float(dx), *shape, _st(vel))


@block("S2.14.4", "Apply CSF impulse (sigma*kappa*grad chi~ * dt/rho) to MAC u faces")
def k3_csf_apply_u(u, chi, kappa, marker, sigma_dt_over_rho, dx, nx, ny, nz,
                   off_x=0, off_y=0, off_z=0):
    """k3_csf_apply_u (solver3d.py:316-345)."""
    _csf_apply(0, u, chi, kappa, marker, sigma_dt_over_rho, dx, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.14.4", "Apply CSF impulse to MAC v faces")
def k3_csf_apply_v(v, chi, kappa, marker, sigma_dt_over_rho, dx, nx, ny, nz,
                   off_x=0, off_y=0, off_z=0):
    """k3_csf_apply_v (solver3d.py:348-365)."""
    _csf_apply(1, v, chi, kappa, marker, sigma_dt_over_rho, dx, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.14.4", "Apply CSF impulse to MAC w faces")
def k3_csf_apply_w(w_, chi, kappa, marker, sigma_dt_over_rho, dx, nx, ny, nz,
                   off_x=0, off_y=0, off_z=0):
    """k3_csf_apply_w (solver3d.py:368-385)."""
    _csf_apply(2, w_, chi, kappa, marker, sigma_dt_over_rho, dx, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.14.5", "Host helper: capillary-wave dt_max for explicit CSF")
def csf_max_stable_dt(rho: float, dx: float, surface_tension: float, safety: float = 0.9) -> float:
    """csf_max_stable_dt (solver3d.py:2284-2307): safety * sqrt(rho dx^3 /
    (2 pi sigma)), +inf without surface tension."""
    import math
    if surface_tension <= 0.0:
        return float("inf")
    return float(safety) * math.sqrt(max(rho, 1e-12) * (dx ** 3) / (2.0 * math.pi * surface_tension))


_CSF_RED: dict = {}


def _csf_sum(comp, before, after, marker, total, count, nx, ny, nz, offs):
    _offsets("S2.14.6", *offs)
    shape = _cells(marker)
    if (int(nx), int(ny), int(nz)) != shape:
        raise BlockError("S2.14.6", f"grid ({nx},{ny},{nz}) does not match marker {shape}")
    dev = _t(marker).device
    red = _CSF_RED.get((shape, dev))
    if red is None:
        n = ctypes.c_int64()
        _lib.call("S2.14.6", "gf_k3_csf_sum_workspace", *shape, ctypes.byref(n))
        red = torch.empty(n.value, dtype=torch.float64, device=dev)
        _CSF_RED[(shape, dev)] = red
    _lib.call("S2.14.6", "gf_k3_csf_sum", comp, _p(before), _p(after), _p(marker), _p(total),
              _p(count), *shape, _lib.ptr(red), _st(marker))


@block("S2.14.6", "Force-balance CSF: sum the u impulse over faces next to fluid")
def k3_csf_sum_u(u_before, u_after, marker, sum_du, count, nx, ny, nz, off_x=0, off_y=0, off_z=0):
    """k3_csf_sum_u (solver3d.py:400-420): adds sum(after - before) and the
    face count over u faces next to fluid to sum_du[0] and count[0]."""
    _csf_sum(0, u_before, u_after, marker, sum_du, count, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.14.6", "Force-balance CSF: sum the v impulse")
def k3_csf_sum_v(v_before, v_after, marker, sum_dv, count, nx, ny, nz, off_x=0, off_y=0, off_z=0):
    """k3_csf_sum_v (solver3d.py:423-443)."""
    _csf_sum(1, v_before, v_after, marker, sum_dv, count, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.14.6", "Force-balance CSF: sum the w impulse")
def k3_csf_sum_w(w_before, w_after, marker, sum_dw, count, nx, ny, nz, off_x=0, off_y=0, off_z=0):
    """k3_csf_sum_w (solver3d.py:446-464)."""
    _csf_sum(2, w_before, w_after, marker, sum_dw, count, nx, ny, nz, (off_x, off_y, off_z))


def _csf_bias(comp, vel, marker, total, count, nx, ny, nz, offs):
    _offsets("S2.14.6", *offs)
    shape = _cells(marker)
    if (int(nx), int(ny), int(nz)) != shape:
        raise BlockError("S2.14.6", f"grid ({nx},{ny},{nz}) does not match marker {shape}")
    _lib.call("S2.14.6", "gf_k3_csf_subtract_bias", comp, _p(vel), _p(marker), _p(total),
              _p(count), *shape, _st(vel))


@block("S2.14.6", "CSF force-balance on u (device-side, no host sync)")
def k3_csf_subtract_bias_u_dev(u, marker, sum_dev, count_dev, nx, ny, nz,
                               off_x=0, off_y=0, off_z=0):
    """k3_csf_subtract_bias_u_dev (solver3d.py:478-500)."""
    _csf_bias(0, u, marker, sum_dev, count_dev, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.14.6", "CSF force-balance on v (device-side)")
def k3_csf_subtract_bias_v_dev(v, marker, sum_dev, count_dev, nx, ny, nz,
                               off_x=0, off_y=0, off_z=0):
    """k3_csf_subtract_bias_v_dev (solver3d.py:503-524)."""
    _csf_bias(1, v, marker, sum_dev, count_dev, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.14.6", "CSF force-balance on w (device-side)")
def k3_csf_subtract_bias_w_dev(w_, marker, sum_dev, count_dev, nx, ny, nz,
                               off_x=0, off_y=0, off_z=0):
    """k3_csf_subtract_bias_w_dev (solver3d.py:527-547)."""
    _csf_bias(2, w_, marker, sum_dev, count_dev, nx, ny, nz, (off_x, off_y, off_z))


# ------------------------------------------ colour / scalar transfer blocks
_ATTR_WS: dict = {}


def _attr_ws(channels, shape, dev):
    key = (channels, shape, dev)
    ws = _ATTR_WS.get(key)
    if ws is None:
        cells = shape[0] * shape[1] * shape[2]
        ws = (torch.empty((channels + 1) * cells, dtype=torch.int64, device=dev),
              torch.empty(4, dtype=torch.int32, device=dev))
        _ATTR_WS[key] = ws
    return ws


def _attr_grid_shape(block_id, grids, nx, ny, nz, offs):
    _offsets(block_id, *offs)
    shape = _grid3(grids[0])
    if (int(nx), int(ny), int(nz)) != shape or any(_grid3(g) != shape for g in grids):
        raise BlockError(block_id, f"grids do not all have the shape ({nx},{ny},{nz})")
    return shape


def _p2g_attr(block_id, pos, attr, channels, grids, dx, nx, ny, nz, offs):
    shape = _attr_grid_shape(block_id, grids, nx, ny, nz, offs)
    n = int(_t(pos).reshape(-1, 3).shape[0])
    if _t(attr).numel() != n * channels:
        raise BlockError(block_id, f"attribute holds {_t(attr).numel()} values for {n} particles")
    acc, scratch = _attr_ws(channels, shape, _t(pos).device)
    g = [_p(x) for x in grids]
    if channels == 1:
        g = [g[0], None, None, g[1]]
    _lib.call(block_id, "gf_k3_p2g_attr", _p(pos), _p(attr), channels, n, *g, *shape, float(dx),
              _lib.ptr(acc), _lib.ptr(scratch), _st(pos))


def _g2p_attr(block_id, pos, attr, channels, grids, dx, nx, ny, nz, offs):
    shape = _attr_grid_shape(block_id, grids, nx, ny, nz, offs)
    n = int(_t(pos).reshape(-1, 3).shape[0])
    if _t(attr).numel() != n * channels:
        raise BlockError(block_id, f"attribute holds {_t(attr).numel()} values for {n} particles")
    g = [_p(x) for x in grids]
    if channels == 1:
        g = [g[0], None, None, g[1]]
    _lib.call(block_id, "gf_k3_g2p_attr", _p(pos), _p(attr), channels, n, *g, *shape, float(dx),
              _st(pos))


@block("S2.15.1", "P2G scatter of per-particle RGB color to cell-centred grid")
def k3_p2g_color(pos, color, cgrid_r, cgrid_g, cgrid_b, cgrid_w, dx, nx, ny, nz,
                 off_x=0, off_y=0, off_z=0):
    """k3_p2g_color (solver3d.py:562-600): adds the trilinear sums of
    colour*w and w at the cell centres to the four grids."""
    _p2g_attr("S2.15.1", pos, color, 3, (cgrid_r, cgrid_g, cgrid_b, cgrid_w), dx, nx, ny, nz,
              (off_x, off_y, off_z))


@block("S2.15.2", "Normalize grid color by deposited weight")
def k3_normalize_color(cgrid_r, cgrid_g, cgrid_b, cgrid_w):
    """k3_normalize_color (solver3d.py:607-631)."""
    shape = _grid3(cgrid_w)
    _lib.call("S2.15.2", "gf_k3_normalize_attr", _lib.GF_ATTR_COLOR, 3, _p(cgrid_r),
              _p(cgrid_g), _p(cgrid_b), _p(cgrid_w), *shape, _st(cgrid_w))


@block("S2.15.3", "G2P gather grid color back to particle")
def k3_g2p_color(pos, color, cgrid_r, cgrid_g, cgrid_b, cgrid_w, dx, nx, ny, nz,
                 off_x=0, off_y=0, off_z=0):
    """k3_g2p_color (solver3d.py:633-664)."""
    _g2p_attr("S2.15.3", pos, color, 3, (cgrid_r, cgrid_g, cgrid_b, cgrid_w), dx, nx, ny, nz,
              (off_x, off_y, off_z))


@block("S2.18.1", "P2G scatter of per-particle scalar onto grid")
def k3_p2g_scalar(pos, attr, sgrid_v, sgrid_w, dx, nx, ny, nz, off_x=0, off_y=0, off_z=0):
    """k3_p2g_scalar (solver3d.py:670-707)."""
    _p2g_attr("S2.18.1", pos, attr, 1, (sgrid_v, sgrid_w), dx, nx, ny, nz, (off_x, off_y, off_z))


@block("S2.18.2", "Normalize grid scalar by deposited weight")
def k3_normalize_scalar(sgrid_v, sgrid_w):
    """k3_normalize_scalar (solver3d.py:710-727)."""
    shape = _grid3(sgrid_w)
    _lib.call("S2.18.2", "gf_k3_normalize_attr", _lib.GF_ATTR_SCALAR, 1, _p(sgrid_v), None, None,
              _p(sgrid_w), *shape, _st(sgrid_w))


@block("S2.18.3", "G2P gather grid scalar back to particle")
def k3_g2p_scalar(pos, attr, sgrid_v, sgrid_w, dx, nx, ny, nz, off_x=0, off_y=0, off_z=0):
    """k3_g2p_scalar (solver3d.py:729-755)."""
    _g2p_attr("S2.18.3", pos, attr, 1, (sgrid_v, sgrid_w), dx, nx, ny, nz, (off_x, off_y, off_z))
