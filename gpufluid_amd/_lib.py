"""ctypes binding of libgpufluid_hip.so (include/gpufluid_hip.h).

The library is the only compute path: if it is missing or fails to load,
every entry point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import threading
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_uint8, c_void_p
from pathlib import Path

from .blocks import BlockError

LIB_PATH = Path(__file__).resolve().with_name("libgpufluid_hip.so")

GF_TRANSFER_FLIP = 0
GF_TRANSFER_APIC = 1
GF_PCG_MIXED = 0
GF_PCG_FP64 = 1
GF_ATTR_COLOR = 0
GF_ATTR_SCALAR = 1
GF_NACT_DEVICE = -1  # n_active argument: read the active tile count on the device
GF_EINVAL, GF_ERANGE, GF_ESTATE = -1, -2, -3  # negative return codes of the C-ABI
TILE = 8
TILE_CELLS = 512
SCENE_PAR = 16  # GF_SCENE_PAR: floats per scene in gf_scene_t.scene_par

_lock = threading.Lock()
_lib = None


class GfScene(ctypes.Structure):
    """Mirror of gf_scene_t (include/gpufluid_hip.h)."""

    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("tnx", c_int32), ("tny", c_int32), ("tnz", c_int32),
        ("dx", c_float), ("dom", c_float * 3),
        ("n", c_int32), ("cap", c_int32),
        ("pos", c_void_p), ("vel", c_void_p), ("affine_c", c_void_p), ("pid", c_void_p),
        ("pos_b", c_void_p), ("vel_b", c_void_p), ("affine_c_b", c_void_p), ("pid_b", c_void_p),
        ("keys", c_void_p), ("tmp", c_void_p), ("counts", c_void_p), ("cell_start", c_void_p),
        ("scan_part", c_void_p),
        ("tile_occ", c_void_p), ("tile_map", c_void_p), ("active", c_void_p),
        ("n_active", c_void_p), ("nbr", c_void_p),
        ("marker", c_void_p),
        ("u", c_void_p), ("v", c_void_p), ("w", c_void_p),
        ("us", c_void_p), ("vs", c_void_p), ("ws", c_void_p),
        ("solid_u", c_void_p), ("solid_v", c_void_p), ("solid_w", c_void_p),
        ("flags", c_void_p), ("b", c_void_p),
        ("x32", c_void_p), ("r32", c_void_p), ("p32", c_void_p), ("ap32", c_void_p),
        ("x64", c_void_p), ("r64", c_void_p), ("p64", c_void_p), ("ap64", c_void_p),
        ("pc", c_void_p), ("ptmp", c_void_p),
        ("partials", c_void_p), ("scalars", c_void_p),
        ("p32b", c_void_p), ("p64b", c_void_p), ("cg32", c_void_p), ("cg64", c_void_p),
        ("n_scenes", c_int32), ("tile_scene", c_void_p), ("tile_order", c_void_p),
        ("scene_par", c_void_p), ("scene_slot", c_void_p), ("pcg_part", c_void_p),
        ("scene_pstart", c_void_p),
    ]


class GfAdjoint(ctypes.Structure):
    """Mirror of gf_adjoint_t (include/gpufluid_hip.h)."""

    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("dx", c_float), ("dom", c_float * 3), ("n", c_int32),
        ("pos", c_void_p), ("vel", c_void_p), ("pid", c_void_p),
        ("marker", c_void_p), ("tile_map", c_void_p),
        ("u", c_void_p), ("v", c_void_p), ("w", c_void_p),
        ("us", c_void_p), ("vs", c_void_p), ("ws", c_void_p),
        ("mu", c_void_p), ("mv", c_void_p), ("mw", c_void_p),
        ("au", c_void_p), ("av", c_void_p), ("aw", c_void_p),
        ("aus", c_void_p), ("avs", c_void_p), ("aws", c_void_p),
        ("apos", c_void_p), ("avel", c_void_p),
    ]


class GfSizes(ctypes.Structure):
    _fields_ = [("cells", c_int64), ("ntiles", c_int64), ("compact", c_int64),
                ("scan_parts", c_int64), ("partials", c_int64), ("scalars", c_int64),
                ("pcg_part", c_int64)]


V, I, F, D = c_void_p, c_int, c_float, c_double
PI, PD = POINTER(c_int), POINTER(c_double)
SP = POINTER(GfScene)
AP = POINTER(GfAdjoint)

# name -> argtypes (all return int)
SIGNATURES = {
    "gf_version": [],
    "gf_tune_set": [ctypes.c_char_p, I],
    "gf_tune_get": [ctypes.c_char_p, PI],
    "gf_init": [I, PI],
    "gf_scene_sizes": [I, I, I, POINTER(GfSizes)],
    "gf_scene_sort_sync": [SP, I, PI, V],
    "gf_scene_p2g": [SP, I, F, F, F, V],
    "gf_scene_p2g_stage_size": [I, I, I, V, V],
    "gf_scene_p2g_staged": [SP, I, F, F, F, V, V, V],
    "gf_scene_p2g_raw": [SP, I, V, V, V, V],
    "gf_scene_jacobi": [SP, I, I, V],
    "gf_scene_gsrb": [SP, I, I, V],
    "gf_scene_pcg_sync": [SP, I, I, D, I, I, I, PI, PD, PI, PD, V],
    "gf_scene_pcg_stats": [SP, PI, PD, PI, V],
    "gf_scene_pcg_flags": [SP, F, V, V, V],
    "gf_scene_pcg_enqueue": [SP, I, I, D, I, I, I, I, I, V],
    "gf_scene_pcg_enqueue_noalloc": [SP, I, I, D, I, I, I, I, I, V],
    "gf_scene_g2p": [SP, V, I, F, F, F, F, V],
    "gf_scene_attr_transfer": [SP, V, I, I, V, V, V, V],
    "gf_anim_obstacles": [V, I, V, V, V, V, V, I, I, I, D, V],
    "gf_k3_p2g_attr": [V, V, I, I, V, V, V, V, I, I, I, F, V, V, V],
    "gf_k3_normalize_attr": [I, I, V, V, V, V, I, I, I, V],
    "gf_k3_g2p_attr": [V, V, I, I, V, V, V, V, I, I, I, F, V],
    "gf_k3_indicator_from_marker": [V, V, I, I, I, V],
    "gf_k3_box_blur_centred": [V, V, I, I, I, V],
    "gf_k3_csf_normal": [V, V, V, V, I, I, I, F, V],
    "gf_k3_csf_curvature": [V, V, V, V, I, I, I, F, V],
    "gf_k3_csf_apply": [I, V, V, V, V, F, F, I, I, I, V],
    "gf_k3_csf_sum_workspace": [I, I, I, V],
    "gf_k3_csf_sum": [I, V, V, V, V, V, I, I, I, V, V],
    "gf_k3_csf_subtract_bias": [I, V, V, V, V, I, I, I, V],
    "gf_k3_mark_active_blocks": [V, V, I, I, I, I, I, I, I, V],
    "gf_k3_jacobi_pressure_blocksparse": [V, V, V, V, V, I, I, I, I, I, I, I, V],
    "gf_k3_jacobi_pressure_per_tile": [V, V, V, V, V, I, I, V, I, I, I, V],
    "gf_k3_gauss_seidel_rb_per_tile": [V, V, V, V, I, I, I, V, I, I, I, V],
    "gf_k3_apply_A_per_tile": [V, V, V, V, I, I, V, V, I, I, I, V],
    "gf_k3_apply_invM_per_tile": [V, V, V, V, V, I, I, V, V, I, I, I, V],
    "gf_k3_dot_fluid_per_tile": [V, V, V, V, V, I, I, V, V, I, I, I, V],
    "gf_k3_axpy_devscalar_per_tile": [V, F, V, V, V, V, V, I, I, V, V, I, I, I, V],
    "gf_mesh_workspace": [I, I, I, V],
    "gf_mesh_bin_sync": [V, I, I, I, I, D, V, V, V],
    "gf_mesh_mark": [V, I, I, I, I, D, V, V, c_int64, V, V, V, V, F, F, F, V],
    "gf_reseed_workspace": [I, I, I, I, V],
    "gf_reseed_plan_sync": [V, I, V, I, I, I, F, I, I, I, V, PI, V],
    "gf_reseed_compact": [V, I, I, I, I, I, V, V, V],
    "gf_reseed_emit": [I, I, I, I, F, V, V, V, V],
    "gf_outflow_mark_sync": [V, I, V, I, V, V, V, PI, V],
    "gf_compact_rows": [V, I, I, V, V, V, V],
    "gf_scene_g2p_pre": [SP, V, V, V, V, I, F, F, F, V],
    "gf_scene_divergence": [SP, F, F, V],
    "gf_csf_workspace": [I, I, I, V, V],
    "gf_csf_apply": [V, V, V, V, I, I, I, F, F, I, V, V, V],
    "gf_scene_compact_to_dense": [SP, V, V, V],
    "gf_scene_unsort": [SP, V, I, V, V],
    "gf_k3_clear_grid": [V, V, V, V, V, V, V, I, I, I, V],
    "gf_k3_mark_fluid": [V, I, V, F, I, I, I, V],
    "gf_k3_normalize": [V, V, V, V, V, V, V, V, V, I, I, I, V],
    "gf_k3_add_gravity": [V, I, I, I, F, F, V],
    "gf_k3_enforce_solid_bc": [V, V, V, V, V, V, V, I, I, I, V],
    "gf_k3_compute_divergence": [V, V, V, V, V, I, I, I, F, F, F, V],
    "gf_k3_jacobi_pressure": [V, V, V, V, I, I, I, V],
    "gf_k3_gauss_seidel_rb": [V, V, V, I, I, I, I, V],
    "gf_k3_jacobi_visc": [V, V, V, I, I, I, F, V],
    "gf_k3_jacobi_visc3": [V, V, V, V, V, V, V, V, V, I, I, I, F, V],
    "gf_k3_jacobi_visc3_box": [V, V, V, V, V, V, V, V, V, I, I, I, I, I, I, I, I, I, F, V],
    "gf_k3_apply_A": [V, V, V, V, I, I, I, V],
    "gf_k3_compute_diag": [V, V, I, I, I, V],
    "gf_k3_apply_invM": [V, V, V, V, V, I, I, I, V],
    "gf_k3_dot_fluid": [V, V, V, V, V, I, I, I, V],
    "gf_k3_axpy_devscalar": [V, F, V, V, V, V, V, I, I, I, V],
    "gf_k3_scalar_op": [I, V, V, V, F, V, V],
    "gf_k3_subtract_pressure_grad": [V, V, V, V, V, I, I, I, F, F, F, V],
    "gf_k3_g2p_and_advect": [V, V, V, V, V, V, V, V, I, F, F, I, I, I, F, POINTER(c_float), V],
    "gf_k3_g2p_apic_advect": [V, V, V, V, V, V, I, F, F, I, I, I, POINTER(c_float), V],
    "gf_k3_vmax_sq_reduce": [V, I, V, V],
    "gf_adj_clear": [AP, V],
    "gf_adj_weights": [AP, V],
    "gf_adj_g2p": [AP, F, F, V],
    "gf_adj_pressure_rhs": [AP, F, F, V, V],
    "gf_adj_faces": [AP, V, F, F, V],
    "gf_adj_p2g": [AP, V],
}


def lib() -> ctypes.CDLL:
    """Load the HIP library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH.name} is not built: run `python -m gpufluid_amd.build` "
                    "(the HIP library is the only compute path)")
            h = ctypes.CDLL(str(LIB_PATH))
            for name, args in SIGNATURES.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = c_int
            h.gf_last_error.argtypes = []
            h.gf_last_error.restype = ctypes.c_char_p
            _lib = h
    return _lib


def last_error() -> str:
    return lib().gf_last_error().decode(errors="replace")


def call(block_id: str, name: str, *args) -> None:
    """Invoke an entry point; map a nonzero return to BlockError(block_id)."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise BlockError(block_id, f"{name} failed ({rc}): {last_error()}")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def sizes(nx: int, ny: int, nz: int) -> GfSizes:
    s = GfSizes()
    call("F3.2", "gf_scene_sizes", nx, ny, nz, ctypes.byref(s))
    return s
