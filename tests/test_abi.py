"""C-ABI boundary checks that need no GPU.

* every function declared in include/gpufluid_hip.h is exported by the built
  libgpufluid_hip.so and has a ctypes signature with the same arity;
* the ctypes mirrors of gf_scene_t / gf_scene_size_info have the C layout
  (offsets and sizes printed by a gcc-compiled probe of the header);
* the S2 block registry exposes every block id the reference registers on this
  path, with the reference's positional argument order.
"""
from __future__ import annotations

import ctypes
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gpufluid_hip.h"


def _declarations():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^(?:int|const char\*)\s+(gf_\w+)\s*\(([^;]*?)\)\s*;", text, re.M | re.S):
        params = " ".join(m.group(2).split())
        n = 0 if params in ("", "void") else params.count(",") + 1
        out[m.group(1)] = n
    return out


def _library():
    from gpufluid_amd import build
    if not build.LIB.exists():
        build.build()
    return build.LIB


def test_header_declares_the_substep_entry_points():
    decl = _declarations()
    for name in ("gf_scene_sort_sync", "gf_scene_p2g", "gf_scene_jacobi", "gf_scene_pcg_sync",
                 "gf_scene_g2p", "gf_k3_jacobi_pressure", "gf_k3_g2p_and_advect", "gf_last_error"):
        assert name in decl, name


def test_ctypes_signatures_match_header_arity():
    from gpufluid_amd import _lib
    decl = _declarations()
    for name, n in decl.items():
        if name == "gf_last_error":
            continue
        assert name in _lib.SIGNATURES, f"{name} declared but has no ctypes signature"
        assert len(_lib.SIGNATURES[name]) == n, (name, len(_lib.SIGNATURES[name]), n)
    for name in _lib.SIGNATURES:
        assert name in decl, f"{name} has a ctypes signature but is not declared in the header"


def test_library_exports_every_declared_symbol():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    lib = _library()
    out = subprocess.run([nm, "-D", "--defined-only", str(lib)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(set(_declarations()) - exported)
    assert not missing, f"declared but not exported: {missing}"


def test_library_loads_and_reports_version():
    from gpufluid_amd import _lib
    _library()
    lib = _lib.lib()
    assert lib.gf_version() >= 10000
    assert isinstance(lib.gf_last_error(), bytes)


def test_scene_sizes_without_gpu():
    from gpufluid_amd import _lib
    _library()
    sz = _lib.sizes(256, 256, 256)
    assert sz.cells == 256 ** 3
    assert sz.ntiles == 32 ** 3
    assert sz.compact == 32 ** 3 * 512
    # a grid that overflows int32 cell indexing is rejected with a message
    bad = _lib.GfSizes()
    rc = _lib.lib().gf_scene_sizes(2048, 2048, 1024, ctypes.byref(bad))
    assert rc == _lib.GF_ERANGE
    assert b"int32" in _lib.lib().gf_last_error()


def _c_layout(struct_name, fields):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    import tempfile
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {",
             f'  printf("size %zu\\n", sizeof({struct_name}));']
    for f in fields:
        lines.append(f'  printf("{f} %zu\\n", offsetof({struct_name}, {f}));')
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as td:
        src = Path(td) / "probe.c"
        exe = Path(td) / "probe"
        src.write_text("\n".join(lines) + "\n")
        subprocess.run([gcc, "-std=c11", str(src), "-o", str(exe)], check=True, capture_output=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


@pytest.mark.parametrize("cname,pyname", [("gf_scene_t", "GfScene"),
                                          ("gf_scene_size_info", "GfSizes"),
                                          ("gf_adjoint_t", "GfAdjoint")])
def test_ctypes_struct_layout_matches_c(cname, pyname):
    from gpufluid_amd import _lib
    cls = getattr(_lib, pyname)
    names = [f[0] for f in cls._fields_]
    c = _c_layout(cname, names)
    assert c["size"] == ctypes.sizeof(cls)
    for n in names:
        assert c[n] == getattr(cls, n).offset, n


# Block ids the reference registers on this path (solver3d.py:114, 142, 156,
# 799, 871, 938, 1407, 1463, 1588, 1732, 1800, 2012).
S2_BLOCKS = {
    "S2.1": "k3_p2g", "S2.2": "k3_normalize", "S2.3": "k3_add_gravity",
    "S2.4": "k3_enforce_solid_bc", "S2.5": "k3_compute_divergence",
    "S2.6.1": "k3_jacobi_pressure", "S2.6.3": "k3_apply_A", "S2.7": "k3_subtract_pressure_grad",
    "S2.8": "k3_g2p_and_advect", "S2.12": "k3_g2p_apic_advect",
}


def test_block_registry_covers_the_substep_path():
    from gpufluid_amd import blocks, kernels  # noqa: F401  (importing kernels registers them)
    for bid, name in S2_BLOCKS.items():
        names = {b.qualname for b in blocks.find(bid)}
        assert name in names, (bid, name, names)
        assert getattr(kernels, name).__block_id__ == bid


def test_block_signatures_follow_reference_argument_order():
    import inspect
    from gpufluid_amd import kernels
    # k3_jacobi_pressure(p_in, p_out, div, marker, off_x, off_y, off_z): solver3d.py:878-885
    assert list(inspect.signature(kernels.k3_jacobi_pressure).parameters) == [
        "p_in", "p_out", "div", "marker", "off_x", "off_y", "off_z"]
    # k3_compute_divergence(u, v, w, div, marker, dx, dt, rho, offs): solver3d.py:846-853
    assert list(inspect.signature(kernels.k3_compute_divergence).parameters)[:8] == [
        "u", "v", "w", "div", "marker", "dx", "dt", "rho"]
