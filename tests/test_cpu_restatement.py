"""The C restatement of the reference substep (oracle/flip_cpu.c), which bench.py
times as the CPU baseline, against the numpy FP64 oracle.

Runs on the host: the program is compiled by oracle/Makefile and fed a small
seeded scene through its particle file; its output particles after 1 and 3
substeps must match flip_oracle.substep (exact pressure solve) to 1e-5
relative, the substep bar of the GPU path.  flip_cpu.c computes in FP32 with
the reference's Jacobi-PCG (tolerance 1e-6 here, checked every 10
iterations), so the remaining difference is FP32 rounding and the 1e-6 solve.
"""
from __future__ import annotations

import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import gfutil as gu
from oracle import flip_oracle as fo

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "oracle" / "_build" / "flip_cpu"


def _exe():
    if shutil.which("make") is None or shutil.which("gcc") is None:
        pytest.skip("make/gcc not available")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    return EXE


def _run(tmp_path, n, dx, dt, pos, vel, substeps, tol=1e-6, max_iter=1000, threads=2,
         transfer="flip"):
    exe = _exe()
    src = tmp_path / "in.f32"
    out = tmp_path / "out.f32"
    np.concatenate([pos, vel], axis=1).astype(np.float32).tofile(src)
    r = subprocess.run([str(exe), str(n), repr(dx), repr(dt), repr(tol), str(max_iter),
                        str(substeps), str(src), str(len(pos)), str(out), transfer],
                       capture_output=True, text=True, check=True,
                       env={"OMP_NUM_THREADS": str(threads), "PATH": "/usr/bin:/bin"})
    info = json.loads(r.stdout.strip().splitlines()[-1])
    data = np.fromfile(out, dtype=np.float32)
    npart = len(pos)
    p, v = data[:3 * npart].reshape(npart, 3), data[3 * npart:6 * npart].reshape(npart, 3)
    if transfer == "apic":
        return p, v, info, data[6 * npart:].reshape(npart, 3, 3)
    return p, v, info


@pytest.mark.parametrize("substeps", [1, 3])
def test_flip_cpu_matches_the_fp64_oracle(tmp_path, substeps):
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=7, vel_sigma=0.3)
    dt = 1.0 / 240
    cp, cv, info = _run(tmp_path, 16, dx, dt, pos, vel, substeps)
    assert info["iters"] > 0
    st = gu.oracle_state(shape, dx, pos, vel)
    for _ in range(substeps):
        st, _ = fo.substep(st, dt, pressure="exact")
    ev, ep = gu.rel_err(cv, st["vel"]), gu.rel_err(cp, st["pos"])
    print(f"flip_cpu {substeps} substep(s), PCG {info['iters']} it: rel err vel {ev:.2e} pos {ep:.2e}")
    assert ev < 1e-5 and ep < 1e-5


def test_flip_cpu_thread_count_does_not_change_the_result(tmp_path):
    """The race-free slab schedule makes the output independent of the
    number of threads (a data race would not be)."""
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=8, vel_sigma=0.3)
    a = _run(tmp_path, 16, dx, 1 / 240, pos, vel, 2, threads=1)
    b = _run(tmp_path, 16, dx, 1 / 240, pos, vel, 2, threads=4)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_flip_cpu_catches_a_planted_sign_error(tmp_path):
    """The comparison is sharp enough to see a flipped pressure sign: the
    oracle substep with the reference PCG's literal sign (x = -p) differs from
    flip_cpu by far more than the tolerance."""
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=7, vel_sigma=0.3)
    dt = 1.0 / 240
    cp, cv, _ = _run(tmp_path, 16, dx, dt, pos, vel, 1)
    st, _ = fo.substep(gu.oracle_state(shape, dx, pos, vel), dt, pressure="pcg", iters=1000,
                       tol=1e-10, pcg_sign="reference")
    assert gu.rel_err(cv, st["vel"]) > 1e-2


@pytest.mark.parametrize("substeps", [1, 3])
def test_flip_cpu_apic_matches_the_fp64_oracle(tmp_path, substeps):
    """The APIC transfer of the C restatement (the CPU baseline of the APIC
    configuration) against the oracle's APIC substep: positions, velocities
    and the affine matrices after 1 and 3 substeps from C = 0."""
    shape, dx, pos, vel = gu.dam_scene(n=16, seed=9, vel_sigma=0.3)
    dt = 1.0 / 240
    cp, cv, info, cc = _run(tmp_path, 16, dx, dt, pos, vel, substeps, transfer="apic")
    assert info["iters"] > 0 and "APIC" in info["sample"]
    st = gu.oracle_state(shape, dx, pos, vel)
    for _ in range(substeps):
        st, _ = fo.substep(st, dt, pressure="exact", transfer="apic")
    ev, ep = gu.rel_err(cv, st["vel"]), gu.rel_err(cp, st["pos"])
    ec = gu.rel_err(cc, st["affine"])
    print(f"flip_cpu apic {substeps} substep(s): rel err vel {ev:.2e} pos {ep:.2e} C {ec:.2e}")
    assert ev < 1e-5 and ep < 1e-5 and ec < 1e-4
    # the affine terms matter: the FLIP run of the same scene differs
    fp, fv, _ = _run(tmp_path, 16, dx, dt, pos, vel, substeps)
    if substeps > 1:
        assert gu.rel_err(fv, cv) > 1e-3
