"""Substep throughput on one MI355X: 256^3 pool + sphere (P2G + PCG + G2P).

python bench.py [--steps K] [--warmup W] [--config c3|c2|c1] [--no-cpu-baseline]

One step = one full substep on ~40.2 M particles: stable counting sort by cell,
fused P2G (+normalize, gravity, BC, divergence), mixed-precision PCG to a 1e-8
relative residual (cap 500 iterations), fused gradient subtract + G2P + advection.
Particles start resident in HBM.  Prints one JSON line (rank 0, N = 1).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "Mparticle-substeps/s at 256³ (P2G+PCG+G2P), % HBM peak; adjoint/fwd time"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per unit (fluid cells N_f, particles N_p, active-tile
# cells N_a) of each kernel's compulsory inputs and outputs.
BYTES = {
    "sort": lambda np_, nf, na: 68 * np_ + 12 * na,
    "p2g": lambda np_, nf, na: 24 * np_ + 29 * nf,
    "g2p": lambda np_, nf, na: 48 * np_ + 41 * nf,
    # two-kernel PCG iteration (FP32 vectors, fluid cells only):
    # direction+SpMV: z, p_old, flags in; p_new out (Ap only reduced)
    "pcg_dirspmv": lambda np_, nf, na: 13 * nf,
    # update (Ap recomputed from p): p, z, x32, flags in; z, x32 out
    "pcg_update": lambda np_, nf, na: 21 * nf,
    # FP64 replacement: x64, x32, b, flags in; z out
    "pcg_replace": lambda np_, nf, na: 21 * nf,
    # one-kernel (Chronopoulos-Gear) iteration: u, w~, s~, p, x32, flags in;
    # u, w~, s~, p, x32 out
    "pcg_iter": lambda np_, nf, na: 41 * nf,
    # its replacement: x64, x32, b, p, flags in -> u, s~ out; then u, flags in -> w~ out
    "pcg_replace_cg1": lambda np_, nf, na: 38 * nf,
    # one PCG iteration as a stage, at the figure of the minimal fused FP32
    # iteration (SpMV p, flags -> Ap: 9; x, r read+write, p, Ap read: 24;
    # r, flags, p read+write: 13), whatever kernels implement it
    "pcg_iteration": lambda np_, nf, na: 46 * nf,
}


# Measured HBM traffic per dispatch (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled
# by a plain-copy calibration) for this workload; the file states its commands.
TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_substep_pool_sphere_256.txt"
TRAFFIC_KERNEL = {"pcg_dirspmv": ["k_pcg_dirspmv<float,true>"], "pcg_update": ["k_pcg_update<float,true>"],
                  "pcg_replace": ["k_pcg_replace<true>"], "pcg_iter": ["k_cg1_iter<float,true>"],
                  "p2g": ["k_p2g_cells_flip<false>", "k_p2g_finish"], "g2p": ["k_g2p_tile<false>"]}


def measured_traffic(stage):
    """HBM bytes per launch of a stage's kernels from the committed PMC summary, or None."""
    try:
        lines = TRAFFIC_FILE.read_text().splitlines()
    except OSError:
        return None
    found = {}
    for line in lines:
        f = line.split()
        if len(f) == 5:
            found[f[0]] = int(f[4])
    want = TRAFFIC_KERNEL.get(stage, [])
    if not want or any(k not in found for k in want):
        return None
    return sum(found[k] for k in want)


def _spread(iters, residuals, cap):
    """min / median / max of the timed solves' iteration counts and FP64 true
    relative residuals, and how many stopped at the cap."""
    import numpy as np
    r = np.asarray(residuals, dtype=np.float64)
    return {"iters_min": int(min(iters)), "iters_median": float(np.median(iters)),
            "iters_max": int(max(iters)), "at_cap": int(sum(i >= cap for i in iters)),
            "rel_residual_min": float(r.min()), "rel_residual_median": float(np.median(r)),
            "rel_residual_max": float(r.max()), "rel_residuals": [float(x) for x in r]}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_scene(name):
    from gpufluid_amd import seeding
    return {"c1": seeding.config_c1, "c2": seeding.config_c2, "c3": seeding.config_c3}[name]()


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_topology():
    """lscpu's view of this host: model, threads per core, and the CPUs this
    process may run on (its affinity mask, which a container or a scheduler's
    CPU share sets)."""
    info = {"model": _cpu_model(), "cpus_allowed": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Socket(s)":
                info["sockets"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    tpc = info.get("threads_per_core", 1)
    info["physical_cores_allowed"] = max(1, info["cpus_allowed"] // max(1, tpc))
    return info


def _run_cpu(exe, td, scene, substeps, threads, max_iter=None):
    import numpy as np
    p = scene.params
    pfile = Path(td) / f"{scene.name}.f32"
    if not pfile.exists():
        np.concatenate([scene.pos, scene.vel], axis=1).astype(np.float32).tofile(pfile)
    cmd = [str(exe), str(scene.shape[0]), repr(scene.dx), repr(p.get("dt", 1 / 240)),
           repr(p.get("tol", 1e-8)), str(max_iter or p.get("max_iter", 500)), str(substeps),
           str(pfile), str(len(scene.pos)), "-", p.get("transfer", "flip")]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        log(f"cpu baseline failed: {r.stderr[-400:]}")
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["threads"] = threads
    return out


def cpu_baseline(scene):
    """Time the C restatement of the reference algorithm (oracle/flip_cpu.c,
    checked against the numpy oracle by tests/test_cpu_restatement.py) on this
    host's cores: one substep of the benchmark scene on every CPU this process
    may use, one substep of the 128^3 APIC dam break (C2) on all of them, and
    the 32^3 dam break (C1) for its full 100 substeps on all of them and on
    one thread.  Compiled for the host CPU in a temporary
    directory (the prebuilt oracle/_build/flip_cpu is the fallback)."""
    import tempfile
    from gpufluid_amd import seeding
    src = ROOT / "oracle" / "flip_cpu.c"
    topo = _cpu_topology()
    threads = topo["cpus_allowed"]
    with tempfile.TemporaryDirectory() as td:
        exe = Path(td) / "flip_cpu"
        r = subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-std=c11", str(src), "-o",
                            str(exe), "-lm"], capture_output=True, text=True)
        if r.returncode != 0:
            exe = ROOT / "oracle" / "_build" / "flip_cpu"
            if not exe.exists():
                log("cpu baseline: no compiler and no prebuilt binary")
                return None
        log(f"cpu baseline: {threads} threads on {topo}")
        main = _run_cpu(exe, td, scene, 1, threads)
        if main is None:
            return None
        c1 = seeding.config_c1()
        c1_all = _run_cpu(exe, td, c1, c1.params["substeps"], threads, max_iter=1000)
        c1_one = _run_cpu(exe, td, c1, c1.params["substeps"], 1, max_iter=1000)
        c2 = seeding.config_c2()
        c2_all = _run_cpu(exe, td, c2, 1, threads)
    main["cores"] = threads
    main["topology"] = topo
    main["c1_100_substeps"] = {"all_threads": c1_all, "one_thread": c1_one}
    main["c2_one_substep"] = c2_all
    return main


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiable-rollout leg (64^3, 40 substeps)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-scenes leg (256 scenes, 20 substeps)")
    ap.add_argument("--no-attributes", action="store_true",
                    help="skip the colour/temperature transfer leg (3 substeps on the scene)")
    ap.add_argument("--no-viscosity", action="store_true",
                    help="skip the viscosity leg (3 viscous substeps on the scene)")
    ap.add_argument("--no-surface-tension", action="store_true",
                    help="skip the surface-tension leg (3 substeps with CSF on the scene)")
    ap.add_argument("--no-reseed", action="store_true",
                    help="skip the GPU reseed leg (one reseed of the scene's particles)")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the small-grid graph-replay leg (64^3, direct vs hipGraph substeps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (3 launches) so that profiler "
                         "byte counters can be scaled against a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the substep path runs on a single device "
              "(--gpus 1); several GPUs are not supported.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from gpufluid_amd import FlipSolver3D, _lib

    t0 = time.time()
    scene = build_scene(args.config)
    p = scene.params
    dt = p.get("dt", 1.0 / 240)
    tol = p.get("tol", 1e-8)
    max_iter = p.get("max_iter", 500)
    log(f"scene {scene.name}: {scene.shape} grid, {scene.n_particles} particles "
        f"(seeded in {time.time() - t0:.1f}s)")
    s = FlipSolver3D(*scene.shape, dx=scene.dx, flip_blend=p.get("flip_blend", 0.95),
                     transfer_mode=p.get("transfer", "flip"), enable_timing=True,
                     pcg_precision=args.precision, pcg_check_every=args.check_every,
                     device="cuda:0")
    s.set_particles(scene.pos, scene.vel)
    s.pcg_kernel_timing = True

    def one_step():
        s.step(dt, pressure_iters=max_iter, pressure_solver="pcg", pressure_tol=tol)

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 256 MiB device-to-device copies")
        del src, dst
    for i in range(args.warmup):
        one_step()
        log(f"warmup {i}: PCG {s.last_pressure_iters} it, residual {s.last_pressure_residual:.2e}")
    # The timed steps run uninstrumented (no stage events, no PCG kernel
    # events); the stage breakdown comes from instrumented steps afterwards.
    s._prof.enabled = False
    s.pcg_kernel_timing = False
    stats = []
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    w0 = time.perf_counter()
    for i in range(args.steps):
        one_step()
        stats.append(dict(iters=s.last_pressure_iters, res=s.last_pressure_residual,
                          nrep=s.last_pressure_replacements, na=s.n_active_tiles,
                          nf=s._sc_fluid_cells))
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    total_ms = start.elapsed_time(end)
    ms_step = total_ms / args.steps
    npart = s.n_particles
    value = npart * args.steps / (total_ms / 1e3) / 1e6
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step (wall {wall * 1e3 / args.steps:.3f}), "
        f"{value:.1f} Mparticle-substeps/s")

    s._prof.enabled = True
    s.pcg_kernel_timing = True
    s._prof.reset()
    n_instr = max(1, min(args.steps, 3))
    istats = []
    for i in range(n_instr):
        one_step()
        istats.append(dict(iters=s.last_pressure_iters, nrep=s.last_pressure_replacements,
                           kms=s.last_pcg_kernel_ms))
    log(f"instrumented {n_instr} steps for the stage breakdown")
    sec = s._prof.summarize()
    nf = float(np.mean([st["nf"] for st in stats]))
    na = float(np.mean([st["na"] for st in stats])) * 512
    iters = [st["iters"] for st in stats]
    # iteration kernel (two-kernel form: direction+SpMV), update (two-kernel
    # form only): ms summed over the sampled iterations (one per PCG chunk);
    # number of sampled iterations; replacements: ms summed over all
    kms = np.sum([st["kms"] for st in istats], axis=0)
    n_it = sum(st["iters"] for st in istats)
    n_sampled = int(kms[2])
    n_rep = sum(st["nrep"] for st in istats)
    two_kernel = kms[1] > 0
    if two_kernel:
        pcg_kern = {"pcg_dirspmv": (kms[0], n_sampled), "pcg_update": (kms[1], n_sampled),
                    "pcg_replace": (kms[3], n_rep)}
    else:
        pcg_kern = {"pcg_iter": (kms[0], n_sampled), "pcg_replace_cg1": (kms[3], n_rep)}
    kern = {
        **pcg_kern,
        "sort": (sec.get("clear", 0.0), n_instr), "p2g": (sec.get("p2g", 0.0), n_instr),
        "g2p": (sec.get("g2p_advect", 0.0), n_instr),
    }
    stages = {}
    for k, (ms, calls) in kern.items():
        if calls <= 0 or ms <= 0:
            continue
        avg_ms = ms / calls
        b = BYTES[k](npart, nf, na)
        stages[k] = {"avg_ms": round(avg_ms, 5), "launches": int(calls),
                     "alg_bytes": int(b), "GBps": round(b / (avg_ms * 1e-3) / 1e9, 1),
                     "frac_hbm": round(b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # The PCG iteration as one stage: its kernels' sampled average durations
    # added, at the minimal fused iteration's 46 bytes per fluid cell.
    it_keys = [k for k in pcg_kern if k in ("pcg_dirspmv", "pcg_update", "pcg_iter")]
    rep_key = "pcg_replace" if two_kernel else "pcg_replace_cg1"
    it_ms = sum(stages[k]["avg_ms"] for k in it_keys if k in stages)
    if it_ms > 0:
        b = BYTES["pcg_iteration"](npart, nf, na)
        stages["pcg_iteration"] = {"avg_ms": round(it_ms, 5), "launches": n_sampled,
                                   "kernels": it_keys, "alg_bytes": int(b),
                                   "GBps": round(b / (it_ms * 1e-3) / 1e9, 1),
                                   "frac_hbm": round(b / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # whole PCG: iterations at their sampled average plus the replacements, in
    # the kernels' own compulsory bytes (34 per fluid cell per iteration for the
    # two-kernel form, fewer than the 46 of the stage figure)
    pcg_ms = it_ms * n_it + float(kms[3])
    own_bytes = (sum(BYTES[k](npart, nf, na) for k in it_keys) * n_it
                 + BYTES[rep_key](npart, nf, na) * n_rep)
    dom = "pcg_iteration"
    d = stages[dom]
    c3_mixed = args.precision == "mixed" and args.config == "c3"
    traffic = [measured_traffic(k) for k in it_keys] if c3_mixed else [None]
    roofline = {"bound": "hbm", "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": d["frac_hbm"],
                "traffic": sum(traffic) if None not in traffic else None,
                "traffic_source": TRAFFIC_FILE.name, "kernel": dom,
                "launch_kernels": [TRAFFIC_KERNEL[k][0] for k in it_keys],
                "alg_bytes_per_launch": d["alg_bytes"], "avg_launch_ms": d["avg_ms"],
                "timed_launches": n_sampled,
                "timing": "HIP events on the kernels' dispatch packets (hipExtLaunchKernel), "
                          "one iteration per PCG chunk sampled across the timed steps",
                "kernel_own_bytes": {k: {"alg_bytes": stages[k]["alg_bytes"],
                                         "avg_ms": stages[k]["avg_ms"],
                                         "frac_hbm": stages[k]["frac_hbm"],
                                         "traffic": measured_traffic(k) if c3_mixed else None}
                                     for k in it_keys if k in stages},
                "pcg_total_own_GBps": round(own_bytes / (pcg_ms * 1e-3) / 1e9, 1),
                "pcg_total_own_frac": round(own_bytes / (pcg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    # The GPU's own account of the timed window: HIP events on the stream
    # around the timed steps against the host wall clock, and the kernels'
    # execution per step (stage events of the instrumented steps, PCG from the
    # kernels' dispatch-packet events) as a fraction of the event window.
    kernel_ms = sum(sec.get(k, 0.0) for k in ("clear", "p2g", "g2p_advect")) / n_instr + \
        pcg_ms / n_instr
    gpu_window = {"event_ms_per_step": round(ms_step, 4),
                  "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
                  "kernel_ms_per_step": round(float(kernel_ms), 4),
                  "busy_frac": round(float(kernel_ms) / ms_step, 4)}
    log(f"gpu window: {gpu_window}")

    pcie = pcie_inclusive(s, npart, ms_step)
    log(f"pcie-inclusive: {pcie}")

    fp64 = None
    if not args.no_fp64_compare:
        fp64 = compare_fp64(s, tol, max_iter, args.check_every,
                            (s.last_pressure_iters, s.last_pressure_residual))
        log(f"pcg mixed vs fp64: {fp64}")

    attrs = None
    if not args.no_attributes:
        attrs = attribute_leg(s, one_step)
        log(f"attributes: {attrs}")

    visc = None
    if not args.no_viscosity:
        visc = viscosity_leg(s, one_step)
        log(f"viscosity: {visc}")

    rsd = None
    if not args.no_reseed:
        rsd = reseed_leg(s)
        log(f"reseed: {rsd}")

    csf = None
    if not args.no_surface_tension:
        csf = surface_tension_leg(s, one_step)
        log(f"surface tension: {csf}")

    adj = None
    if not args.no_adjoint:
        del s
        torch.cuda.empty_cache()
        adj = adjoint_leg()
        log(f"adjoint: {adj}")

    bat = None
    if not args.no_batch:
        torch.cuda.empty_cache()
        bat = batch_leg()
        log(f"batch: {bat}")

    gr = None
    if not args.no_graph:
        torch.cuda.empty_cache()
        gr = graph_leg()
        log(f"graph replay: {gr}")

    cpu = None
    if not args.no_cpu_baseline:
        c = cpu_baseline(scene)
        if c is not None:
            cpu = {"value": round(c["mps"], 4), "unit": "Mparticle-substeps/s",
                   "cores": c["cores"], "kind": "port", "sample": c.get("sample"),
                   "topology": c.get("topology"), "pcg_iters": c.get("iters"),
                   "seconds": c.get("seconds"), "c1_100_substeps": c.get("c1_100_substeps"),
                   "c2_one_substep": c.get("c2_one_substep"),
                   "label": "reference-algorithm CPU restatement (oracle/flip_cpu.c, FP32 "
                            "Jacobi-PCG checked every 10 iterations, dense grid)"}
            log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mparticle-substeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (f64 dots + residual replacement)",
        "data": "synthetic",
        "config": {"workload": scene.name, "grid": list(scene.shape), "particles": npart,
                   "transfer": p.get("transfer", "flip"), "dt": dt, "pcg_tol": tol,
                   "pcg_max_iter": max_iter, "pcg_precision": args.precision,
                   "pcg_iters": iters, "pcg_replacements": sum(st["nrep"] for st in stats),
                   "pcg_timed_solves": _spread([st["iters"] for st in stats],
                                               [st["res"] for st in stats], max_iter),
                   "fluid_cells": int(nf), "active_tiles": int(na // 512)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stages,
        "gpu_window": gpu_window,
        "pcg_vs_fp64": fp64,
        "pcie_inclusive_value": pcie,
        "attributes": attrs,
        "viscosity": visc,
        "surface_tension": csf,
        "reseed": rsd,
        "adjoint": adj,
        "batched": bat,
        "graph_replay": gr,
    }
    print(json.dumps(out), flush=True)


def attribute_leg(s, one_step, steps=3):
    """Per-particle colour (3 channels) and temperature (1 channel) on every
    particle of the scene: substep time with both transfers, and each
    transfer's own time (HIP events around its launches)."""
    import numpy as np
    import torch
    n = s.n_particles
    rng = np.random.default_rng(0)
    s.attr_color = rng.random((n, 3), dtype=np.float32)
    s.attr_temperature = (300.0 + 50.0 * rng.random(n)).astype(np.float32)
    one_step()
    s._prof.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    sec = s._prof.summarize()
    col, tmp = sec.get("color", 0.0) / steps, sec.get("temperature", 0.0) / steps
    s.attr_color = None
    s.attr_temperature = None
    return {"particles": n, "steps": steps, "ms_per_step": round(ms, 3),
            "color_ms": round(col, 4), "temperature_ms": round(tmp, 4),
            "color_Mparticles_per_s": round(n / (col * 1e-3) / 1e6, 1) if col else None,
            "temperature_Mparticles_per_s": round(n / (tmp * 1e-3) / 1e6, 1) if tmp else None,
            "timing": "wall clock around the substeps; transfers by HIP events (step profiler)"}


def viscosity_leg(s, one_step, steps=3, viscosity=1e-3, iters=12):
    """The scene with viscosity on (S2.13, `iters` Jacobi sweeps per component
    over the dense face arrays): substep time and the viscosity stage's own
    time (HIP events), plus its bytes (12 per face per sweep: u_in, u_old,
    u_out; 16 per face for the two copies of the faces besides)."""
    import torch
    s.viscosity, s.viscosity_iters = viscosity, iters
    one_step()
    s._prof.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    sec = s._prof.summarize()
    s.viscosity = 0.0
    vms = sec.get("viscosity", 0.0) / steps
    faces = sum(int(s._f[k].numel()) for k in ("u", "v", "w"))
    b = faces * (12 * iters + 16)
    return {"viscosity": viscosity, "viscosity_iters": iters, "steps": steps,
            "ms_per_step": round(ms, 3), "viscosity_ms": round(vms, 4),
            "divergence_ms": round(sec.get("divergence", 0.0) / steps, 4),
            "stage_bytes": b, "GBps": round(b / (vms * 1e-3) / 1e9, 1) if vms else None,
            "timing": "wall clock around the substeps; stages by HIP events (step profiler)"}


def surface_tension_leg(s, one_step, steps=3, sigma=0.07, passes=2):
    """The scene with surface tension on (S2.14: indicator, `passes` blurs,
    normal, curvature over the dense cells; impulse and force balance over
    the faces): substep time and the stage's own time (HIP events)."""
    import torch
    s.surface_tension, s.csf_smoothing_passes = sigma, passes
    one_step()
    s._prof.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    sec = s._prof.summarize()
    s.surface_tension = 0.0
    cms = sec.get("surface_tension", 0.0) / steps
    cells = s.nx * s.ny * s.nz
    faces = sum(int(s._f[k].numel()) for k in ("u", "v", "w"))
    # cells: indicator 8, blur 8 per pass, normal 16, curvature 16; faces:
    # apply 8 (+ chi/kappa reads, cached) and balance 8, then the solid BC
    b = cells * (8 + 8 * passes + 16 + 16) + faces * 16
    return {"surface_tension": sigma, "smoothing_passes": passes, "steps": steps,
            "ms_per_step": round(ms, 3), "surface_tension_ms": round(cms, 4),
            "stage_bytes": b, "GBps": round(b / (cms * 1e-3) / 1e9, 1) if cms else None,
            "timing": "wall clock around the substeps; stages by HIP events (step profiler)"}


def reseed_leg(s, repeats=3):
    """One GPU reseed (S2.11.GPU: count, rank-cull, compact, emit; default
    bounds 4..16 per cell) of the scene's particles after the timed steps,
    wall clock around the call (it synchronises once to size the output)."""
    import numpy as np
    import torch
    from gpufluid_amd.reseed import ReseedConfig, reseed_particles_gpu
    pos, vel, marker = s.pos, s.vel, s.marker
    cfg = ReseedConfig()
    out = reseed_particles_gpu(pos, vel, marker, s.dx, cfg, np.random.default_rng(0))
    times = []
    for r in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = reseed_particles_gpu(pos, vel, marker, s.dx, cfg, np.random.default_rng(r))
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(times))
    n = s.n_particles
    return {"particles": n, "emitted": out[4], "culled": out[5], "ms": round(ms, 3),
            "Mparticles_per_s": round(n / (ms * 1e-3) / 1e6, 1),
            "timing": f"wall clock, median of {repeats}"}


def adjoint_leg():
    """Differentiable rollout on the 64^3 box (x in [0.2,0.5], y < 0.5, 292,144
    particles, initial velocities N(0, 0.5^2) from seed 3): 40 FLIP substeps at
    dt = 1/240 with mixed PCG to 1e-8 and checkpoints every 10 substeps.  Loss =
    |centre of mass - (0.7,0.3,0.5)|^2.  Reports forward and reverse wall time
    (the reverse includes the segment recomputation).  The gradient itself is
    checked against FP64 references by tests/test_gpu_adjoint.py (16^3) and
    tests/test_gpu_adjoint_c4.py (this workload)."""
    import numpy as np
    import torch
    from gpufluid_amd import FlipSolver3D, seeding
    from gpufluid_amd.adjoint import FlipRollout
    sc = seeding.config_c4()
    n = sc.shape[0]
    p = sc.params
    target = torch.tensor(p["target"], dtype=torch.float64)
    s = FlipSolver3D(n, n, n, dx=sc.dx, flip_blend=0.95, device="cuda:0")
    roll = FlipRollout(s, p["dt"], p["substeps"], checkpoint_every=p["checkpoint_every"],
                       pressure_iters=500, pressure_tol=1e-8)
    p0 = torch.as_tensor(sc.pos, device="cuda:0")
    v0 = torch.as_tensor(sc.vel, device="cuda:0")
    roll.run(p0, v0)                      # warm-up (first launches, allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = roll.run(p0, v0)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    npart = x.shape[0]
    com = x.double().mean(0).cpu()
    gx = (2.0 * (com - target) / npart).float().to("cuda:0").expand(npart, 3).contiguous()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, gvel = roll.backward(gx)
    torch.cuda.synchronize()
    bwd = time.perf_counter() - t0
    return {"workload": "rollout_box_64 (292,144 particles, 40 FLIP substeps, checkpoint every 10)",
            "forward_ms": round(fwd * 1e3, 2), "reverse_ms": round(bwd * 1e3, 2),
            "adjoint_over_forward": round(bwd / fwd, 3),
            "forward_pcg_iters_mean": float(np.mean(roll.forward_iters)),
            "adjoint_pcg_iters_mean": float(np.mean(roll.adjoint_iters)),
            "grad_norm": float(gvel.double().norm()),
            "gradient_checks": "tests/test_gpu_adjoint.py (16^3, FP64 reference and central "
                               "differences); tests/test_gpu_adjoint_c4.py (this workload, FP64)"}


def batch_leg(substeps=20, alone_substeps=2, tol=1e-8, max_iter=500):
    """256 independent scenes (edges uniform in {24,32,48,64,96}, fluid boxes
    with sides U[0.2,0.7] from seed 4, 8 particles per cell, FLIP, dt 1/240,
    mixed PCG to 1e-8 per scene): packed into one grid, PCG iteration form
    and grid size chosen by the autotuner, ``substeps`` substeps replayed as captured hipGraph pairs
    with an autotuned PCG budget.  Compared with the same scenes run one at a
    time through FlipSolver3D (``alone_substeps`` timed substeps each)."""
    import ctypes as C
    import numpy as np
    import torch
    from gpufluid_amd import FlipSolver3D, _lib, autotune, seeding
    from gpufluid_amd.batch import BatchedScenes
    scenes = seeding.config_c5()
    data = [(sc.shape[0], sc.pos, sc.vel) for sc in scenes]
    npart = sum(sc.n_particles for sc in scenes)
    dt = scenes[0].params["dt"]
    b = BatchedScenes(data, dt=dt, device="cuda:0")
    for _ in range(2):
        b.step(pressure_iters=max_iter, pressure_tol=tol)   # warm-up, first stats
    s = b.s

    def solve():
        d = s._descriptor()
        _lib.call("bench", "gf_scene_pcg_sync", C.byref(d), 0, _lib.GF_PCG_MIXED, tol, max_iter,
                  s.pcg_check_every, -1, None, None, None, None, s._stream())

    (variant, wg), tune_times = autotune.tune_pcg(solve, key=("batch", b.pack.shape, b.n_scenes))
    budget = autotune.graph_budget(b.last_iters, s.pcg_check_every, max_iter)
    b.step2_graph(max_iter, tol, budget_iters=budget)      # capture + one untimed pair
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(substeps // 2):
        b.step2_graph(max_iter, tol, budget_iters=budget)
        iters.append(int(max(b.last_iters_pair[0].max(), b.last_iters_pair[1].max())))
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - t0
    worst_res = float(b.last_residual.max())
    # direct launches (host-synchronised PCG), same batch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(2):
        b.step(pressure_iters=max_iter, pressure_tol=tol)
    torch.cuda.synchronize()
    t_direct = (time.perf_counter() - t0) / 2
    del b, s
    torch.cuda.empty_cache()
    autotune.set_pcg_workgroups(0)
    autotune.set_pcg_variant(0)
    # one scene at a time
    t_alone = 0.0
    for sc in scenes:
        n = sc.shape[0]
        f = FlipSolver3D(n, n, n, device="cuda:0")
        f.set_particles(sc.pos, sc.vel)
        f.step(dt, pressure_iters=max_iter, pressure_solver="pcg", pressure_tol=tol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(alone_substeps):
            f.step(dt, pressure_iters=max_iter, pressure_solver="pcg", pressure_tol=tol)
        torch.cuda.synchronize()
        t_alone += time.perf_counter() - t0
        del f
    graph_v = npart * substeps / t_graph / 1e6
    alone_v = npart * alone_substeps / t_alone / 1e6
    return {"workload": f"{len(scenes)} scenes, edges 24-96, {npart} particles, FLIP, mixed PCG to 1e-8",
            "particles": npart, "substeps_timed": substeps,
            "batched_graph_value": round(graph_v, 2),
            "batched_direct_value": round(npart / t_direct / 1e6, 2),
            "one_at_a_time_value": round(alone_v, 2),
            "one_at_a_time_sample": f"{alone_substeps} substeps per scene after one warm-up substep",
            "unit": "Mparticle-substeps/s",
            "speedup_vs_one_at_a_time": round(graph_v / alone_v, 2),
            "pcg_workgroups_autotuned": wg,
            "pcg_variant_autotuned": {1: "two_kernel", 2: "one_kernel"}[variant],
            "autotune_seconds": {str(k): round(v, 5) for k, v in tune_times.items()},
            "graph_pcg_budget_iters": budget, "max_pcg_iters_per_substep": iters,
            "worst_scene_rel_residual": worst_res}


def graph_leg(n=64, steps=30, warmup=3):
    """Small-grid launch overhead: the reference's graph workload (64^3,
    seed_box (0.3,0.3,0.3)-(0.6,0.6,0.6) at 4 particles per cell, dt 0.005)
    stepped directly and replayed from captured hipGraphs, with 20 Jacobi
    sweeps and with PCG (tolerance 1e-4, cap 80).  Wall clock around `steps`
    substeps after `warmup` (the first graph substeps capture)."""
    import torch
    from gpufluid_amd import FlipSolver3D
    out = {"grid": [n, n, n], "steps": steps, "dt": 0.005,
           "timing": "wall clock around the substeps, synchronised once at each end"}
    for solver, iters in (("jacobi", 20), ("pcg", 80)):
        row = {}
        for mode in ("direct", "graph"):
            s = FlipSolver3D(n, n, n, dx=1.0 / n, device="cuda:0",
                             enable_cuda_graphs=mode == "graph")
            s.seed_box((0.30, 0.30, 0.30), (0.60, 0.60, 0.60), ppc=4)
            for _ in range(warmup):
                s.step(0.005, pressure_iters=iters, pressure_solver=solver)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                s.step(0.005, pressure_iters=iters, pressure_solver=solver)
            torch.cuda.synchronize()
            row[f"{mode}_ms"] = round((time.perf_counter() - t0) * 1e3 / steps, 4)
            row["particles"] = s.n_particles
            if mode == "graph":
                row["captures"] = s._cuda_graph_misses
                row["replays"] = s._cuda_graph_hits
            if solver == "pcg":
                row[f"{mode}_pcg_iters"] = s.last_pressure_iters
        row["speedup"] = round(row["direct_ms"] / row["graph_ms"], 3)
        out[solver] = row
    return out


def pcie_inclusive(s, npart, ms_step):
    """Rate a caller would see if it handed pos+vel over from host memory every
    substep: one pinned upload and one download of (N,3)+(N,3) FP32, timed
    with HIP events, added to the device-resident step time."""
    import torch
    dev = s.device
    host = torch.empty((2, npart, 3), dtype=torch.float32, pin_memory=True)
    buf = torch.empty((2, npart, 3), dtype=torch.float32, device=dev)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    c = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    buf.copy_(host, non_blocking=True)
    b.record()
    host.copy_(buf, non_blocking=True)
    c.record()
    torch.cuda.synchronize()
    h2d, d2h = a.elapsed_time(b), b.elapsed_time(c)
    del host, buf
    ms = ms_step + h2d + d2h
    return {"value": round(npart / (ms * 1e-3) / 1e6, 2), "unit": "Mparticle-substeps/s",
            "h2d_ms": round(h2d, 3), "d2h_ms": round(d2h, 3), "bytes_each_way": npart * 24}


def _solve_timed(s, prec, tol, max_iter, check_every):
    """One PCG solve of the system the solver's last substep built (same
    right-hand side, flags and tiles), timed with HIP events on its stream."""
    import torch
    from gpufluid_amd import _lib
    sc = s._sc
    d = s._descriptor()
    it, rr, nrep = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    stream = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    _lib.call("S2.6.3", "gf_scene_pcg_sync", ctypes.byref(d), sc.n_active_host, prec, tol,
              max_iter, check_every, -1, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(nrep),
              None, s._stream())
    b.record(stream)
    torch.cuda.synchronize()
    return {"ms": round(a.elapsed_time(b), 3), "iters": it.value, "rel_residual": rr.value,
            "replacements": nrep.value}


def compare_fp64(s, tol, max_iter, check_every, own):
    """The last substep's pressure system solved again:
    * with the mixed PCG under the substep's own tolerance and cap: the solver
      is deterministic, so this reproduces the substep's own solve (``own``:
      its iterations and residual) exactly — the check that the re-solve sees
      the same system;
    * with the mixed and the plain FP64 PCG under the same cap (equal
      iteration budget: the per-iteration cost ratio);
    * with both run to the same relative residual with no practical cap (time
      to an equal residual): 1e-6, and the substep's own tolerance."""
    from gpufluid_amd import _lib
    s._sc.ensure_fp64()
    res = {}
    mixed = _solve_timed(s, _lib.GF_PCG_MIXED, tol, max_iter, check_every)
    res["same_system_resolve"] = {"mixed": mixed, "substep_iters": own[0],
                                  "substep_rel_residual": own[1],
                                  "identical": (mixed["iters"], mixed["rel_residual"]) == own}
    fp64 = _solve_timed(s, _lib.GF_PCG_FP64, tol, max_iter, check_every)
    res["equal_iteration_cap"] = {"cap": max_iter, "mixed": mixed, "fp64": fp64,
                                  "speedup": round(fp64["ms"] / mixed["ms"], 3)}
    eq = {}
    for t in sorted({1e-6, tol}, reverse=True):
        m = _solve_timed(s, _lib.GF_PCG_MIXED, t, 20000, check_every)
        f = _solve_timed(s, _lib.GF_PCG_FP64, t, 20000, check_every)
        eq[f"{t:g}"] = {"mixed": m, "fp64": f, "speedup": round(f["ms"] / m["ms"], 3)}
    res["equal_residual"] = eq
    return res


if __name__ == "__main__":
    main()
