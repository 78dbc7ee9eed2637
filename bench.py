#!/usr/bin/env python3
"""Benchmark: MiniPIC electrostatic timestep on one MI355X.

Workload (the 1-GPU configuration): unit cube, 64^3 hexes x 6 Kuhn tets =
1,572,864 tets; 64 electrons per tet = 100,663,296 particles, Maxwellian
vth = 0.1 (temp = 2 vth^2), weights modulated by 1 + 0.05 sin(2 pi x),
uniform neutralising background, phi = 0 on the boundary, CG tol 1e-10,
dt = 0.05.  Synthetic inputs generated on the device (seeded).

One step = PIC::time_step: charge deposit, CG solve, E gather, Boris kick,
tet-walk relocation, KE, amortised counting sort.  The metric is
particle-steps per second over exactly --steps timed steps after --warmup.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "minipic-crossroads_amd")
ORACLE = os.path.join(ROOT, "oracle")

HBM_PEAK_GBS = 8000.0
DENSITY = 2.0 / 3.0 * math.pi * math.pi  # main.cpp:68
# Algorithmic bytes per particle of the fused FP64 push: reads x, x_ref, v
# (3 x 24 B), elem (4 B), weight (8 B), species (1 B); writes x, x_ref, v, elem.
PUSH_BYTES_FP64 = 3 * 24 + 4 + 8 + 1 + 3 * 24 + 4


def measured_traffic():
    """HBM bytes per push launch measured by the rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs recorded in profiles/ (None when no such summary exists)."""
    path = os.path.join(ROOT, "profiles", "k_push_fused_fp64_c2_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def gpu_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def run_gpu(args):
    mp = load("pyminipic", os.path.join(PKG, "pyminipic.py"))
    n = args.nx
    t0 = time.perf_counter()
    mesh = mp.Mesh(n, n, n, 6, background=DENSITY)
    log("mesh %d^3 x6: %d tets, %d nodes, nnz %d (%.1f s)" % (n, mesh.ne, mesh.nn, mesh.nnz, time.perf_counter() - t0))
    cfg = mp.default_config(sort_period=args.sort_period, cg_tol=1e-10)
    ctx = mp.Context(mesh.desc, cfg)
    ctx.reserve(mesh.ne * args.nppe)
    ctx.fill(args.nppe, 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=1, pert_amp=0.05)
    npart = ctx.num_particles()
    log("filled %d particles" % npart)
    for i in range(args.warmup):
        st = ctx.time_step(args.dt)
        log("warmup %d: cg %d iters, push %.2f ms, solve %.2f ms" % (i, st["cg_iters"], st["ms_push"], st["ms_solve"]))
    gpu_sync()
    stats = []
    push_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = ctx.time_step(args.dt)
        stats.append(st)
        push_ms.append(ctx.kernel_time(0))
    gpu_sync()
    wall = time.perf_counter() - t0
    for i, st in enumerate(stats):
        log("step %d: cg %d iters %.2f ms, deposit %.2f, push %.2f, sort %.2f, KE %.6e" % (
            i, st["cg_iters"], st["ms_solve"], st["ms_deposit"], st["ms_push"], st["ms_sort"], st["kinetic_energy"]))
    ctx.close()
    return dict(npart=npart, wall=wall, stats=stats, push_ms=push_ms, ne=mesh.ne, nn=mesh.nn, nnz=mesh.nnz)


def run_cpu_baseline(args):
    """The FP64 oracle (restatement of the reference's OpenMP path, host solve
    single-threaded as in the reference) on one step of the same workload."""
    O = load("minipic_oracle", os.path.join(ORACLE, "minipic_oracle.py"))
    import numpy as np
    threads = O.lib().orc_get_threads()
    n = args.nx
    m = O.Mesh(n, n, n, 6)
    p = O.Particles(m.ne * args.nppe).fill(m, args.nppe, 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=1, pert_amp=0.05)
    bg = m.background(DENSITY)
    phi = np.zeros(m.nn)
    S = O.charge_scale(m, p, bg)
    t0 = time.perf_counter()
    steps = args.cpu_steps
    for _ in range(steps):
        O.step(m, p, phi, args.dt, bg=bg, S=S)
    dt = time.perf_counter() - t0
    return dict(value=p.n * steps / dt, unit="particle-steps/s", cores=threads, kind="port",
                sample="%d step(s) of the full %d^3 x6, %d/tet workload (%d particles) from a cold potential, "
                       "%.1f s; particle phases on %d OpenMP threads, CG single-threaded as the reference's host solve"
                       % (steps, n, args.nppe, p.n, dt, threads))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--nppe", type=int, default=64)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--sort-period", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 2 GiB-in/2 GiB-out elementwise kernel (torch) whose FETCH_SIZE/WRITE_SIZE calibrate HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding).",
              file=sys.stderr)
        sys.exit(2)
    g = run_gpu(args)
    if args.copy_calibration:
        import torch
        a = torch.empty(2 ** 28, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(2):
            torch.add(a, 1.0, out=b)  # an elementwise kernel: reads 2 GiB, writes 2 GiB
        torch.cuda.synchronize()
        log("copy calibration: 2 elementwise kernels of 2 GiB in, 2 GiB out done")
    npart, K = g["npart"], args.steps
    value = npart * K / g["wall"]
    push_ms = sum(g["push_ms"]) / len(g["push_ms"])
    achieved = npart * PUSH_BYTES_FP64 / (push_ms * 1e-3) / 1e9
    cg_iters = [s["cg_iters"] for s in g["stats"]]
    cg_ms = [s["ms_solve"] for s in g["stats"]]
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = run_cpu_baseline(args)
        log("cpu baseline: %.3e particle-steps/s" % cpu["value"])
    out = {
        "metric": "particle-steps/sec incl. deposit+CG at 100M particles, mixed vs FP64 on 1 GPU",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": g["wall"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded device fill)",
        "config": {
            "workload": "Kuhn %d^3 x6 tets (%d), %d electrons/tet (%d particles), FP64 form, CG tol 1e-10, dt %g, "
                        "sort every %d steps" % (args.nx, g["ne"], args.nppe, npart, args.dt, args.sort_period),
            "form": "fp64", "tets": g["ne"], "nodes": g["nn"], "nnz": g["nnz"], "particles": npart,
            "cg_iters_mean": sum(cg_iters) / K, "cg_ms_per_step": sum(cg_ms) / K,
            "push_ms": push_ms,
        },
        "roofline": {"bound": "hbm", "kernel": "k_push_fused", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "bytes_per_particle": PUSH_BYTES_FP64},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
