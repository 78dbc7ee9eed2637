#!/usr/bin/env python3
"""Benchmark: MiniPIC electrostatic timestep on one MI355X.

Workload (the 1-GPU configuration): unit cube, 64^3 hexes x 6 Kuhn tets =
1,572,864 tets; 64 electrons per tet = 100,663,296 particles, Maxwellian
vth = 0.1 (temp = 2 vth^2), weights modulated by 1 + 0.05 sin(2 pi x),
uniform neutralising background, phi = 0 on the boundary, CG tol 1e-10,
dt = 0.05.  Synthetic inputs generated on the device (seeded).

One step = PIC::time_step: charge deposit, CG solve, E gather, Boris kick,
tet-walk relocation, KE, counting sort amortised over --sort-period steps
(default 2: the reorder only restores memory locality; the physics does not
depend on it, and the reference's OpenMP path never reorders).  The metric is
particle-steps per second over exactly --steps timed steps after --warmup.
Both forms run the same workload from the same fill: the precision form
(FP32 push/gather/walk, FP64 point location near faces, 64-bit fixed-point
deposit, FP32 CG inside FP64 refinement) is the headline value, the FP64 form
(bit-identical to the reference's arithmetic) is reported beside it with the
ratio, and the precision form's errors against the FP64 form are reported.

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
# Algorithmic bytes per particle of the fused push: reads x, x_ref, v, elem,
# weight, species; writes x, x_ref, v, elem.
PUSH_BYTES_FP64 = 3 * 24 + 4 + 8 + 1 + 3 * 24 + 4        # 161 B
PUSH_BYTES_MIXED = 3 * 12 + 4 + 8 + 1 + 3 * 12 + 4        # 89 B (FP32 state, FP64 weight)
# Infinity Cache (MALL) of the MI355X: a matrix below this size can stay resident
MALL_BYTES = 256 << 20


def cg_bytes_per_iter(n, nnz, word):
    """Algorithmic bytes of one CG iteration: the CSR SpMV (values of `word`
    bytes, int32 columns and row offsets, the input vector read once and the
    output written once) plus five vector passes (p update, p.q, x update, r
    update, r.r) of `word` bytes per row.  FP64 CG: word 8 (the SpMV part is
    12 nnz + 4 (n+1) + 16 n); FP32 inner CG: word 4.  nnz counts the matrix's
    nonzero entries (and the diagonal), the entries the device reads: the
    stored exact zeros of the CSR are left out of its ELL."""
    spmv = (word + 4) * nnz + 4 * (n + 1) + 2 * word * n
    return spmv, spmv + 5 * word * n


def mg_levels_sizes(cells):
    """Unknowns of the coarse levels of the multigrid solve on a cubic brick of
    `cells` cells per direction: c cells coarsen to (c + 1) // 2 while c >= 3;
    the level with at most 512 unknowns is the coarsest (dense inverse)."""
    sizes = []
    c = cells
    while c >= 3:
        c = (c + 1) // 2
        sizes.append((c - 1) ** 3)
        if sizes[-1] <= 512:
            break
    return sizes


def mg_bytes_per_iter(cells, n, nnz):
    """Algorithmic bytes of one multigrid-preconditioned CG iteration of the
    precision form (k_mg.hip), split as (FP64 CG part, V-cycle part):
      q = A p: FP64 values and int32 columns (12 nnz), z (4 n) and the old p
        (8 n) read, p and q written (16 n); the update: x, p, r, q read (32 n),
        x, r written (16 n), the V-cycle's b0 and x0 written (8 n);
      level 0 of the V-cycle: two FP32 SpMVs (8 nnz each) with b0, x0, t0, dinv,
        r and z read or written once (32 n), restriction and prolongation
        (t0, x0 read, t0 written: 12 n);
      coarse levels of n_l unknowns: the 27-point stencil's FP32 values read
        twice (2 x 108 n_l; the columns are the tile's halo in LDS) and the
        level's vectors (restriction input, b, x, t, D^-1 and the coarser z:
        40 n_l); the coarsest level's dense inverse (4 n_c^2)."""
    cg = 12 * nnz + 28 * n + 56 * n
    sizes = mg_levels_sizes(cells)
    vc = 16 * nnz + 44 * n
    for nl in sizes[:-1]:
        vc += 2 * 108 * nl + 40 * nl
    if sizes:
        vc += 4 * sizes[-1] ** 2 + 16 * sizes[-1]
    return cg, vc


def mg_roofline(sm, cells, n, nnz, levels):
    cg, vc = mg_bytes_per_iter(cells, n, nnz)
    per_iter = cg + vc
    it, ms = sm["cg_iters_mean"], sm["cg_ms_per_step"]
    achieved = per_iter * it / (ms * 1e-3) / 1e9 if ms > 0 and it > 0 else 0.0
    return {"kernel": "multigrid-preconditioned CG: k_mg_ap/k_mg_update (FP64) + the FP32 V-cycle (%d levels)" % levels,
            "bound": "hbm", "bytes_per_iter": per_iter, "cg_bytes_per_iter": cg, "vcycle_bytes_per_iter": vc,
            "iters_per_step": it, "ms_per_step": ms, "us_per_iter": ms * 1e3 / it if it else None,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "time_covers": "the whole solve of a step (the first V-cycle, the iterations and the final check of the "
                           "true FP64 residual), host waits included"}


def cg_roofline(sm, n, nnz, word, kernel):
    spmv, per_iter = cg_bytes_per_iter(n, nnz, word)
    it, ms = sm["cg_iters_mean"], sm["cg_ms_per_step"]
    achieved = per_iter * it / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"kernel": kernel, "bound": "hbm", "bytes_per_iter": per_iter, "spmv_bytes_per_iter": spmv,
            "iters_per_step": it, "ms_per_step": ms, "us_per_iter": ms * 1e3 / it if it else None,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "matrix_bytes": (word + 4) * nnz,
            "matrix_fits_infinity_cache": (word + 4) * nnz < MALL_BYTES,
            "time_covers": "the whole solve of a step (FP64 form: CG; precision form: FP32 inner CG and the "
                           "FP64 residuals and updates of the refinement), host waits included"}


def measured_traffic(kernel):
    """HBM bytes per push launch measured by the rocprofv3 FETCH_SIZE/WRITE_SIZE
    runs recorded in profiles/ (None when no such summary exists)."""
    path = os.path.join(ROOT, "profiles", "%s_c2_traffic.md" % kernel)
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
    """hipDeviceSynchronize (the library's calls are stream-ordered on its own stream)."""
    import ctypes as C
    assert C.CDLL("libamdhip64.so").hipDeviceSynchronize() == 0


def run_gpu(args, form):
    """Fill, one solve on the fill (same particles in both forms), then warmup
    and exactly args.steps timed time steps."""
    import numpy as np
    mp = load("pyminipic", os.path.join(PKG, "pyminipic.py"))
    name = "mixed" if form == mp.FORM_MIXED else "fp64"
    n = args.nx
    t0 = time.perf_counter()
    mesh = mp.Mesh(n, n, n, 6, background=DENSITY)
    log("[%s] mesh %d^3 x6: %d tets, %d nodes, nnz %d (%.1f s)" % (name, n, mesh.ne, mesh.nn, mesh.nnz,
                                                                  time.perf_counter() - t0))
    cfg = mp.default_config(sort_period=args.sort_period, cg_tol=1e-10, form=form)
    ctx = mp.Context(mesh.desc, cfg)
    ctx.reserve(mesh.ne * args.nppe)
    ctx.fill(args.nppe, 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=1, pert_amp=0.05)
    npart = ctx.num_particles()
    ctx.weight_charge()
    ctx.solve(1e-10, 1000)
    log("[%s] filled %d particles" % (name, npart))
    for i in range(args.warmup):
        st = ctx.time_step(args.dt)
        log("[%s] warmup %d: cg %d iters, push %.2f ms, solve %.2f ms" % (name, i, st["cg_iters"], st["ms_push"],
                                                                       st["ms_solve"]))
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
        log("[%s] step %d: cg %d iters %.2f ms, deposit %.2f, push %.2f, sort %.2f, KE %.9e" % (
            name, i, st["cg_iters"], st["ms_solve"], st["ms_deposit"], st["ms_push"], st["ms_sort"],
            st["kinetic_energy"]))
    mg_levels = ctx.multigrid_levels()[0] if form == mp.FORM_MIXED else 0
    out = dict(npart=npart, wall=wall, stats=stats, push_ms=push_ms, ne=mesh.ne, nn=mesh.nn, nnz=mesh.nonzeros(),
               nnz_stored=mesh.nnz,
               owner_mismatch=ctx.check_ownership(), mg_levels=mg_levels, cells=n)
    ctx.close()
    return out


def summary(g, K, bytes_per_particle):
    push_ms = sum(g["push_ms"]) / len(g["push_ms"])
    achieved = g["npart"] * bytes_per_particle / (push_ms * 1e-3) / 1e9
    st = g["stats"]
    return dict(value=g["npart"] * K / g["wall"], ms_per_step=g["wall"] / K * 1e3, push_ms=push_ms,
                push_gbs=achieved, cg_iters_mean=sum(s["cg_iters"] for s in st) / K,
                cg_ms_per_step=sum(s["ms_solve"] for s in st) / K,
                sort_ms_per_step=sum(s["ms_sort"] for s in st) / K,
                sort_ms_per_reorder=(sum(s["ms_sort"] for s in st) / max(1, sum(1 for s in st if s["ms_sort"] > 0))),
                field_ms_per_step=sum(s["ms_field"] for s in st) / K,
                deposit_ms_per_step=sum(s["ms_deposit"] for s in st) / K)


def run_error_curves(args):
    """The precision form's error against the FP64 form step by step, with a
    control: the FP64 form started from the same fill rounded to FP32 and back
    (position and velocity; the reference position recomputed in FP64 from the
    rounded position).  When the control drifts from the FP64 run as fast as
    the precision form does, the drift is the divergence of the trajectories
    (particles put on the other side of a face by a rounding), not a bias of
    the precision form.  Untimed; --err-steps steps from the fill solve."""
    import numpy as np
    mp = load("pyminipic", os.path.join(PKG, "pyminipic.py"))
    n = args.nx
    mesh = mp.Mesh(n, n, n, 6, background=DENSITY)
    runs = {}
    for name, form, roundtrip in (("fp64", mp.FORM_FP64, False), ("mixed", mp.FORM_MIXED, False),
                                  ("fp64_fp32_roundtrip", mp.FORM_FP64, True)):
        ctx = mp.Context(mesh.desc, mp.default_config(sort_period=args.sort_period, cg_tol=1e-10, form=form))
        ctx.reserve(mesh.ne * args.nppe)
        ctx.fill(args.nppe, 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=1, pert_amp=0.05)
        if roundtrip:
            ctx.set_form(mp.FORM_MIXED)
            ctx.set_form(mp.FORM_FP64)
        ctx.weight_charge()
        ctx.solve(1e-10, 1000)
        phis, ke = [ctx.phi()], []
        for _ in range(args.err_steps):
            st = ctx.time_step(args.dt)
            phis.append(ctx.phi())
            ke.append(st["kinetic_energy"])
        runs[name] = dict(phi=phis, ke=ke, rhs=ctx.rhs(), owner=ctx.check_ownership())
        ctx.close()
        log("[errors] %s: %d steps" % (name, args.err_steps))
    ref = runs["fp64"]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    curve = lambda r: [float("%.3g" % rel(a, b)) for a, b in zip(r["phi"], ref["phi"])]
    fe = lambda r: float(np.dot(r["phi"][-1], r["rhs"]))
    m, c = runs["mixed"], runs["fp64_fp32_roundtrip"]
    return {
        "phi_l2_rel_same_particles": rel(m["phi"][0], ref["phi"][0]),
        "phi_l2_rel_after_steps": rel(m["phi"][-1], ref["phi"][-1]),
        "field_energy_rel_after_steps": abs(fe(m) / fe(ref) - 1.0),
        "kinetic_energy_rel_after_steps": abs(m["ke"][-1] - ref["ke"][-1]) / ref["ke"][-1],
        "owner_mismatch_fp64_location": m["owner"],
        "steps_compared": args.err_steps,
        "phi_l2_rel_per_step": curve(m),
        "control": {
            "what": "FP64 form from the same fill rounded to FP32 and back (x, v), against the FP64 form",
            "phi_l2_rel_after_steps": rel(c["phi"][-1], ref["phi"][-1]),
            "field_energy_rel_after_steps": abs(fe(c) / fe(ref) - 1.0),
            "kinetic_energy_rel_after_steps": abs(c["ke"][-1] - ref["ke"][-1]) / ref["ke"][-1],
            "phi_l2_rel_per_step": curve(c),
        },
        "reference": "FP64 form of this build (bit-identical to the FP64 oracle)",
    }


def run_ensemble(args):
    """The batching form on the ensemble configuration: 256 problems, Kuhn sides
    6-24, 8-64 electrons per tet, vth 0.05-0.2 (seeded), one set of launches per
    step; tuned, replayed as hipGraphs; particle-steps/s over --ens-steps steps."""
    import numpy as np
    mp = load("pyminipic", os.path.join(PKG, "pyminipic.py"))
    rng = np.random.default_rng(2024)
    sides = rng.integers(6, 25, size=256)
    nppe = rng.integers(8, 65, size=256)
    vth = rng.uniform(0.05, 0.2, size=256)
    meshes = [mp.Mesh(int(n), int(n), int(n), 6) for n in sides]
    b = mp.Batch(meshes, mp.default_config(cg_tol=1e-10))
    b.fill([dict(nppe=int(nppe[k]), temp=2 * float(vth[k]) ** 2, density=DENSITY, seed_offset=k + 1)
            for k in range(256)])
    npart = int((sides.astype(np.int64) ** 3 * 6 * nppe).sum())
    b.tune(args.dt)
    b.graph(True)
    for _ in range(3):
        st = b.step(args.dt)
    gpu_sync()
    t0 = time.perf_counter()
    for _ in range(args.ens_steps):
        st = b.step(args.dt)
    gpu_sync()
    wall = time.perf_counter() - t0
    ps = b.problem_stats()
    out = dict(value=npart * args.ens_steps / wall, unit="particle-steps/s", ms_per_step=wall / args.ens_steps * 1e3,
               problems=256, particles=npart, tets=int((sides.astype(np.int64) ** 3 * 6).sum()),
               steps=args.ens_steps, push_grid=st["push_grid"], solve_threads=st["solve_threads"],
               max_inner_iters=st["max_inner_iters"], all_converged=bool(st["all_converged"]),
               max_relres=float(np.max(ps["relres"])), ownership_mismatches=b.check_ownership())
    b.close()
    log("[ensemble] %d problems, %d particles: %.2f ms/step, %.3e particle-steps/s" % (
        256, npart, out["ms_per_step"], out["value"]))
    return out


def copy_calibration(nbytes=1 << 31):
    """Two device-to-device copies of nbytes (the runtime's blit kernel): the
    FETCH_SIZE / WRITE_SIZE counters read for a copy of known size calibrate the
    push kernels' HBM traffic."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: 2 device-to-device copies of %d bytes done" % nbytes)


def run_cpu_baseline(args):
    """The FP64 oracle (restatement of the reference's OpenMP path, host solve
    single-threaded as in the reference) on --cpu-steps steps of the same
    workload (3 by default: about 10 s)."""
    O = load("minipic_oracle", os.path.join(ORACLE, "minipic_oracle.py"))
    import numpy as np
    threads = O.lib().orc_get_threads()
    # C1 timed in full: Kuhn 8^3 x6, 16 electrons per tet, dt 0.1, 10 steps
    # from the fill (cold potential, as the reference's run starts)
    m1 = O.Mesh(8, 8, 8, 6)
    p1 = O.Particles(m1.ne * 16).fill(m1, 16, 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=1)
    bg1 = m1.background(DENSITY)
    phi1 = np.zeros(m1.nn)
    S1 = O.charge_scale(m1, p1, bg1)
    t0 = time.perf_counter()
    for _ in range(10):
        O.step(m1, p1, phi1, 0.1, bg=bg1, S=S1)
    t1 = time.perf_counter() - t0
    c1 = dict(value=p1.n * 10 / t1, unit="particle-steps/s", cores=threads, kind="port",
              sample="the whole Kuhn 8^3 x6, 16/tet configuration (%d particles, dt 0.1, 10 steps from the fill), "
                     "%.2f s" % (p1.n, t1))
    # the GPU workload: one untimed solve after the fill (the GPU run's warm
    # start), then --cpu-steps timed steps
    n = args.nx
    m = O.Mesh(n, n, n, 6)
    p = O.Particles(m.ne * args.nppe).fill(m, args.nppe, 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=1, pert_amp=0.05)
    bg = m.background(DENSITY)
    phi = np.zeros(m.nn)
    S = O.charge_scale(m, p, bg)
    tw = time.perf_counter()
    O.cg(m, O.deposit(m, p, bg, S), phi, 1e-10, 1000)
    tw = time.perf_counter() - tw
    t0 = time.perf_counter()
    steps = args.cpu_steps
    for _ in range(steps):
        O.step(m, p, phi, args.dt, bg=bg, S=S)
    dt = time.perf_counter() - t0
    return dict(value=p.n * steps / dt, unit="particle-steps/s", cores=threads, kind="port",
                sample="%d step(s) of the full %d^3 x6, %d/tet workload (%d particles) warm-started like the GPU run "
                       "(an untimed %.1f s solve after the fill), %.1f s; particle phases on %d OpenMP threads, CG "
                       "single-threaded as the reference's host solve" % (steps, n, args.nppe, p.n, tw, dt, threads),
                c1_full=c1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--nppe", type=int, default=64)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--sort-period", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--err-steps", type=int, default=25, help="untimed steps of the error study (0 = skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ens-steps", type=int, default=10, help="timed steps of the ensemble (batching form); 0 = skip")
    ap.add_argument("--ensemble-only", action="store_true", help="run only the ensemble leg and print its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run two 2 GiB device-to-device copies whose FETCH_SIZE/WRITE_SIZE calibrate HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)
    if args.ensemble_only:
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    import numpy as np
    gm = run_gpu(args, 1)
    gd = run_gpu(args, 0)
    if args.copy_calibration:
        copy_calibration()
    K = args.steps
    sm, sd = summary(gm, K, PUSH_BYTES_MIXED), summary(gd, K, PUSH_BYTES_FP64)
    errors = run_error_curves(args) if args.err_steps > 0 else {}
    errors["owner_mismatch_fp64_location_timed_run"] = gm["owner_mismatch"]
    errors["relocated_last_step"] = gm["stats"][-1]["relocated"]
    ens = run_ensemble(args) if args.ens_steps > 0 else None
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = run_cpu_baseline(args)
        log("cpu baseline: %.3e particle-steps/s" % cpu["value"])
    out = {
        "metric": "particle-steps/sec incl. deposit+CG at 100M particles, mixed vs FP64 on 1 GPU",
        "value": sm["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": sm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (precision form: f32 push/gather/walk, f64 CG preconditioned by an f32 multigrid V-cycle, "
                 "f64 point location, i64 deposit)",
        "data": "synthetic (seeded device fill)",
        "config": {
            "workload": "Kuhn %d^3 x6 tets (%d), %d electrons/tet (%d particles), vth 0.1, density perturbation "
                        "0.05, CG tol 1e-10, dt %g, reorder every %d steps" % (
                            args.nx, gm["ne"], args.nppe, gm["npart"], args.dt, args.sort_period),
            "form": "mixed", "tets": gm["ne"], "nodes": gm["nn"], "nnz": gm["nnz"], "nnz_stored": gm["nnz_stored"], "particles": gm["npart"],
            "cg_iters_mean": sm["cg_iters_mean"], "cg_ms_per_step": sm["cg_ms_per_step"],
            "push_ms": sm["push_ms"], "sort_ms_per_step": sm["sort_ms_per_step"],
            "sort_ms_per_reorder": sm["sort_ms_per_reorder"], "field_ms_per_step": sm["field_ms_per_step"],
            "timed_steps_of_config": "%d timed steps (+%d warmup) of the configuration's 100; a rate" % (
                K, args.warmup),
        },
        "fp64": {"value": sd["value"], "ms_per_step": sd["ms_per_step"], "cg_iters_mean": sd["cg_iters_mean"],
                 "cg_ms_per_step": sd["cg_ms_per_step"], "push_ms": sd["push_ms"],
                 "sort_ms_per_step": sd["sort_ms_per_step"], "sort_ms_per_reorder": sd["sort_ms_per_reorder"],
                 "field_ms_per_step": sd["field_ms_per_step"],
                 "cg_roofline": cg_roofline(sd, gd["nn"], gd["nnz"], 8, "k_cg_k1 + k_cg_k2"),
                 "push_roofline": {"kernel": "k_push_fused", "achieved": sd["push_gbs"], "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": sd["push_gbs"] / HBM_PEAK_GBS,
                                   "bytes_per_particle": PUSH_BYTES_FP64,
                                   "traffic": measured_traffic("k_push_fused")}},
        "cg_roofline": (mg_roofline(sm, gm["cells"], gm["nn"], gm["nnz"], gm["mg_levels"]) if gm["mg_levels"] else
                        cg_roofline(sm, gm["nn"], gm["nnz"], 4, "k_ir_cg (FP32 inner) + k_ir_resid/k_ir_start/"
                                                               "k_ir_update (FP64 refinement)")),
        "mixed_over_fp64": sm["value"] / sd["value"],
        "errors_vs_fp64": errors,
        "roofline": {"bound": "hbm", "kernel": "k_push32", "achieved": sm["push_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": sm["push_gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic("k_push32"),
                     "bytes_per_particle": PUSH_BYTES_MIXED},
        "cpu_baseline": cpu,
        "ensemble": ens,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
