"""The batching form: an ensemble of independent problems in one set of launches.

Each problem of an ensemble is checked against the same problem run alone in
the precision form (identical fills, right-hand sides from the same charges),
against the FP64 oracle (charge deposit, solve residual), and for the
invariants the configuration fixes (particle counts, ownership, convergence).
Graph replays are compared with eager steps after every replay.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY = 2.0 / 3.0 * math.pi * math.pi


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


SIDES = [3, 5, 4, 6, 2]
NPPE = [4, 10, 7, 12, 16]
VTH = [0.05, 0.2, 0.1, 0.15, 0.08]
# every side >= 3: each problem solves by CG preconditioned by a multigrid
# V-cycle on its own hierarchy (odd and even cell counts); SIDES has a
# 2-cell box, which does not coarsen, so its ensemble runs the refinement solve
MG_SIDES = [3, 5, 4, 6, 7]
SIDE_SETS = pytest.mark.parametrize("sides,mg", [(SIDES, 0), (MG_SIDES, 1)], ids=["refinement", "multigrid"])


def fills(drift=0.0):
    return [dict(nppe=NPPE[k], temp=2 * VTH[k] ** 2, density=DENSITY, pert_amp=0.05, drift=drift, seed_offset=k + 1)
            for k in range(len(SIDES))]


def make_batch(mp, sides=SIDES, **cfg):
    meshes = [mp.Mesh(n, n, n, 6) for n in sides]
    b = mp.Batch(meshes, mp.default_config(**cfg))
    b.fill(fills())
    return b


def by_id(g):
    o = np.argsort(g["id"], kind="stable")
    return {k: (v[:, o] if v.ndim == 2 else v[o]) for k, v in g.items()}


@SIDE_SETS
def test_batch_fill_and_charge_match_single_problems(mp, oracle, sides, mg):
    say("batch fill and charge vs single problems and the oracle, sides %s" % sides)
    b = make_batch(mp, sides)
    assert b.step(0.05)["multigrid"] == mg
    st = b.problem_stats()
    for k, n in enumerate(sides):
        pm = mp.Mesh(n, n, n, 6, background=DENSITY)
        ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED))
        ctx.reserve(pm.ne * NPPE[k])
        ctx.fill(NPPE[k], 0, DENSITY, temp=2 * VTH[k] ** 2, seed_offset=k + 1, pert_amp=0.05)
        alone = ctx.download()
        # rhs of the first batch step = charge of the fill (+ background), as the oracle deposits it
        om = oracle.Mesh(n, n, n, 6)
        p = oracle.Particles(om.ne * NPPE[k]).fill(om, NPPE[k], 0, DENSITY, temp=2 * VTH[k] ** 2, seed_offset=k + 1,
                                                   pert_amp=0.05)
        bg = om.background(DENSITY)
        ref_rhs = oracle.deposit(om, p, bg)
        rhs = b.rhs(k)
        # the net charge is a small difference of the electrons' and the background's:
        # errors are measured against the background's magnitude
        assert np.allclose(rhs, ref_rhs, rtol=0, atol=1e-12 * np.abs(bg).max()), k
        # the solve met the reference's stopping rule |r| < tol |b| on that rhs
        phi = b.phi(k)
        res = np.linalg.norm(om.spmv(phi) - rhs) / np.linalg.norm(rhs)
        assert res < 1e-10 and st["converged"][k] == 1, (k, res)
        assert st["count"][k] == p.n == alone["x"].shape[1]
        # the fill's weights: the single-problem fill's; the modulation's sin on the
        # device may differ from the host's by an ulp
        assert np.allclose(np.sort(alone["w"]), np.sort(p.w[:p.n]), rtol=4e-16, atol=0)
        ctx.close()
    assert b.check_ownership() == 0
    b.close()


@SIDE_SETS
def test_batch_steps_match_single_problems(mp, sides, mg):
    say("batch steps vs single problems (precision form), sides %s" % sides)
    b = make_batch(mp, sides)
    for _ in range(3):
        sb = b.step(0.05)
        assert sb["multigrid"] == mg and sb["all_converged"] == 1
        if mg:
            assert 0 < sb["max_inner_iters"] <= 30, sb["max_inner_iters"]
    st = b.problem_stats()
    assert np.all(st["relres"] < 1e-10)
    for k, n in enumerate(sides):
        pm = mp.Mesh(n, n, n, 6, background=DENSITY)
        ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED))
        ctx.reserve(pm.ne * NPPE[k])
        ctx.fill(NPPE[k], 0, DENSITY, temp=2 * VTH[k] ** 2, seed_offset=k + 1, pert_amp=0.05)
        for _ in range(3):
            s1 = ctx.time_step(0.05)
        phi_a, phi_b = ctx.phi(), b.phi(k)
        # two solvers stopped at |r| < 1e-10 |b|: potentials agree to the solve tolerance
        assert np.linalg.norm(phi_a - phi_b) <= 1e-7 * np.linalg.norm(phi_a), k
        assert abs(st["ke"][k] - s1["kinetic_energy"]) <= 1e-6 * s1["kinetic_energy"], k
        g = by_id(b.download(k, int(st["count"][k])))
        a = by_id(ctx.download())
        assert np.array_equal(g["id"], a["id"].astype(np.uint32))
        assert np.mean(g["elem"] == a["elem"]) > 0.999
        assert np.abs(g["x"] - a["x"]).max() < 1e-4
        ctx.close()
    assert b.check_ownership() == 0
    b.close()


def test_batch_solve_storage_paths(mp):
    """The refinement solve of the batch (multigrid switched off) keeps the
    inner residual in LDS up to 16384 nodes (64 KB), in a larger LDS tile up to
    36864 nodes, in device memory beyond: sides 26 (19683 nodes) and 34 (42875
    nodes) take the last two, each with a small problem beside it, and match
    the same problems run alone.  The multigrid solve of the same ensembles
    matches them too."""
    for sides, mgc in (((26, 3), 0), ((34, 2), 0), ((26, 3), 1)):
        say("batch solve, sides %s, multigrid %d" % (sides, mgc))
        nppe = [2, 5]
        meshes = [mp.Mesh(n, n, n, 6) for n in sides]
        b = mp.Batch(meshes, mp.default_config(multigrid=mgc))
        b.fill([dict(nppe=nppe[k], temp=2 * 0.1 ** 2, density=DENSITY, pert_amp=0.05, seed_offset=k + 1)
                for k in range(2)])
        st = b.step(0.05)
        assert st["all_converged"] == 1 and st["multigrid"] == mgc
        ps = b.problem_stats()
        for k, n in enumerate(sides):
            assert ps["count"][k] == n ** 3 * 6 * nppe[k]
            pm = mp.Mesh(n, n, n, 6, background=DENSITY)
            ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED))
            ctx.reserve(pm.ne * nppe[k])
            ctx.fill(nppe[k], 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=k + 1, pert_amp=0.05)
            s1 = ctx.time_step(0.05)
            phi_a, phi_b = ctx.phi(), b.phi(k)
            assert np.linalg.norm(phi_a - phi_b) <= 1e-7 * np.linalg.norm(phi_a), (sides, k)
            assert abs(ps["ke"][k] - s1["kinetic_energy"]) <= 1e-6 * s1["kinetic_energy"], (sides, k)
            ctx.close()
        b.close()


@SIDE_SETS
def test_batch_graph_replays_equal_eager(mp, sides, mg):
    say("batch graph replays vs eager steps, sides %s" % sides)
    ea = make_batch(mp, sides)
    gr = make_batch(mp, sides)
    gr.graph(True)
    for s in range(5):
        se = ea.step(0.05)
        sg = gr.step(0.05)  # first two steps capture (one graph per buffer parity), then replays
        assert sg["graph"] == 1 and sg["multigrid"] == se["multigrid"] == mg
        pe, pg = ea.problem_stats(), gr.problem_stats()
        for k in range(len(sides)):
            assert np.array_equal(ea.phi(k), gr.phi(k)), (s, k)
        assert np.array_equal(pe["ke"], pg["ke"]) and np.array_equal(pe["iters"], pg["iters"]), s
    for k in range(len(sides)):
        n = int(ea.problem_stats()["count"][k])
        a, g = by_id(ea.download(k, n)), by_id(gr.download(k, n))
        for key in ("x", "v", "elem"):
            assert np.array_equal(a[key], g[key]), (k, key)
    # after a download (which reorders the list) the graph path continues correctly
    ea.step(0.05)
    gr.step(0.05)
    for k in range(len(sides)):
        assert np.array_equal(ea.phi(k), gr.phi(k)), k
    ea.close()
    gr.close()


@SIDE_SETS
def test_batch_tune(mp, sides, mg):
    say("batch autotuner, sides %s" % sides)
    b = make_batch(mp, sides)
    phi0 = [b.phi(k) for k in range(len(sides))]
    b.tune(0.05)
    for k in range(len(sides)):
        assert np.array_equal(b.phi(k), phi0[k])  # tuning advances nothing
    st = b.step(0.05)
    assert st["tuned"] == 1 and st["push_grid"] in (2048, 4096, 8192, 16384)
    assert st["solve_threads"] in (256, 512, 1024) and st["all_converged"] == 1 and st["multigrid"] == mg
    assert b.check_ownership() == 0
    b.close()


def test_batch_two_stream_fill(mp):
    say("batch two-stream fill")
    meshes = [mp.Mesh(n, n, n, 6) for n in (4, 3)]
    b = mp.Batch(meshes, mp.default_config())
    b.fill([dict(nppe=8, temp=2 * 0.05 ** 2, density=DENSITY, drift=0.2, seed_offset=1),
            dict(nppe=5, temp=2 * 0.05 ** 2, density=DENSITY, drift=0.2, seed_offset=2)])
    st = b.problem_stats()
    b.step(0.05)
    st = b.problem_stats()
    for k, m in enumerate(meshes):
        g = b.download(k, int(st["count"][k]))
        vx = g["v"][0]
        half = (vx > 0).mean()
        assert 0.35 < half < 0.65  # two counter-drifting halves
    assert b.check_ownership() == 0
    b.close()


def test_batch_ensemble_c4(mp):
    """The ensemble configuration: 256 problems, sides 6-24 (6 tets per hex),
    8-64 electrons per tet, vth 0.05-0.2, distinct seeds; tuned tiles, graph
    replays, 50 steps of dt 0.05."""
    rng = np.random.default_rng(2024)
    sides = rng.integers(6, 25, size=256)
    nppe = rng.integers(8, 65, size=256)
    vth = rng.uniform(0.05, 0.2, size=256)
    say("C4 ensemble: %d problems, %d tets, %d particles" % (
        256, int((sides ** 3 * 6).sum()), int((sides ** 3 * 6 * nppe).sum())))
    meshes = [mp.Mesh(int(n), int(n), int(n), 6) for n in sides]
    b = mp.Batch(meshes, mp.default_config(cg_tol=1e-10))
    b.fill([dict(nppe=int(nppe[k]), temp=2 * float(vth[k]) ** 2, density=DENSITY, seed_offset=k + 1)
            for k in range(256)])
    count0 = sides.astype(np.int64) ** 3 * 6 * nppe
    b.tune(0.05)
    b.graph(True)
    for s in range(50):
        st = b.step(0.05)
        assert st["all_converged"] == 1 and st["multigrid"] == 1, s
        if s % 10 == 0:
            say("  step %d: %.2f ms, max inner iterations %d" % (s, st["ms_step"], st["max_inner_iters"]))
    ps = b.problem_stats()
    assert np.array_equal(ps["count"], count0)
    assert np.all(ps["relres"] < 1e-10)
    assert np.all(np.isfinite(ps["ke"])) and np.all(ps["ke"] > 0)
    assert b.check_ownership() == 0
    b.close()
