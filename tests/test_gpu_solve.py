"""Solver paths: persistent (one cooperative launch per solve, an option) against
per-iteration solves (the default), and the reference's BiCGStab.

Both forms must give the same bits either way: the persistent kernels keep the
per-iteration kernels' operation order and reduction tree.  The FP64 results are
also checked against the oracle's CG and BiCGStab.
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


def run(mp, nx, form, persistent, steps):
    mesh = mp.Mesh(nx, nx, nx, 6, background=DENSITY)
    ctx = mp.Context(mesh.desc, mp.default_config(form=form, persistent_solve=persistent, cg_tol=1e-10))
    ctx.reserve(mesh.ne * 16)
    ctx.fill(16, 0, DENSITY, temp=0.02, seed_offset=3, pert_amp=0.05)
    ctx.weight_charge()
    ok, iters, relres = ctx.solve(1e-10, 1000)
    out = dict(phi0=ctx.phi(), iters0=iters, steps=[])
    for _ in range(steps):
        st = ctx.time_step(0.05)
        out["steps"].append((st["cg_iters"], st["cg_relres"], st["kinetic_energy"]))
    out["phi"] = ctx.phi()
    ctx.close()
    return out


@pytest.mark.parametrize("form", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("nx", [6, 24])
def test_persistent_equals_per_iteration(mp, form, nx):
    say("persistent vs per-iteration solve, form %d, Kuhn %d^3" % (form, nx))
    a = run(mp, nx, form, 1, 3)
    b = run(mp, nx, form, 0, 3)
    assert a["iters0"] == b["iters0"]
    assert np.array_equal(a["phi0"], b["phi0"])
    assert a["steps"] == b["steps"]
    assert np.array_equal(a["phi"], b["phi"])


def test_persistent_fp64_matches_oracle(mp, oracle):
    say("persistent FP64 CG vs oracle CG, Kuhn 16^3")
    om = oracle.Mesh(16, 16, 16, 6)
    pm = mp.Mesh(16, 16, 16, 6)
    ctx = mp.Context(pm.desc, mp.default_config(persistent_solve=1))
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal(om.nn)
    rhs[pm.array("boundary") != 0] = 0.0
    ctx.set_rhs(rhs)
    x0 = rng.standard_normal(om.nn) * 1e-3
    ctx.set_phi(x0)
    ok, iters, relres = ctx.solve(1e-11, 1000)
    phi = x0.copy()
    conv, it_o, rr_o = oracle.cg(om, rhs, phi, tol=1e-11)
    assert ok and conv and iters == it_o
    assert np.array_equal(ctx.phi(), phi)
    ctx.close()


def test_bicgstab_matches_oracle(mp, oracle):
    say("BiCGStab (the reference's ES::solve) vs oracle, 5-tet 6^3 box")
    om = oracle.Mesh(6, 6, 6, 5, lo=(-1, -1, -1), hi=(1, 1, 1))
    pm = mp.Mesh(6, 6, 6, 5, lo=(-1, -1, -1), hi=(1, 1, 1))
    ctx = mp.Context(pm.desc, mp.default_config(solver=mp.SOLVER_BICGSTAB))
    rng = np.random.default_rng(11)
    rhs = rng.standard_normal(om.nn)
    rhs[pm.array("boundary") != 0] = 0.0
    for tol in (1e-6, 1e-11):
        ctx.set_rhs(rhs)
        ctx.set_phi(rng.standard_normal(om.nn))  # discarded: BiCGStab starts from 0
        ok, iters, relres = ctx.solve(tol, 1000)
        phi = rng.standard_normal(om.nn)
        status, it_o, rr_o = oracle.bicgstab(om, rhs, phi, tol=tol)
        assert ok and status == 0 and iters == it_o, (tol, iters, it_o)
        assert np.array_equal(ctx.phi(), phi)
        assert relres == rr_o
    ctx.close()


def test_bicgstab_time_steps_match_oracle(mp, oracle):
    """PIC::time_step as the reference runs it (PIC.cpp:245-267): BiCGStab at
    1e-6 from zero, the box problem (e- and e+ fills at temp 1), 5-tet mesh."""
    say("time steps with BiCGStab 1e-6 vs oracle, reference box problem")
    om = oracle.Mesh(4, 4, 4, 5, lo=(-1, -1, -1), hi=(1, 1, 1))
    pm = mp.Mesh(4, 4, 4, 5, lo=(-1, -1, -1), hi=(1, 1, 1))
    ctx = mp.Context(pm.desc, mp.default_config(solver=mp.SOLVER_BICGSTAB, cg_tol=1e-6))
    p = oracle.Particles(om.ne * 16).fill(om, 8, 0, DENSITY, temp=1.0)
    p.fill(om, 8, 1, DENSITY, temp=1.0)
    ctx.reserve(p.cap)
    ctx.fill(8, 0, DENSITY, temp=1.0)
    ctx.fill(8, 1, DENSITY, temp=1.0)
    S = ctx.charge_scale()
    phi = np.zeros(om.nn)
    for s in range(3):
        st = ctx.time_step(0.1)
        rhs = oracle.deposit(om, p, None, S)
        status, it_o, rr_o = oracle.bicgstab(om, rhs, phi, tol=1e-6)
        oracle.push(om, p, phi, 0.1)
        assert st["cg_iters"] == it_o and st["cg_relres"] == rr_o, (s, st["cg_iters"], it_o)
        assert np.array_equal(ctx.phi(), phi), s
    g = ctx.download()
    o = np.argsort(g["id"])
    assert np.array_equal(g["x"][:, o], p.x[:, :p.n])
    assert np.array_equal(g["elem"][o], p.elem[:p.n])
    ctx.close()
