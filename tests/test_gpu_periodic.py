"""The reference's two-stream problem: a mesh periodic in x (Mesh::Mesh with
periodic = true), two counter-streaming electron beams and ions, BiCGStab at
1e-6 (main.cpp:89-106, PIC.cpp:245-267).  The FP64 form against the oracle, bit
for bit, step by step."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY = 2.0 / 3.0 * math.pi * math.pi


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def two_stream_fills(nppe):
    # main.cpp:98-106: nppe/2 electrons at +0.5 and at -0.5 (temp -0.05: the
    # sin perturbation of particle_fill.hpp:106-110), nppe ions at rest
    return [(nppe // 2, 0, DENSITY / 2, (0.5, 0.0, 0.0), -0.05), (nppe // 2, 0, DENSITY / 2, (-0.5, 0.0, 0.0), -0.05),
            (nppe, 2, DENSITY, (0.0, 0.0, 0.0), 0.0)]


@pytest.mark.parametrize("split", [5, 6])
def test_two_stream_time_steps_match_oracle(mp, oracle, split):
    say("two-stream problem, periodic %d-tet mesh 8x4x4, 12 steps, BiCGStab 1e-6 vs oracle" % split)
    nx, ny, nz, nppe, dt = 8, 4, 4, 8, 0.2
    lo, hi = (-1, -1, -1), (1, 1, 1)
    om = oracle.Mesh(nx, ny, nz, split, lo=lo, hi=hi, periodic=True)
    pm = mp.Mesh(nx, ny, nz, split, lo=lo, hi=hi, periodic=True)
    ctx = mp.Context(pm.desc, mp.default_config(solver=mp.SOLVER_BICGSTAB, cg_tol=1e-6))
    p = oracle.Particles(om.ne * nppe * 2)
    ctx.reserve(p.cap)
    for n, t, dens, vel, temp in two_stream_fills(nppe):
        p.fill(om, n, t, dens, vel=vel, temp=temp)
        ctx.fill(n, t, dens, vel=vel, temp=temp)
    # the device fill equals the oracle's bit for bit (the beams' velocity
    # perturbation uses the fill's own sine on both sides); the steps run from it
    g = ctx.download()
    o = np.argsort(g["id"])
    for k in ("x", "xr", "v"):
        assert np.array_equal(g[k][:, o], getattr(p, k)[:, :p.n]), k
    for k in ("elem", "w", "type"):
        assert np.array_equal(g[k][o], getattr(p, k)[:p.n]), k
    S = ctx.charge_scale()
    phi = np.zeros(om.nn)
    x0 = p.x[:, :p.n].copy()
    images = np.where(om.pair >= 0)[0]
    for s in range(12):
        st = ctx.time_step(dt)
        rhs = oracle.deposit(om, p, None, S)
        assert np.all(rhs[images] == 0.0)  # image charge is on the partner rows (PIC.cpp:132-134)
        status, it_o, rr_o = oracle.bicgstab(om, rhs, phi, tol=1e-6)
        oracle.push(om, p, phi, dt)
        assert st["cg_iters"] == it_o and st["cg_relres"] == rr_o, (s, st["cg_iters"], it_o)
        assert np.array_equal(ctx.phi(), phi), s
        assert np.array_equal(ctx.rhs(), rhs), s
        assert st["killed"] == 0
    g = ctx.download()
    o = np.argsort(g["id"])
    assert np.array_equal(g["x"][:, o], p.x[:, :p.n])
    assert np.array_equal(g["v"][:, o], p.v[:, :p.n])
    assert np.array_equal(g["elem"][o], p.elem[:p.n])
    # beams crossed the periodic faces: x jumped by about the period for some particles
    jumped = np.abs(p.x[0, :p.n] - x0[0] - p.v[0, :p.n] * 12 * dt) > 1.0
    assert jumped.sum() > 0
    assert ctx.check_ownership() == 0
    ctx.close()


def test_periodic_mesh_needs_bicgstab(mp):
    say("periodic mesh refuses the CG solver and the precision form")
    pm = mp.Mesh(4, 2, 2, 5, lo=(-1, -1, -1), hi=(1, 1, 1), periodic=True)
    with pytest.raises(mp.MpicError):
        mp.Context(pm.desc, mp.default_config())
    ctx = mp.Context(pm.desc, mp.default_config(solver=mp.SOLVER_BICGSTAB))
    with pytest.raises(mp.MpicError):
        ctx.set_form(mp.FORM_MIXED)
    ctx.close()
    with pytest.raises(mp.MpicError):
        mp.Batch([pm], mp.default_config())
