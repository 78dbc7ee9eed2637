"""Degenerate geometry and failure paths: the HIP path against the FP64 oracle.

Crafted particles sit on the meshes where every inverse Jacobian is dyadic:
Kuhn 2^3 on [0,1]^3 (h = 1/2) and the reference's 5-tet 2^3 brick on [-1,1]^3.
Positions on vertices, edges and faces, landings exactly on a face, exact
corner hits, velocities along faces (zero reference-velocity components) and
particles at rest are then represented exactly, so the walk's ties
(find_intersection's strictly-smaller replacement, particle_move.hpp:141-178;
exit fractions of -0 / d and 0 / d; several faces at the same s) must be broken
the same way on both sides.  The FP64 form is compared bit for bit, walk
counters included; the precision form is held to its ownership rule.

Failure paths: a fill beyond the capacity is refused and leaves the list alone
(particle_fill.hpp:61-64); CG stops at max_iter (CG_solve.hpp:73,113) and
returns at once on a zero residual (CG_solve.hpp:71-72; the reference's CG
returns false there, this build reports the exact solution as converged with 0
iterations); a particle that exceeds the crossing cap stops, the step completes
and every form returns MPIC_ERR_WALK_CAP.  All loops involved are bounded.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY = 2.0 / 3.0 * math.pi * math.pi  # main.cpp:68


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def containing(om, x):
    """First element whose barycentric coordinates of x are all >= 0, and x_ref."""
    X = om.nodes[om.elem_nodes]
    for e in range(om.ne):
        xr = om.jinv[e] @ (x - X[e, 0])
        if min(1.0 - xr.sum(), *xr) >= 0.0:
            return e, xr
    raise AssertionError("point %s outside the mesh" % (x,))


def crafted(oracle, om, points, vels, scale=1.0):
    cases = [(np.array(x, float), scale * np.array(v, float)) for x in points for v in vels]
    p = oracle.Particles(len(cases))
    for i, (x, v) in enumerate(cases):
        e, xr = containing(om, x)
        p.x[:, i], p.xr[:, i], p.v[:, i], p.elem[i] = x, xr, v, e
        p.w[i], p.type[i], p.id[i] = 1.0, 0, i
    p.n = len(cases)
    return p


KUHN_POINTS = [
    (0.5, 0.5, 0.5),        # interior vertex of all 48 tets' corners
    (0.5, 0.5, 0.25),       # midpoint of an interior edge
    (0.125, 0.125, 0.125),  # on the cube's main diagonal: edge of six tets
    (0.25, 0.25, 0.125),    # on a diagonal face (x = y)
    (0.5, 0.25, 0.125),     # on an interior cube face (x = 1/2)
    (0.0, 0.25, 0.375),     # on the wall x = 0
    (0.0625, 0.0625, 0.0625),
    (0.125, 0.3125, 0.375),
]
KUHN_VELS = [
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, 0, -1), (1, 1, 1), (-1, -1, -1), (1, -1, 0.5),
    (0.25, 0.5, -0.75), (0.5, 0.5, 0), (0, 0, 0), (-0.5, 0.25, 0.25), (1.5, 0, 0),
]
REF5_POINTS = [
    (0.0, 0.0, 0.0), (0.0, 0.0, 0.5), (0.5, 0.5, 0.5), (0.25, 0.25, 0.0), (-1.0, 0.25, 0.5),
    (-0.875, -0.875, -0.875), (0.5, -0.5, 0.25),
]


def mesh_pair(mp, oracle, case):
    if case == "kuhn2":
        om = oracle.Mesh(2, 2, 2, 6)
        pm = mp.Mesh(2, 2, 2, 6)
        return om, pm, KUHN_POINTS, 1.0, 0.25
    om = oracle.Mesh(2, 2, 2, 5, lo=(-1, -1, -1), hi=(1, 1, 1))
    pm = mp.Mesh(2, 2, 2, 5, lo=(-1, -1, -1), hi=(1, 1, 1))
    return om, pm, REF5_POINTS, 2.0, 0.5


def by_id(g):
    order = np.argsort(g["id"], kind="stable")
    return {k: (v[:, order] if v.ndim == 2 else v[order]) for k, v in g.items()}


@pytest.mark.parametrize("case", ["kuhn2", "ref5"])
@pytest.mark.parametrize("kill", [False, True])
def test_crafted_walks_bitwise(mp, oracle, case, kill):
    """Zero field (phi = 0): gather, kick and walk of the crafted set, FP64 form
    against the oracle bit for bit, twice (the second step starts from the
    first's end points, some of them exactly on faces)."""
    say("crafted walks %s kill=%s" % (case, kill))
    om, pm, points, vscale, dt = mesh_pair(mp, oracle, case)
    p = crafted(oracle, om, points, KUHN_VELS, vscale)
    ctx = mp.Context(pm.desc, mp.default_config(kill_on_wall=int(kill), sort_period=0))
    ctx.reserve(p.cap)
    ctx.upload(p.x, p.xr, p.v, p.elem, p.w, p.type, p.id, n=p.n)
    ctx.set_phi(np.zeros(om.nn))
    phi = np.zeros(om.nn)
    for step in range(2):
        ctx.weight_efield()
        ctx.accel(dt)
        st = ctx.move(dt)
        ost = oracle.push(om, p, phi, dt, kill=kill)
        for k, ok in (("crossings", "crossings"), ("killed", "killed"), ("capped", "capped"),
                      ("nonpositive", "nonpos"), ("reflections", "reflections")):
            assert st[k] == ost[ok], (step, k, st[k], ost[ok])
        g = by_id(ctx.download())
        live = np.nonzero(p.elem[:p.n] >= 0)[0]
        assert np.array_equal(g["id"], live.astype(np.uint32)), step
        for k in ("x", "xr", "v"):
            assert np.array_equal(g[k], getattr(p, k)[:, live]), (step, k)
        assert np.array_equal(g["elem"], p.elem[live]), step
        say("  step %d: crossings %d reflections %d nonpositive %d killed %d" % (
            step, st["crossings"], st["reflections"], st["nonpositive"], st["killed"]))
    # the set really exercises the degenerate paths: exit fractions s <= 0
    # (particles starting on a face and leaving through it), and walls
    assert ost["nonpos"] > 0
    assert (ost["killed"] if kill else ost["reflections"]) > 0
    ctx.close()


@pytest.mark.parametrize("case", ["kuhn2", "ref5"])
def test_crafted_walks_precision_form(mp, oracle, case):
    """The same crafted set in the precision form, through time_step with
    uncharged particles (zero charge -> zero right-hand side -> the solve's
    zero-residual return -> zero field): every particle's element contains its
    position under the FP64 test, and positions stay within FP32 rounding of
    the oracle's walk."""
    say("crafted walks precision form %s" % case)
    om, pm, points, vscale, dt = mesh_pair(mp, oracle, case)
    p = crafted(oracle, om, points, KUHN_VELS, vscale)
    ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED, sort_period=1))
    ctx.set_species(np.zeros(3), np.array([1.0, 1.0, 1837.0]))
    ctx.reserve(p.cap)
    ctx.upload(p.x, p.xr, p.v, p.elem, p.w, p.type, p.id, n=p.n)
    phi = np.zeros(om.nn)
    for step in range(2):
        st = ctx.time_step(dt)
        oracle.push(om, p, phi, dt)
        assert st["cg_iters"] == 0 and st["num_particles"] == p.n
        assert ctx.check_ownership() == 0
        g = ctx.download()
        x = np.full((3, p.n), np.nan)
        x[:, g["id"]] = g["x"]
        drift = np.abs(x - p.x[:, :p.n]).max()
        say("  step %d: drift %.2e crossings %d reflections %d" % (step, drift, st["crossings"], st["reflections"]))
        assert drift < 1e-5
    assert np.abs(ctx.phi()).max() == 0.0
    ctx.close()


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_fill_over_capacity(mp, oracle, form):
    say("fill over capacity " + form)
    pm = mp.Mesh(2, 2, 2, 6)
    f = mp.FORM_MIXED if form == "mixed" else mp.FORM_FP64
    ctx = mp.Context(pm.desc, mp.default_config(form=f))
    ctx.reserve(pm.ne * 4 + 5)
    ctx.fill(4, 0, DENSITY, temp=0.02, seed_offset=1)
    assert ctx.num_particles() == pm.ne * 4
    with pytest.raises(mp.MpicError) as e:
        ctx.fill(4, 1, DENSITY, temp=0.02)
    assert e.value.code == mp.ERR_CAPACITY
    assert ctx.num_particles() == pm.ne * 4
    ctx.time_step(0.1)
    assert ctx.num_particles() == pm.ne * 4
    ctx.close()


def test_cg_max_iter_and_zero_rhs(mp, oracle):
    """CG_solve's exits: max_iter (bitwise equal to the oracle's one iteration)
    and the zero-residual return; the refinement solve's max_iter exit."""
    say("cg max_iter / zero rhs")
    om = oracle.Mesh(4, 4, 4, 6)
    pm = mp.Mesh(4, 4, 4, 6)
    p = oracle.Particles(om.ne * 4).fill(om, 4, 0, DENSITY, temp=0.02, seed_offset=1)
    for form in (mp.FORM_FP64, mp.FORM_MIXED):
        ctx = mp.Context(pm.desc, mp.default_config(form=form))
        ctx.reserve(p.cap)
        ctx.upload(p.x, p.xr, p.v, p.elem, p.w, p.type, p.id, n=p.n)
        ctx.weight_charge()
        with pytest.raises(mp.MpicError) as e:
            ctx.solve(1e-10, 1)
        assert e.value.code == mp.ERR_NOT_CONVERGED
        ctx.set_phi(np.zeros(om.nn))
        ok, it, rr = ctx.solve(1e-10, 1, check=False)
        # FP64: exactly one CG iteration; precision form: max_iter bounds the
        # inner iterations of its refinement passes (stopped above tolerance)
        assert not ok and rr > 1e-10
        if form == mp.FORM_FP64:
            assert it == 1 and rr > 1e-3
        if form == mp.FORM_FP64:
            rhs = ctx.rhs()
            phi = np.zeros(om.nn)
            ctx.set_phi(phi)
            ctx.solve(1e-10, 1, check=False)
            assert oracle.cg(om, rhs, phi, 1e-10, 1)[:2] == (False, 1)
            assert np.array_equal(ctx.phi(), phi)
        # a zero right-hand side with a zero start: exact, 0 iterations
        ctx.set_phi(np.zeros(om.nn))
        ctx.set_rhs(np.zeros(om.nn))
        ok, it, rr = ctx.solve(1e-10, 1000)
        assert ok and it == 0 and rr == 0.0
        assert not np.any(ctx.phi())
        ctx.close()


def cap_set(oracle, om):
    pts = [(0.3, 0.2, 0.1), (0.6, 0.7, 0.2), (0.1, 0.4, 0.8), (0.9, 0.1, 0.55)]
    vel = [(3.7, -2.9, 1.3), (1.0, 2.0, 3.0), (-5.0, 0.5, 2.0), (0.25, -4.5, 1.75)]
    p = oracle.Particles(len(pts))
    for i, (x, v) in enumerate(zip(pts, vel)):
        e, xr = containing(om, np.array(x))
        p.x[:, i], p.xr[:, i], p.v[:, i], p.elem[i] = x, xr, v, e
        p.w[i], p.type[i], p.id[i] = 1.0, 0, i
    p.n = len(pts)
    return p


def test_walk_cap_fp64(mp, oracle):
    """A walk of ~500 box lengths stops at the cap: the move completes, returns
    MPIC_ERR_WALK_CAP and equals the oracle's capped walk bit for bit."""
    say("walk cap fp64")
    om = oracle.Mesh(2, 2, 2, 6)
    pm = mp.Mesh(2, 2, 2, 6)
    p = cap_set(oracle, om)
    ctx = mp.Context(pm.desc, mp.default_config(sort_period=0))
    ctx.reserve(p.cap)
    ctx.upload(p.x, p.xr, p.v, p.elem, p.w, p.type, p.id, n=p.n)
    ctx.set_phi(np.zeros(om.nn))
    ctx.weight_efield()
    ctx.accel(100.0)
    rc, st = ctx.move(100.0, check=False)
    ost = oracle.push(om, p, np.zeros(om.nn), 100.0)
    assert rc == mp.ERR_WALK_CAP
    assert st["capped"] == ost["capped"] == p.n
    assert st["crossings"] == ost["crossings"] and st["reflections"] == ost["reflections"]
    g = by_id(ctx.download())
    for k in ("x", "xr", "v"):
        assert np.array_equal(g[k], getattr(p, k)[:, :p.n]), k
    assert np.array_equal(g["elem"], p.elem[:p.n])
    assert ctx.check_ownership() == 0
    ctx.close()


def test_walk_cap_precision_and_batch(mp, oracle):
    """The precision form and the batching form keep the same contract: the
    step completes, the cap hits are counted, MPIC_ERR_WALK_CAP is returned."""
    say("walk cap precision form / batch")
    om = oracle.Mesh(2, 2, 2, 6)
    pm = mp.Mesh(2, 2, 2, 6)
    p = cap_set(oracle, om)
    ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED, sort_period=1))
    ctx.set_species(np.zeros(3), np.array([1.0, 1.0, 1837.0]))
    ctx.reserve(p.cap)
    ctx.upload(p.x, p.xr, p.v, p.elem, p.w, p.type, p.id, n=p.n)
    rc, st = ctx.time_step(100.0, check=False)
    assert rc == mp.ERR_WALK_CAP
    assert st["capped"] == p.n and st["num_particles"] == p.n
    assert ctx.check_ownership() == 0
    rc, st = ctx.time_step(1e-3, check=False)  # a normal step afterwards
    assert rc == mp.OK and st["capped"] == 0
    ctx.close()
    # batching form: two small problems of fast particles
    meshes = [mp.Mesh(2, 2, 2, 6), mp.Mesh(3, 3, 3, 6)]
    b = mp.Batch(meshes, mp.default_config(cg_tol=1e-10))
    b.fill([dict(nppe=2, temp=2 * 25.0 ** 2, density=DENSITY, seed_offset=k + 1) for k in range(2)])
    rc, bst = b.step(50.0, check=False)
    assert rc == mp.ERR_WALK_CAP and bst["capped"] > 0 and bst["crossings"] > 0
    assert b.check_ownership() == 0
    rc, bst = b.step(1e-4, check=False)
    assert rc == mp.OK and bst["capped"] == 0
    b.close()
