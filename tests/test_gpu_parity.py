"""GPU parity of the HIP path against the FP64 CPU oracle, called through the C ABI.

The FP64 form is expected to be bit-identical to the oracle: same setup arrays,
same operation order, exact fixed-point charge sums and the same CG reduction
tree.  Particles are compared by their identity key (the sort reorders them).
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


def make(mp, oracle, nx, split=6, lo=(0, 0, 0), hi=(1, 1, 1), background=None, jitter=0.0, **cfg):
    om = oracle.Mesh(nx, nx, nx, split, lo=lo, hi=hi, jitter=jitter, jitter_seed=5)
    pm = mp.Mesh(nx, nx, nx, split, lo=lo, hi=hi, jitter=jitter, jitter_seed=5, background=background)
    ctx = mp.Context(pm.desc, mp.default_config(**cfg))
    bg = om.background(background) if background is not None else None
    return om, pm, ctx, bg


def upload(ctx, p):
    ctx.reserve(max(p.cap, 1))
    ctx.upload(p.x, p.xr, p.v, p.elem, p.w, p.type, p.id, n=p.n)


def by_id(g):
    order = np.argsort(g["id"], kind="stable")
    return {k: (v[:, order] if v.ndim == 2 else v[order]) for k, v in g.items()}


def assert_particles_equal(g, p, what=""):
    g = by_id(g)
    alive = p.elem[:p.n] >= 0
    ids = np.nonzero(alive)[0]
    assert np.array_equal(g["id"], ids.astype(np.uint32)), what + " ids"
    for k in ("x", "xr", "v"):
        assert np.array_equal(g[k], getattr(p, k)[:, ids]), what + " " + k
    assert np.array_equal(g["elem"], p.elem[ids]), what + " elem"
    assert np.array_equal(g["w"], p.w[ids]), what + " w"
    assert np.array_equal(g["type"], p.type[ids]), what + " type"


@pytest.mark.parametrize("case", ["kuhn8_c1", "ref5_box"])
def test_fill_bitwise(mp, oracle, case):
    say("fill parity " + case)
    if case == "kuhn8_c1":
        om, pm, ctx, _ = make(mp, oracle, 8)
        p = oracle.Particles(om.ne * 16).fill(om, 16, 0, DENSITY, temp=0.02, seed_offset=1)
        ctx.reserve(p.cap)
        ctx.fill(16, 0, DENSITY, temp=0.02, seed_offset=1)
    else:
        om, pm, ctx, _ = make(mp, oracle, 4, split=5, lo=(-1, -1, -1), hi=(1, 1, 1))
        p = oracle.Particles(om.ne * 10).fill(om, 5, 0, DENSITY, temp=1.0)
        p.fill(om, 5, 1, DENSITY, temp=1.0)
        ctx.reserve(p.cap)
        ctx.fill(5, 0, DENSITY, temp=1.0)
        ctx.fill(5, 1, DENSITY, temp=1.0)
    g = ctx.download()
    assert_particles_equal(g, p, case)
    ctx.close()


def test_deposit_solve_gather_bitwise(mp, oracle):
    say("deposit/solve/gather parity kuhn8")
    om, pm, ctx, bg = make(mp, oracle, 8, background=DENSITY)
    p = oracle.Particles(om.ne * 16).fill(om, 16, 0, DENSITY, temp=0.02, seed_offset=1)
    p.x[0, :p.n] += 0  # same particles on both sides
    # perturb weights so the charge is not neutral
    p.w[:p.n] *= 1.0 + 0.05 * np.sin(2 * np.pi * p.x[0, :p.n])
    upload(ctx, p)
    S = oracle.charge_scale(om, p, bg)
    assert ctx.charge_scale() == S
    ctx.weight_charge()
    rhs = oracle.deposit(om, p, bg, S)
    assert np.array_equal(ctx.rhs(), rhs)
    ok, it, rr = ctx.solve(1e-10, 1000)
    phi = np.zeros(om.nn)
    ok_o, it_o, rr_o = oracle.cg(om, rhs, phi, 1e-10, 1000)
    assert ok and ok_o and it == it_o and it > 5
    assert np.array_equal(ctx.phi(), phi)
    ctx.weight_efield()
    E = oracle.gather(om, p, phi)
    assert np.array_equal(ctx.efield(), E[:, :p.n])
    ctx.close()


@pytest.mark.parametrize("case", ["kuhn8_c1", "ref5_box_two_species", "jitter_kill"])
def test_time_steps_bitwise(mp, oracle, case):
    say("time_step parity " + case)
    if case == "kuhn8_c1":
        om, pm, ctx, bg = make(mp, oracle, 8, background=DENSITY, sort_period=3)
        p = oracle.Particles(om.ne * 16).fill(om, 16, 0, DENSITY, temp=0.02, seed_offset=1)
        dt, steps, kill = 0.1, 10, False
    elif case == "ref5_box_two_species":
        om, pm, ctx, bg = make(mp, oracle, 4, split=5, lo=(-1, -1, -1), hi=(1, 1, 1), sort_period=2)
        p = oracle.Particles(om.ne * 20).fill(om, 10, 0, DENSITY, temp=1.0)
        p.fill(om, 10, 1, DENSITY, temp=1.0)
        dt, steps, kill = 0.1, 6, False
    else:
        om, pm, ctx, bg = make(mp, oracle, 5, jitter=0.2, background=DENSITY, kill_on_wall=1, sort_period=0)
        p = oracle.Particles(om.ne * 8).fill(om, 8, 0, DENSITY, temp=0.5, seed_offset=3)
        dt, steps, kill = 0.2, 4, True
    upload(ctx, p)
    S = ctx.charge_scale()
    phi = np.zeros(om.nn)
    for s in range(steps):
        st = ctx.time_step(dt)
        rhs, info = oracle.step(om, p, phi, dt, bg=bg, S=S, kill=kill)
        assert st["cg_iters"] == info["cg_iters"], s
        assert np.array_equal(ctx.rhs(), rhs), s
        assert np.array_equal(ctx.phi(), phi), s
        for k in ("crossings", "reflections", "killed", "capped"):
            assert st[k] == info[k], (s, k)
        assert math.isclose(st["kinetic_energy"], info["ke"], rel_tol=1e-13), s
        assert_particles_equal(ctx.download(), p, "%s step %d" % (case, s))
    if kill:
        assert ctx.num_particles() == int((p.elem[:p.n] >= 0).sum()) < p.n
    ctx.close()


def test_split_phases_equal_fused(mp, oracle):
    say("split vs fused kuhn6")
    res = []
    for mode in ("fused", "split"):
        om, pm, ctx, bg = make(mp, oracle, 6, background=DENSITY, sort_period=0)
        p = oracle.Particles(om.ne * 12).fill(om, 12, 0, DENSITY, temp=0.05, seed_offset=4)
        upload(ctx, p)
        for s in range(4):
            if mode == "fused":
                ctx.time_step(0.1)
            else:
                ctx.weight_charge()
                ctx.solve(1e-10, 1000)
                ctx.weight_efield()
                ctx.accel(0.1)
                ctx.move(0.1)
        res.append((by_id(ctx.download()), ctx.phi()))
        ctx.close()
    (a, pa), (b, pb) = res
    assert np.array_equal(pa, pb)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_sort_counts_and_determinism(mp, oracle):
    say("sort counts kuhn8")
    om, pm, ctx, bg = make(mp, oracle, 8, background=DENSITY, sort_period=0)
    p = oracle.Particles(om.ne * 16).fill(om, 16, 0, DENSITY, temp=0.05, seed_offset=1)
    upload(ctx, p)
    phi = np.zeros(om.nn)
    for s in range(3):
        ctx.time_step(0.1)
        oracle.step(om, p, phi, 0.1, bg=bg, S=ctx.charge_scale())
    start = ctx.sort(want_start=True)
    counts = np.bincount(p.elem[:p.n], minlength=om.ne)
    assert np.array_equal(np.diff(start), counts)
    g = ctx.download()
    assert np.all(np.diff(g["elem"]) >= 0)
    assert_particles_equal(g, p, "after sort")
    # steps after the sort stay bit-identical
    ctx.time_step(0.1)
    oracle.step(om, p, phi, 0.1, bg=bg, S=ctx.charge_scale())
    assert_particles_equal(ctx.download(), p, "post-sort step")
    ctx.close()


def test_rank_of_a_shuffled_list(mp, oracle):
    """The list uploaded in random order: a ranking workgroup's 2,048 particles
    then hold far more keys (elements, or sort buckets in the time step's
    reordering) than its 512-entry table, and most of them take a slot of their
    own with one atomic.  Counts, order and every particle's state against the
    oracle, by element and through three steps reordering every second step."""
    say("rank of a shuffled list kuhn16")
    om, pm, ctx, bg = make(mp, oracle, 16, background=DENSITY, sort_period=2)
    p = oracle.Particles(om.ne * 16).fill(om, 16, 0, DENSITY, temp=0.05, seed_offset=1)
    order = np.random.default_rng(7).permutation(p.n)
    ctx.reserve(p.cap)
    ctx.upload(p.x[:, order], p.xr[:, order], p.v[:, order], p.elem[order], p.w[order], p.type[order], p.id[order],
               n=p.n)
    start = ctx.sort(want_start=True)
    assert np.array_equal(np.diff(start), np.bincount(p.elem[:p.n], minlength=om.ne))
    g = ctx.download()
    assert np.all(np.diff(g["elem"]) >= 0)
    assert_particles_equal(g, p, "after sort")
    ctx.upload(p.x[:, order], p.xr[:, order], p.v[:, order], p.elem[order], p.w[order], p.type[order], p.id[order],
               n=p.n)
    phi = np.zeros(om.nn)
    for s in range(3):
        ctx.time_step(0.1)
        oracle.step(om, p, phi, 0.1, bg=bg, S=ctx.charge_scale())
        assert_particles_equal(ctx.download(), p, "step %d" % s)
    ctx.close()


def oracle_particles(oracle, g, idx=None):
    """Oracle particle set over downloaded arrays (optionally a subset)."""
    P = oracle.Particles(1)
    sel = slice(None) if idx is None else idx
    for k in ("x", "xr", "v"):
        setattr(P, k, np.ascontiguousarray(g[k][:, sel]))
    for k in ("elem", "w", "type", "id"):
        setattr(P, k, np.ascontiguousarray(g[k][sel]))
    P.n = P.cap = P.elem.shape[0]
    return P


def test_full_size_c2_fill(mp, oracle):
    """The device fill of the C2 workload (Kuhn 64^3, 64 electrons/tet, ~100M,
    with the density perturbation, whose device sin and the rejection test's
    device exp are the fill's only transcendental calls) against the oracle's
    fill, every particle, bit for bit."""
    n = 64
    say("full size C2 fill: device")
    pm = mp.Mesh(n, n, n, 6, background=DENSITY)
    ctx = mp.Context(pm.desc, mp.default_config())
    ctx.reserve(pm.ne * 64)
    ctx.fill(64, 0, DENSITY, temp=0.02, seed_offset=1, pert_amp=0.05)
    g = ctx.download()
    ctx.close()
    say("full size C2 fill: oracle")
    om = oracle.Mesh(n, n, n, 6)
    p = oracle.Particles(om.ne * 64).fill(om, 64, 0, DENSITY, temp=0.02, seed_offset=1, pert_amp=0.05)
    assert p.n == pm.ne * 64
    say("full size C2 fill: compare")
    assert_particles_equal(g, p, "C2 fill")


def test_full_size_c2_properties(mp, oracle):
    """Kuhn 64^3 (1.57M tets), 64 electrons/tet (~100M): device fill, two fused
    steps; per-cell counts, node charge and potential bit-identical to the
    oracle at full size; one push bit-identical on a 1M-particle sample."""
    say("full size C2: setup")
    n = 64
    pm = mp.Mesh(n, n, n, 6, background=DENSITY)
    ctx = mp.Context(pm.desc, mp.default_config(sort_period=1))
    ne = pm.ne
    ctx.reserve(ne * 64)
    ctx.fill(64, 0, DENSITY, temp=0.02, seed_offset=1, pert_amp=0.05)
    N = ctx.num_particles()
    assert N == ne * 64
    say("full size C2: two fused steps")
    for _ in range(2):
        st = ctx.time_step(0.05)
        assert st["capped"] == 0 and st["num_particles"] == N
    say("full size C2: counts, deposit, solve vs oracle")
    g = ctx.download()
    P = oracle_particles(oracle, g)
    start = ctx.sort(want_start=True)
    assert np.array_equal(np.diff(start), np.bincount(P.elem, minlength=ne))
    om = oracle.Mesh(n, n, n, 6)
    bg = om.background(DENSITY)
    ctx.weight_charge()
    rhs = oracle.deposit(om, P, bg, ctx.charge_scale())
    assert np.array_equal(ctx.rhs(), rhs)
    phi0 = ctx.phi()
    ok, it, rr = ctx.solve(1e-10, 1000)
    phi = phi0.copy()
    ok_o, it_o, rr_o = oracle.cg(om, rhs, phi, 1e-10, 1000)
    assert ok and ok_o and it == it_o
    assert np.array_equal(ctx.phi(), phi)
    say("full size C2: push on a 1M sample")
    ctx.weight_efield()
    ctx.accel(0.05)
    ctx.move(0.05)
    g2 = by_id(ctx.download())
    rng = np.random.default_rng(0)
    Q = oracle_particles(oracle, g, np.sort(rng.choice(N, 1_000_000, replace=False)))
    oracle.push(om, Q, phi, 0.05)
    sel = Q.id.astype(np.int64)
    for k in ("x", "xr", "v"):
        assert np.array_equal(g2[k][:, sel], getattr(Q, k)), k
    assert np.array_equal(g2["elem"][sel], Q.elem)
    ctx.close()


def test_factored_and_read_basis_values(mp, oracle):
    """The FP64 push forms each element's weighted basis values as (detJ/24) N_i(x_c)
    when the mesh's values are exactly that (the brick generator's are), and reads
    them otherwise.  With a time step short enough that no particle leaves its
    element, the push's deposit of the new positions must reproduce the deposit of
    the old ones (which always reads the values): bit for bit on the generated
    mesh (factored path), and on a mesh with one element's values scaled by
    1 + 2^-40 (read path; that element's nodes show the scaled values)."""
    say("basis values: factored and read deposits")
    om = oracle.Mesh(3, 3, 3, 6)
    p = oracle.Particles(om.ne * 4).fill(om, 4, 0, DENSITY, temp=0.02, seed_offset=2)
    interior = np.nonzero((om.boundary[om.elem_nodes] == 0).all(axis=1))[0]
    k = int(next(e for e in p.elem[:p.n] if e in set(interior.tolist())))  # an element with particles, no wall node
    rhs = {}
    for name in ("factored", "read"):
        wbv = om.wbv.copy()
        if name == "read":
            wbv[k] *= 1.0 + 2.0 ** -40
        arrs = dict(nodes=om.nodes, elem_nodes=om.elem_nodes, neighbors=om.neighbors, normals=om.normals,
                    jinv=om.jinv, detj=om.detj, grad=om.grad, wbv=wbv, boundary=om.boundary, rowptr=om.rowptr,
                    col=om.col, val=om.val)
        ctx = mp.Context(mp.desc_from_arrays(arrs), mp.default_config(sort_period=0))
        upload(ctx, p)
        first = ctx.time_step(1e-9)
        r1 = ctx.rhs()
        ctx.time_step(1e-9)
        r2 = ctx.rhs()
        assert first["crossings"] == 0, name
        assert np.array_equal(r1, r2), name
        rhs[name] = r1
        ctx.close()
    nodes_k = set(om.elem_nodes[k].tolist())
    differ = set(np.nonzero(rhs["factored"] != rhs["read"])[0].tolist())
    assert differ == nodes_k


def test_derived_and_read_gradients(mp, oracle):
    """The FP64 push forms the element's physical basis gradients from its inverse
    Jacobian when the mesh's gradients are exactly sum_j J^-1[j][d] G_ref[i][j] (the
    generated mesh's are), and reads them otherwise.  Perturbing the gradients of
    an element without particles selects the read path and must leave one time step
    bit-identical to the derived path (and to the oracle); perturbing those of an
    element with particles must change exactly their velocities."""
    say("gradients: derived and read gathers")
    om = oracle.Mesh(3, 3, 3, 6)
    full = oracle.Particles(om.ne * 4).fill(om, 4, 0, DENSITY, temp=0.02, seed_offset=4)
    k_empty, k_busy = 80, 78
    keep = np.nonzero(full.elem[:full.n] != k_empty)[0]
    p = oracle.Particles(len(keep))
    for name in ("x", "xr", "v"):
        getattr(p, name)[:, :len(keep)] = getattr(full, name)[:, keep]
    for name in ("elem", "w", "type"):
        getattr(p, name)[:len(keep)] = getattr(full, name)[keep]
    p.id[:len(keep)] = np.arange(len(keep), dtype=p.id.dtype)
    p.n = len(keep)
    runs = {}
    for name, k in (("derived", None), ("read_empty", k_empty), ("read_busy", k_busy)):
        grad = om.grad.copy()
        if k is not None:
            grad[k] *= 1.0 + 2.0 ** -40
        arrs = dict(nodes=om.nodes, elem_nodes=om.elem_nodes, neighbors=om.neighbors, normals=om.normals,
                    jinv=om.jinv, detj=om.detj, grad=grad, wbv=om.wbv, boundary=om.boundary, rowptr=om.rowptr,
                    col=om.col, val=om.val)
        ctx = mp.Context(mp.desc_from_arrays(arrs), mp.default_config(sort_period=0))
        upload(ctx, p)
        ctx.time_step(0.05)
        runs[name] = by_id(ctx.download())
        ctx.close()
    q = p.copy()
    oracle.step(om, q, np.zeros(om.nn), 0.05)
    for key in ("x", "v", "elem"):
        assert np.array_equal(runs["derived"][key], runs["read_empty"][key]), key
    assert np.array_equal(runs["derived"]["v"], q.v[:, :q.n])
    changed = np.nonzero((runs["derived"]["v"] != runs["read_busy"]["v"]).any(axis=0))[0]
    assert len(changed) > 0 and np.all(p.elem[changed] == k_busy)
