"""The precision form's multigrid-preconditioned CG (csrc/k_mg.hip, csrc/mg_setup.cpp)
against the FP64 oracle's CG (CG_solve.hpp:54-114): same stopping rule
|r| < tol |b|, the true FP64 residual of the device solution checked with the
oracle's SpMV, the two solutions within the tolerance's error bound, and an
iteration count that does not grow with the mesh.  Meshes: Kuhn and 5-tet
splits, a node mapping, interior jitter, a flat box, up to the C2 mesh."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rhs_for(om, seed):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(om.nn) * 1e-3
    b[om.boundary.astype(bool)] = 0.0
    return b


def solve_device(mp, args, kw, b, multigrid=1, phi0=None):
    pm = mp.Mesh(*args, **kw)
    ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED, multigrid=multigrid))
    ctx.reserve(1024)
    levels, why = ctx.multigrid_levels()
    ctx.set_rhs(b)
    ctx.set_phi(np.zeros(pm.nn) if phi0 is None else phi0)
    ok, it, rr = ctx.solve(1e-10, 1000)
    phi = ctx.phi()
    ctx.close()
    return ok, it, rr, phi, levels, why


# (mesh, options, levels): levels coarsen until at most 512 unknowns remain,
# which the coarsest level solves with its dense inverse
CASES = [
    ((16, 16, 16, 6), {}, 2),
    ((16, 16, 8, 5), {"mapping": 2}, 2),
    ((16, 16, 16, 6), {"jitter": 0.2, "jitter_seed": 3}, 2),
    ((32, 32, 32, 6), {}, 3),
    # odd cell counts: coarse point I on fine point 2I, the last coarse cell one fine cell wide
    ((15, 15, 15, 6), {}, 2),
    ((17, 12, 9, 5), {"jitter": 0.2, "jitter_seed": 5}, 2),
    ((23, 23, 23, 6), {}, 3),
]


@pytest.mark.parametrize("args,kw,levels", CASES)
def test_multigrid_solve_against_oracle(mp, oracle, args, kw, levels):
    say("multigrid solve %s %s" % (args, kw))
    om = oracle.Mesh(*args, **kw)
    b = rhs_for(om, 5)
    ok, it, rr, phi, lv, why = solve_device(mp, args, kw, b)
    assert lv == levels, why
    assert ok and rr < 1e-10 and 0 < it <= 30, (ok, it, rr)
    # the true residual in FP64
    true = np.linalg.norm(b - om.spmv(phi)) / np.linalg.norm(b)
    assert true < 1e-10, true
    phi_o = np.zeros(om.nn)
    ok_o, it_o, rr_o = oracle.cg(om, b, phi_o, 1e-10, 5000)
    assert ok_o and it_o > 2 * it
    # both residuals are below 1e-10 |b|: the solutions agree within cond(A) * 2e-10
    err = np.linalg.norm(phi - phi_o) / np.linalg.norm(phi_o)
    assert err < 2e-7, err
    assert not phi[om.boundary.astype(bool)].any()


def test_multigrid_iterations_do_not_grow_and_beat_refinement(mp, oracle):
    its = {}
    for n in (16, 32):
        om = oracle.Mesh(n, n, n, 6)
        b = rhs_for(om, 9)
        ok, it, rr, phi, lv, _ = solve_device(mp, (n, n, n, 6), {}, b)
        ok_r, it_r, rr_r, phi_r, lv_r, why = solve_device(mp, (n, n, n, 6), {}, b, multigrid=0)
        assert ok and ok_r and lv_r == 0 and "switched off" in why
        assert it_r > 3 * it, (it, it_r)
        assert np.linalg.norm(phi - phi_r) / np.linalg.norm(phi_r) < 2e-7
        its[n] = it
        say("n %d: multigrid %d iterations, refinement CG %d" % (n, it, it_r))
    assert its[32] <= its[16] + 3


def test_multigrid_warm_start_and_zero_rhs(mp, oracle):
    args = (16, 16, 16, 6)
    om = oracle.Mesh(*args)
    b = rhs_for(om, 2)
    ok, it, rr, phi, _, _ = solve_device(mp, args, {}, b)
    # from the solution itself: zero iterations
    ok2, it2, rr2, phi2, _, _ = solve_device(mp, args, {}, b, phi0=phi)
    assert ok2 and it2 == 0 and np.array_equal(phi2, phi)
    # zero right-hand side from zero: the reference's early return (CG_solve.hpp:71-72)
    ok3, it3, rr3, phi3, _, _ = solve_device(mp, args, {}, np.zeros(om.nn))
    assert ok3 and it3 == 0 and not phi3.any()


def test_multigrid_iteration_limit(mp, oracle):
    args = (16, 16, 16, 6)
    om = oracle.Mesh(*args)
    b = rhs_for(om, 4)
    pm = mp.Mesh(*args)
    ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED))
    ctx.reserve(1024)
    ctx.set_rhs(b)
    ctx.set_phi(np.zeros(pm.nn))
    ok, it, rr = ctx.solve(1e-10, 3, check=False)
    assert not ok and it == 3 and rr > 1e-10
    ctx.close()


def test_multigrid_not_applicable(mp):
    say("multigrid fallbacks")
    cases = [((16, 16, 2, 6), {}, "3 cells"), ((8, 8, 8, 5), {"lo": (-1, -1, -1), "hi": (1, 1, 1)}, None),
             ((9, 8, 8, 6), {}, None)]
    for args, kw, reason in cases:
        pm = mp.Mesh(*args, **kw)
        ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED))
        ctx.reserve(1024)
        lv, why = ctx.multigrid_levels()
        if reason is None:
            # the reference's 5-tet box: 8 cells -> 4 (27 unknowns, solved directly); 9x8x8 -> 5x4x4 (36)
            assert lv == 2, why
        else:
            assert lv == 0 and reason in why
        ctx.close()
    pm = mp.Mesh(8, 8, 8, 6)
    ctx = mp.Context(pm.desc, mp.default_config())
    lv, why = ctx.multigrid_levels()
    assert lv == 0 and "FP64" in why
    ctx.close()


def test_multigrid_c2_mesh(mp, oracle):
    """The C2 mesh (Kuhn 64^3, 274,625 nodes): four levels, the true residual
    below 1e-10 |b| in at most 25 iterations."""
    say("multigrid C2 mesh")
    args = (64, 64, 64, 6)
    om = oracle.Mesh(*args)
    b = rhs_for(om, 1)
    ok, it, rr, phi, lv, why = solve_device(mp, args, {}, b)
    assert lv == 4, why
    assert ok and it <= 25, (ok, it, rr)
    assert np.linalg.norm(b - om.spmv(phi)) / np.linalg.norm(b) < 1e-10
