"""mpic_solve_noalloc, the allocation-free variant of mpic_solve (FP64 CG; its
working memory is the library's static device array), called through the C
ABI: equal to mpic_solve bit for bit, on two calls in a row on one GPU (the
second on a smaller mesh, after the first left its state in the array), and
refusing sizes beyond its limits with a clear error."""
import ctypes as C
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DENSITY = 2.0 / 3.0 * math.pi * math.pi  # main.cpp:68


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def solve_both(mp, n, nppe):
    """mpic_solve and mpic_solve_noalloc on the same rhs from the same start."""
    pm = mp.Mesh(n, n, n, 6, background=DENSITY)
    ctx = mp.Context(pm.desc, mp.default_config())
    ctx.reserve(pm.ne * nppe)
    ctx.fill(nppe, 0, DENSITY, temp=0.02, seed_offset=3, pert_amp=0.05)
    ctx.weight_charge()
    phi0 = np.zeros(pm.nn)
    ctx.set_phi(phi0)
    ok, it, rr = ctx.solve(1e-10, 1000)
    want = ctx.phi()
    ctx.set_phi(phi0)
    its, res = C.c_int(), C.c_double()
    rc = mp.lib().mpic_solve_noalloc(ctx.h, 1e-10, 1000, C.byref(its), C.byref(res))
    got = ctx.phi()
    out = (rc, its.value, res.value, it, rr, ok, got, want)
    ctx.close()
    return out


def test_solve_noalloc_equals_solve(mp):
    for n, nppe in ((12, 8), (5, 4)):
        say("noalloc solve, Kuhn %d^3" % n)
        rc, its, res, it, rr, ok, got, want = solve_both(mp, n, nppe)
        assert ok and rc == mp.OK, mp.lib().mpic_last_error(None)
        assert its == it and res == rr and it > 10
        assert np.array_equal(got, want)


def test_solve_noalloc_limits(mp):
    say("noalloc limits")
    pm = mp.Mesh(3, 3, 3, 6, background=DENSITY)
    ctx = mp.Context(pm.desc, mp.default_config())
    its, res = C.c_int(), C.c_double()
    rc = mp.lib().mpic_solve_noalloc(ctx.h, 1e-10, 2048, C.byref(its), C.byref(res))
    assert rc == mp.ERR_ARG and b"limit of 2047" in mp.lib().mpic_last_error(ctx.h)
    rc = mp.lib().mpic_solve_noalloc(ctx.h, 1e-10, 2047, C.byref(its), C.byref(res))
    assert rc == mp.OK and its.value == 0  # no particles, no background rhs at the nodes: a zero residual
    ctx.close()
    ctx = mp.Context(pm.desc, mp.default_config(form=mp.FORM_MIXED))
    rc = mp.lib().mpic_solve_noalloc(ctx.h, 1e-10, 100, C.byref(its), C.byref(res))
    assert rc == mp.ERR_STATE
    ctx.close()
