"""ctypes bindings for the FP64 CPU oracle (oracle/minipic_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Particle arrays follow the oracle's layout: positions, reference positions and
velocities are float64 arrays of shape (3, capacity) (component planes);
elem/type are int32, weight float64, id uint32, all of length capacity.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_minipic.so")

# Species table of the reference, particle_type.hpp:65-74: e-, e+, H+.
SPECIES_Q = np.zeros(16)
SPECIES_M = np.ones(16)
SPECIES_Q[:3] = [-1.0, 1.0, 1.0]
SPECIES_M[:3] = [1.0, 1.0, 1837.0]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        d, i32, i64, u64 = C.c_double, C.c_int, C.c_int64, C.c_uint64
        L.orc_mesh_create.restype = P
        L.orc_mesh_create.argtypes = [i32, i32, i32, i32, P, P, d, u64, i32, i32]
        L.orc_mesh_pairs.argtypes = [P, P]
        L.orc_mesh_destroy.argtypes = [P]
        L.orc_mesh_counts.argtypes = [P, P]
        L.orc_mesh_export.argtypes = [P] * 13
        L.orc_elem_stiffness.argtypes = [P, i64, P]
        L.orc_background.argtypes = [P, d, P]
        L.orc_fill.restype = i64
        L.orc_fill.argtypes = [P, i32, i32, d, P, d, u64, d, P, i64, i64, P, P, P, P, P, P, P]
        L.orc_charge_scale.restype = i32
        L.orc_charge_scale.argtypes = [P, i64, P, P, P, P, P]
        L.orc_deposit.argtypes = [P, i64, P, P, P, P, P, i32, P]
        L.orc_deposit_naive.argtypes = [P, i64, P, P, P, P, P]
        L.orc_dot.restype = d
        L.orc_dot.argtypes = [i64, P, P]
        L.orc_spmv.argtypes = [P, P, P]
        L.orc_cg.restype = i32
        L.orc_cg.argtypes = [P, P, P, d, i32, P, P]
        L.orc_bicgstab.restype = i32
        L.orc_bicgstab.argtypes = [P, P, P, d, i32, P, P]
        L.orc_gather.argtypes = [P, i64, i64, P, P, P, P]
        L.orc_push.argtypes = [P, i64, i64, P, P, P, P, P, P, P, d, d, i32, P]
        L.orc_ke.restype = d
        L.orc_ke.argtypes = [i64, i64, P, P, P, P, P]
        L.orc_to_fixed.argtypes = [d, i32, P, P]
        L.orc_from_fixed.restype = d
        L.orc_from_fixed.argtypes = [i64, u64, i32]
        L.orc_scale_for_bound.restype = i32
        L.orc_scale_for_bound.argtypes = [d]
        L.orc_step.restype = i32
        L.orc_step.argtypes = [P, i64, i64, P, P, P, P, P, P, P, P, P, i32, d, d, i32, i32, P, P, P]
        L.orc_set_threads.argtypes = [i32]
        L.orc_fill_sin.argtypes = [i64, P, P]
        L.orc_get_threads.restype = i32
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def fill_sin(x):
    """The fill's sine (fill_sin of the oracle), elementwise over a float64 array."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_fill_sin(x.size, ptr(x), ptr(out))
    return out


class Mesh:
    """Brick mesh: nx*ny*nz hexes on [lo, hi], split 5 (reference, mesh.cpp:83)
    or 6 (Kuhn) tets per hex, optional seeded interior jitter and the
    reference's node mapping 0-3 (mesh.cpp:763-792)."""

    def __init__(self, nx, ny, nz, split=6, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), jitter=0.0, jitter_seed=0,
                 mapping=0, periodic=False):
        L = lib()
        self.lo = np.array(lo, dtype=np.float64)
        self.hi = np.array(hi, dtype=np.float64)
        self.periodic = bool(periodic)
        self.h = L.orc_mesh_create(nx, ny, nz, split, ptr(self.lo), ptr(self.hi), jitter, jitter_seed, mapping,
                                   int(periodic))
        if not self.h:
            raise ValueError("bad mesh parameters or degenerate element")
        cnt = np.zeros(4, dtype=np.int64)
        L.orc_mesh_counts(self.h, ptr(cnt))
        self.nn, self.ne, self.nnz, self.nbnd = (int(c) for c in cnt)
        nn, ne, nnz = self.nn, self.ne, self.nnz
        self.nodes = np.zeros((nn, 3))
        self.elem_nodes = np.zeros((ne, 4), np.int32)
        self.neighbors = np.zeros((ne, 4), np.int32)
        self.normals = np.zeros((ne, 4, 3))
        self.jinv = np.zeros((ne, 3, 3))
        self.detj = np.zeros(ne)
        self.grad = np.zeros((ne, 4, 3))
        self.wbv = np.zeros((ne, 4, 4))
        self.boundary = np.zeros(nn, np.uint8)
        self.rowptr = np.zeros(nn + 1, np.int32)
        self.col = np.zeros(nnz, np.int32)
        self.val = np.zeros(nnz)
        L.orc_mesh_export(self.h, *(ptr(a) for a in (self.nodes, self.elem_nodes, self.neighbors, self.normals,
                                                      self.jinv, self.detj, self.grad, self.wbv, self.boundary,
                                                      self.rowptr, self.col, self.val)))
        self.pair = np.zeros(nn, np.int32)
        L.orc_mesh_pairs(self.h, ptr(self.pair))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def elem_stiffness(self, e):
        K = np.zeros((4, 4))
        lib().orc_elem_stiffness(self.h, e, ptr(K))
        return K

    def background(self, f):
        bg = np.zeros(self.nn)
        lib().orc_background(self.h, f, ptr(bg))
        return bg

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.nn)
        lib().orc_spmv(self.h, ptr(x), ptr(y))
        return y

    def csr_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.nn, self.nn))


class Particles:
    """Particle SoA in the oracle's layout."""

    def __init__(self, cap):
        self.cap = cap
        self.n = 0
        self.x = np.zeros((3, cap))
        self.xr = np.zeros((3, cap))
        self.v = np.zeros((3, cap))
        self.elem = np.full(cap, -1, np.int32)
        self.w = np.zeros(cap)
        self.type = np.zeros(cap, np.int32)
        self.id = np.zeros(cap, np.uint32)

    def copy(self):
        p = Particles(self.cap)
        p.n = self.n
        for k in ("x", "xr", "v", "elem", "w", "type", "id"):
            setattr(p, k, getattr(self, k).copy())
        return p

    def fill(self, mesh, nppe, ptype, density, vel=(0.0, 0.0, 0.0), temp=1.0, seed_offset=0, pert_amp=0.0):
        vel = np.array(vel, dtype=np.float64)
        end = lib().orc_fill(mesh.h, nppe, ptype, density, ptr(vel), temp, seed_offset, pert_amp, ptr(SPECIES_M),
                             self.n, self.cap, ptr(self.x), ptr(self.xr), ptr(self.v), ptr(self.elem), ptr(self.w),
                             ptr(self.type), ptr(self.id))
        if end < 0:
            raise ValueError("capacity exceeded")
        self.n = int(end)
        return self


def charge_scale(mesh, parts, bg=None):
    return lib().orc_charge_scale(mesh.h, parts.n, ptr(parts.elem), ptr(parts.w), ptr(parts.type), ptr(SPECIES_Q),
                                  ptr(bg))


def deposit(mesh, parts, bg=None, S=None):
    if S is None:
        S = charge_scale(mesh, parts, bg)
    rhs = np.zeros(mesh.nn)
    lib().orc_deposit(mesh.h, parts.n, ptr(parts.elem), ptr(parts.w), ptr(parts.type), ptr(SPECIES_Q), ptr(bg), S,
                      ptr(rhs))
    return rhs


def deposit_naive(mesh, parts):
    rhs = np.zeros(mesh.nn)
    lib().orc_deposit_naive(mesh.h, parts.n, ptr(parts.elem), ptr(parts.w), ptr(parts.type), ptr(SPECIES_Q), ptr(rhs))
    return rhs


def dot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().orc_dot(len(a), ptr(a), ptr(b))


def cg(mesh, rhs, phi, tol=1e-10, max_iter=1000):
    """Solves in place into phi (warm start).  Returns (converged, iters, relres)."""
    it = np.zeros(1, np.int32)
    rr = np.zeros(1)
    st = lib().orc_cg(mesh.h, ptr(rhs), ptr(phi), tol, max_iter, ptr(it), ptr(rr))
    return st == 0, int(it[0]), float(rr[0])


def bicgstab(mesh, rhs, phi, tol=1e-6, max_iter=1000):
    """BiCGStab_solve from x = 0, in place into phi.  Returns (status, iters, relres):
    status 0 converged, 1 max_iter, 2 early exit."""
    it = np.zeros(1, np.int32)
    rr = np.zeros(1)
    st = lib().orc_bicgstab(mesh.h, ptr(rhs), ptr(phi), tol, max_iter, ptr(it), ptr(rr))
    return int(st), int(it[0]), float(rr[0])


def gather(mesh, parts, phi):
    E = np.zeros((3, parts.cap))
    lib().orc_gather(mesh.h, parts.n, parts.cap, ptr(parts.xr), ptr(parts.elem), ptr(phi), ptr(E))
    return E


def push(mesh, parts, phi, dt, scale=1.0, kill=False):
    qm = np.where(SPECIES_M != 0, SPECIES_Q / SPECIES_M, 0.0)
    st = np.zeros(5, np.int64)
    lib().orc_push(mesh.h, parts.n, parts.cap, ptr(parts.x), ptr(parts.xr), ptr(parts.v), ptr(parts.elem), ptr(phi),
                   ptr(parts.type), ptr(qm), dt, scale, int(kill), ptr(st))
    return dict(crossings=int(st[0]), killed=int(st[1]), capped=int(st[2]), nonpos=int(st[3]), reflections=int(st[4]))


def ke(parts):
    return lib().orc_ke(parts.n, parts.cap, ptr(parts.v), ptr(parts.elem), ptr(parts.w), ptr(parts.type),
                        ptr(SPECIES_M))


def step(mesh, parts, phi, dt, bg=None, S=None, cg_tol=1e-10, max_iter=1000, kill=False):
    """One PIC::time_step.  Returns (rhs, info dict)."""
    if S is None:
        S = charge_scale(mesh, parts, bg)
    rhs = np.zeros(mesh.nn)
    out = np.zeros(9)
    st = lib().orc_step(mesh.h, parts.n, parts.cap, ptr(parts.x), ptr(parts.xr), ptr(parts.v), ptr(parts.elem),
                        ptr(parts.w), ptr(parts.type), ptr(SPECIES_Q), ptr(SPECIES_M), ptr(bg), S, dt, cg_tol,
                        max_iter, int(kill), ptr(phi), ptr(rhs), ptr(out))
    info = dict(converged=(st == 0), cg_iters=int(out[0]), relres=float(out[1]), ke=float(out[2]),
                crossings=int(out[3]), killed=int(out[4]), capped=int(out[5]), nonpos=int(out[6]),
                reflections=int(out[7]), S=int(out[8]))
    return rhs, info


def to_fixed(v, S):
    hi = C.c_int64()
    lo = C.c_uint64()
    lib().orc_to_fixed(v, S, C.byref(hi), C.byref(lo))
    return (hi.value << 64) | lo.value


def from_fixed(a, S):
    hi = a >> 64
    lo = a & ((1 << 64) - 1)
    return lib().orc_from_fixed(hi, lo, S)


def scale_for_bound(b):
    return lib().orc_scale_for_bound(b)


def set_threads(n):
    lib().orc_set_threads(n)
