"""ctypes binding of libminipic.so (the C ABI of include/minipic_c.h).

This is the Python host-side mirror used by the tests and bench.py; C++ users
use include/minipic/*.hpp.  The library must be built (make -C this folder);
loading fails loudly if it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libminipic.so")

OK = 0
ERR_ARG, ERR_CAPACITY, ERR_NOT_CONVERGED, ERR_HIP, ERR_WALK_CAP, ERR_STATE, ERR_NO_DEVICE = -1, -2, -3, -4, -5, -6, -7
FORM_FP64, FORM_MIXED = 0, 1
SOLVER_CG, SOLVER_BICGSTAB = 0, 1
MAX_SPECIES = 16

# Every extern "C" symbol declared in include/minipic_c.h.
EXPORTS = [
    "mpic_config_default", "mpic_last_error", "mpic_mesh_brick", "mpic_mesh_brick_periodic", "mpic_mesh_set_background",
    "mpic_mesh_describe",
    "mpic_mesh_free", "mpic_create", "mpic_destroy", "mpic_set_form", "mpic_particles_reserve",
    "mpic_particles_upload", "mpic_particles_download", "mpic_num_particles", "mpic_fill", "mpic_weight_charge",
    "mpic_solve", "mpic_weight_efield", "mpic_accel", "mpic_move", "mpic_compute_ke", "mpic_sort_by_element",
    "mpic_time_step", "mpic_get_rhs", "mpic_get_phi", "mpic_set_phi", "mpic_get_efield", "mpic_charge_scale",
    "mpic_check_ownership", "mpic_kernel_time", "mpic_mesh_load_vector", "mpic_set_kill_on_wall",
    "mpic_set_species", "mpic_get_config", "mpic_set_rhs",
    "mpic_batch_create", "mpic_batch_destroy", "mpic_batch_fill", "mpic_batch_tune", "mpic_batch_step",
    "mpic_batch_graph", "mpic_batch_problem_stats", "mpic_batch_get_phi", "mpic_batch_get_rhs",
    "mpic_batch_download", "mpic_batch_check_ownership", "mpic_batch_last_error",
    "mpic_particles_view", "mpic_particles_view_commit", "mpic_stream", "mpic_solve_noalloc",
    "mpic_multigrid_levels",
]


class MeshDesc(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("num_elems", C.c_int32), ("nnz", C.c_int64),
                ("nodes", C.c_void_p), ("elem_nodes", C.c_void_p), ("elem_nbrs", C.c_void_p),
                ("normals", C.c_void_p), ("inv_jacobian", C.c_void_p), ("det_jacobian", C.c_void_p),
                ("grad_basis", C.c_void_p), ("wbv", C.c_void_p), ("boundary", C.c_void_p),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("values", C.c_void_p),
                ("background", C.c_void_p), ("node_pair", C.c_void_p), ("lattice", C.c_int32 * 3)]


class Config(C.Structure):
    _fields_ = [("form", C.c_int), ("cg_tol", C.c_double), ("cg_max_iter", C.c_int), ("kill_on_wall", C.c_int),
                ("sort_period", C.c_int), ("sort_bucket", C.c_int), ("num_species", C.c_int),
                ("species_q", C.c_double * MAX_SPECIES), ("species_m", C.c_double * MAX_SPECIES),
                ("solver", C.c_int), ("persistent_solve", C.c_int), ("multigrid", C.c_int)]


class SoaHost(C.Structure):
    _fields_ = [("x", C.c_void_p), ("x_ref", C.c_void_p), ("v", C.c_void_p), ("elem", C.c_void_p),
                ("weight", C.c_void_p), ("type", C.c_void_p), ("id", C.c_void_p), ("stride", C.c_int64)]


class BatchFill(C.Structure):
    _fields_ = [("nppe", C.c_int), ("temp", C.c_double), ("density", C.c_double), ("pert_amp", C.c_double),
                ("drift", C.c_double), ("seed_offset", C.c_uint64)]


class BatchStats(C.Structure):
    _fields_ = [("push_grid", C.c_int), ("solve_threads", C.c_int), ("tuned", C.c_int), ("graph", C.c_int),
                ("max_inner_iters", C.c_int), ("all_converged", C.c_int), ("num_particles", C.c_int64),
                ("ms_step", C.c_double), ("crossings", C.c_int64), ("capped", C.c_int64),
                ("nonpositive", C.c_int64), ("ke_rescaled", C.c_int), ("multigrid", C.c_int)]


class MoveStats(C.Structure):
    _fields_ = [("crossings", C.c_int64), ("killed", C.c_int64), ("capped", C.c_int64),
                ("nonpositive", C.c_int64), ("reflections", C.c_int64), ("relocated", C.c_int64)]


class StepStats(C.Structure):
    _fields_ = [("cg_iters", C.c_int), ("cg_relres", C.c_double), ("kinetic_energy", C.c_double),
                ("move", MoveStats), ("num_particles", C.c_int64), ("charge_scale_bits", C.c_int),
                ("ms_deposit", C.c_double), ("ms_solve", C.c_double), ("ms_push", C.c_double),
                ("ms_sort", C.c_double), ("ms_ke", C.c_double), ("ms_field", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "move"}
        d.update({k: getattr(self.move, k) for k, _ in MoveStats._fields_})
        return d


_lib = None
LOAD_FN = C.CFUNCTYPE(C.c_double, C.POINTER(C.c_double), C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libminipic.so is not built: run make -C %s" % HERE)
        L = C.CDLL(LIB_PATH)
        P, i32, i64, d, u64 = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_uint64
        sig = {
            "mpic_config_default": (None, [P]),
            "mpic_last_error": (C.c_char_p, [P]),
            "mpic_mesh_brick": (i32, [i32, i32, i32, i32, P, P, d, u64, i32, P]),
            "mpic_mesh_brick_periodic": (i32, [i32, i32, i32, i32, P, P, i32, P]),
            "mpic_mesh_set_background": (i32, [P, d]),
            "mpic_mesh_describe": (i32, [P, P]),
            "mpic_mesh_free": (None, [P]),
            "mpic_create": (i32, [P, P, P]),
            "mpic_destroy": (None, [P]),
            "mpic_set_form": (i32, [P, i32]),
            "mpic_particles_reserve": (i32, [P, i64]),
            "mpic_particles_upload": (i32, [P, P, i64]),
            "mpic_particles_download": (i32, [P, P, P]),
            "mpic_num_particles": (i32, [P, P]),
            "mpic_fill": (i32, [P, i32, i32, d, P, d, u64, d]),
            "mpic_weight_charge": (i32, [P]),
            "mpic_solve": (i32, [P, d, i32, P, P]),
            "mpic_weight_efield": (i32, [P]),
            "mpic_accel": (i32, [P, d, d]),
            "mpic_move": (i32, [P, d, P]),
            "mpic_compute_ke": (i32, [P, P]),
            "mpic_sort_by_element": (i32, [P, P]),
            "mpic_time_step": (i32, [P, d, P]),
            "mpic_get_rhs": (i32, [P, P]),
            "mpic_get_phi": (i32, [P, P]),
            "mpic_set_phi": (i32, [P, P]),
            "mpic_get_efield": (i32, [P, P, i64]),
            "mpic_charge_scale": (i32, [P, P]),
            "mpic_check_ownership": (i32, [P, P]),
            "mpic_kernel_time": (i32, [P, i32, P]),
            "mpic_mesh_load_vector": (i32, [P, P, P, P]),
            "mpic_set_kill_on_wall": (i32, [P, i32]),
            "mpic_set_species": (i32, [P, i32, P, P]),
            "mpic_get_config": (i32, [P, P]),
            "mpic_set_rhs": (i32, [P, P]),
            "mpic_batch_create": (i32, [i32, P, P, P]),
            "mpic_batch_destroy": (None, [P]),
            "mpic_batch_fill": (i32, [P, P]),
            "mpic_batch_tune": (i32, [P, d]),
            "mpic_batch_step": (i32, [P, d, P]),
            "mpic_batch_graph": (i32, [P, i32]),
            "mpic_batch_problem_stats": (i32, [P, P, P, P, P, P]),
            "mpic_batch_get_phi": (i32, [P, i32, P]),
            "mpic_batch_get_rhs": (i32, [P, i32, P]),
            "mpic_batch_download": (i32, [P, i32, P, P]),
            "mpic_batch_check_ownership": (i32, [P, P]),
            "mpic_batch_last_error": (C.c_char_p, [P]),
            "mpic_particles_view": (i32, [P, P]),
            "mpic_particles_view_commit": (i32, [P, P, P]),
            "mpic_stream": (i32, [P, P]),
            "mpic_solve_noalloc": (i32, [P, d, i32, P, P]),
            "mpic_multigrid_levels": (i32, [P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class MpicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("minipic error %d: %s" % (code, msg))
        self.code = code


def default_config(**kw):
    cfg = Config()
    lib().mpic_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


class Mesh:
    """Host mesh + discretisation built by the library (Mesh/ES/PIC setup)."""

    def __init__(self, nx, ny, nz, split=6, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), jitter=0.0, jitter_seed=0,
                 background=None, mapping=0, periodic=False):
        self._lo = np.array(lo, dtype=np.float64)
        self._hi = np.array(hi, dtype=np.float64)
        h = C.c_void_p()
        if periodic:
            rc = lib().mpic_mesh_brick_periodic(nx, ny, nz, split, _p(self._lo), _p(self._hi), mapping, C.byref(h))
        else:
            rc = lib().mpic_mesh_brick(nx, ny, nz, split, _p(self._lo), _p(self._hi), jitter, jitter_seed, mapping,
                                       C.byref(h))
        if rc != OK:
            raise MpicError(rc, lib().mpic_last_error(None).decode())
        self.h = h
        if background is not None:
            lib().mpic_mesh_set_background(self.h, background)
        self.desc = MeshDesc()
        lib().mpic_mesh_describe(self.h, C.byref(self.desc))
        self.nn, self.ne, self.nnz = self.desc.num_nodes, self.desc.num_elems, self.desc.nnz

    def nonzeros(self):
        """Stored matrix entries that are not exactly zero, and the diagonal: the
        entries the device solves read (the CSR of a uniform Kuhn brick stores
        15 a row, 7 of them nonzero)."""
        rp, col, val = self.array("row_ptr"), self.array("col_idx"), self.array("values")
        rows = np.repeat(np.arange(self.nn, dtype=np.int64), np.diff(rp))
        return int(np.count_nonzero((val != 0.0) | (col == rows)))

    def array(self, name):
        """numpy copy of one described array."""
        nn, ne, nnz = self.nn, self.ne, self.nnz
        spec = {"nodes": (C.c_double, (nn, 3)), "elem_nodes": (C.c_int32, (ne, 4)), "elem_nbrs": (C.c_int32, (ne, 4)),
                "normals": (C.c_double, (ne, 4, 3)), "inv_jacobian": (C.c_double, (ne, 3, 3)),
                "det_jacobian": (C.c_double, (ne,)), "grad_basis": (C.c_double, (ne, 4, 3)),
                "wbv": (C.c_double, (ne, 4, 4)), "boundary": (C.c_uint8, (nn,)), "row_ptr": (C.c_int32, (nn + 1,)),
                "col_idx": (C.c_int32, (nnz,)), "values": (C.c_double, (nnz,)), "background": (C.c_double, (nn,)),
                "node_pair": (C.c_int32, (nn,))}
        ct, shape = spec[name]
        ptr = getattr(self.desc, name)
        if not ptr:
            return None
        cnt = int(np.prod(shape))
        buf = (ct * cnt).from_address(ptr)
        return np.ctypeslib.as_array(buf).reshape(shape).copy()

    def load_vector(self, f):
        """ES::set_analytic_RHS load vector of the source density f(x) -> float."""
        cb = LOAD_FN(lambda x, _u: float(f((x[0], x[1], x[2]))))
        out = np.zeros(self.nn)
        rc = lib().mpic_mesh_load_vector(self.h, cb, None, _p(out))
        if rc != OK:
            raise MpicError(rc, lib().mpic_last_error(None).decode())
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().mpic_mesh_free(self.h)
            self.h = None


def desc_from_arrays(arrs):
    """MeshDesc over caller-owned numpy arrays (keys as in the oracle's Mesh)."""
    d = MeshDesc()
    d.num_nodes = arrs["nodes"].shape[0]
    d.num_elems = arrs["elem_nodes"].shape[0]
    d.nnz = arrs["col"].shape[0]
    d.nodes, d.elem_nodes, d.elem_nbrs = _p(arrs["nodes"]), _p(arrs["elem_nodes"]), _p(arrs["neighbors"])
    d.normals, d.inv_jacobian, d.det_jacobian = _p(arrs["normals"]), _p(arrs["jinv"]), _p(arrs["detj"])
    d.grad_basis, d.wbv, d.boundary = _p(arrs["grad"]), _p(arrs["wbv"]), _p(arrs["boundary"])
    d.row_ptr, d.col_idx, d.values = _p(arrs["rowptr"]), _p(arrs["col"]), _p(arrs["val"])
    d.background = _p(arrs.get("background"))
    return d


class Context:
    """One problem on the GPU (the device half of DataWarehouse + PIC)."""

    def __init__(self, desc, config=None):
        self.cfg = config if config is not None else default_config()
        h = C.c_void_p()
        rc = lib().mpic_create(C.byref(desc), C.byref(self.cfg), C.byref(h))
        if rc != OK:
            raise MpicError(rc, lib().mpic_last_error(None).decode())
        self.h = h
        self.nn, self.ne = desc.num_nodes, desc.num_elems

    def close(self):
        if getattr(self, "h", None):
            lib().mpic_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            raise MpicError(rc, lib().mpic_last_error(self.h).decode())
        return rc

    def reserve(self, cap):
        self._chk(lib().mpic_particles_reserve(self.h, cap))

    def upload(self, x, xr, v, elem, w, typ, ids=None, n=None):
        # the library reads rows of a [3][stride] layout and plain vectors: any
        # other memory order (a transposed or fancy-indexed view) is copied first
        x, xr, v = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, xr, v))
        if not (x.shape == xr.shape == v.shape and x.shape[0] == 3):
            raise ValueError("x, xr and v must be [3][stride] arrays of one shape")
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        typ = np.ascontiguousarray(typ, dtype=np.int32)
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        stride = x.shape[1]
        n = stride if n is None else n
        s = SoaHost(_p(x), _p(xr), _p(v), _p(elem), _p(w), _p(typ), _p(ids), stride)
        self._chk(lib().mpic_particles_upload(self.h, C.byref(s), n))

    def download(self):
        n = self.num_particles()
        x, xr, v = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
        elem, w = np.zeros(n, np.int32), np.zeros(n)
        typ, ids = np.zeros(n, np.int32), np.zeros(n, np.uint32)
        s = SoaHost(_p(x), _p(xr), _p(v), _p(elem), _p(w), _p(typ), _p(ids), n)
        cnt = C.c_int64()
        self._chk(lib().mpic_particles_download(self.h, C.byref(s), C.byref(cnt)))
        return dict(x=x, xr=xr, v=v, elem=elem, w=w, type=typ, id=ids)

    def num_particles(self):
        n = C.c_int64()
        self._chk(lib().mpic_num_particles(self.h, C.byref(n)))
        return n.value

    def fill(self, nppe, ptype, density, vel=(0.0, 0.0, 0.0), temp=1.0, seed_offset=0, pert_amp=0.0):
        v = np.array(vel, dtype=np.float64)
        self._chk(lib().mpic_fill(self.h, nppe, ptype, density, _p(v), temp, seed_offset, pert_amp))

    def weight_charge(self):
        self._chk(lib().mpic_weight_charge(self.h))

    def solve(self, tol=1e-10, max_iter=1000, check=True):
        it, rr = C.c_int(), C.c_double()
        rc = lib().mpic_solve(self.h, tol, max_iter, C.byref(it), C.byref(rr))
        if check:
            self._chk(rc)
        return rc == OK, it.value, rr.value

    def multigrid_levels(self):
        """(levels, reason): the precision form's multigrid levels, or 0 and why not."""
        lv = C.c_int()
        self._chk(lib().mpic_multigrid_levels(self.h, C.byref(lv)))
        return lv.value, ("" if lv.value else lib().mpic_last_error(self.h).decode())

    def weight_efield(self):
        self._chk(lib().mpic_weight_efield(self.h))

    def accel(self, dt, scale=1.0):
        self._chk(lib().mpic_accel(self.h, dt, scale))

    def move(self, dt, check=True):
        """MoveParticle::execute.  check=False returns (status, stats) instead of
        raising on a non-OK status (e.g. ERR_WALK_CAP, which still completes the move)."""
        st = MoveStats()
        rc = lib().mpic_move(self.h, dt, C.byref(st))
        out = {k: getattr(st, k) for k, _ in MoveStats._fields_}
        if not check:
            return rc, out
        self._chk(rc)
        return out

    def compute_ke(self):
        ke = C.c_double()
        self._chk(lib().mpic_compute_ke(self.h, C.byref(ke)))
        return ke.value

    def sort(self, want_start=False):
        start = np.zeros(self.ne + 1, np.int32) if want_start else None
        self._chk(lib().mpic_sort_by_element(self.h, _p(start)))
        return start

    def time_step(self, dt, check=True):
        """PIC::time_step.  check=False returns (status, stats) instead of raising
        on a non-OK status (e.g. ERR_WALK_CAP, which still completes the step)."""
        st = StepStats()
        rc = lib().mpic_time_step(self.h, dt, C.byref(st))
        if not check:
            return rc, st.as_dict()
        self._chk(rc)
        return st.as_dict()

    def rhs(self):
        a = np.zeros(self.nn)
        self._chk(lib().mpic_get_rhs(self.h, _p(a)))
        return a

    def phi(self):
        a = np.zeros(self.nn)
        self._chk(lib().mpic_get_phi(self.h, _p(a)))
        return a

    def set_phi(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        self._chk(lib().mpic_set_phi(self.h, _p(phi)))

    def set_rhs(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        self._chk(lib().mpic_set_rhs(self.h, _p(rhs)))

    def set_species(self, q, m):
        q = np.ascontiguousarray(q, dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.float64)
        self._chk(lib().mpic_set_species(self.h, len(q), _p(q), _p(m)))

    def set_kill_on_wall(self, kill):
        self._chk(lib().mpic_set_kill_on_wall(self.h, int(bool(kill))))

    def config(self):
        cfg = Config()
        self._chk(lib().mpic_get_config(self.h, C.byref(cfg)))
        return cfg

    def efield(self):
        n = self.num_particles()
        E = np.zeros((3, n))
        self._chk(lib().mpic_get_efield(self.h, _p(E), n))
        return E

    def charge_scale(self):
        s = C.c_int()
        self._chk(lib().mpic_charge_scale(self.h, C.byref(s)))
        return s.value

    def set_form(self, form):
        self._chk(lib().mpic_set_form(self.h, form))

    def check_ownership(self):
        n = C.c_int64()
        self._chk(lib().mpic_check_ownership(self.h, C.byref(n)))
        return n.value

    def kernel_time(self, which):
        ms = C.c_double()
        self._chk(lib().mpic_kernel_time(self.h, which, C.byref(ms)))
        return ms.value


class Batch:
    """An ensemble of independent problems (the batching form)."""

    def __init__(self, meshes, config=None):
        self.meshes = list(meshes)
        self.nprob = len(self.meshes)
        arr = (C.c_void_p * self.nprob)(*[m.h.value for m in self.meshes])
        cfg = config if config is not None else default_config()
        h = C.c_void_p()
        rc = lib().mpic_batch_create(self.nprob, arr, C.byref(cfg), C.byref(h))
        if rc != OK:
            raise MpicError(rc, lib().mpic_batch_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != OK:
            raise MpicError(rc, lib().mpic_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().mpic_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def fill(self, fills):
        """fills: one dict per problem with nppe, temp, density, pert_amp, drift, seed_offset."""
        arr = (BatchFill * self.nprob)()
        for k, f in enumerate(fills):
            arr[k] = BatchFill(f["nppe"], f.get("temp", 0.0), f["density"], f.get("pert_amp", 0.0),
                               f.get("drift", 0.0), f.get("seed_offset", 0))
        self._chk(lib().mpic_batch_fill(self.h, arr))

    def tune(self, dt):
        self._chk(lib().mpic_batch_tune(self.h, dt))

    def graph(self, enable):
        self._chk(lib().mpic_batch_graph(self.h, int(enable)))

    def step(self, dt, check=True):
        """One time step of every problem; check=False returns (status, stats)."""
        st = BatchStats()
        rc = lib().mpic_batch_step(self.h, dt, C.byref(st))
        out = {k: getattr(st, k) for k, _ in BatchStats._fields_}
        if not check:
            return rc, out
        self._chk(rc)
        return out

    def problem_stats(self):
        n = self.nprob
        ke, rr = np.zeros(n), np.zeros(n)
        it, cv = np.zeros(n, np.int32), np.zeros(n, np.int32)
        cnt = np.zeros(n, np.int64)
        self._chk(lib().mpic_batch_problem_stats(self.h, _p(ke), _p(it), _p(rr), _p(cv), _p(cnt)))
        return dict(ke=ke, iters=it, relres=rr, converged=cv, count=cnt)

    def phi(self, p):
        a = np.zeros(self.meshes[p].nn)
        self._chk(lib().mpic_batch_get_phi(self.h, p, _p(a)))
        return a

    def rhs(self, p):
        a = np.zeros(self.meshes[p].nn)
        self._chk(lib().mpic_batch_get_rhs(self.h, p, _p(a)))
        return a

    def download(self, p, count):
        n = count
        x, xr, v = np.zeros((3, n)), np.zeros((3, n)), np.zeros((3, n))
        elem, w = np.zeros(n, np.int32), np.zeros(n)
        typ, ids = np.zeros(n, np.int32), np.zeros(n, np.uint32)
        s = SoaHost(_p(x), _p(xr), _p(v), _p(elem), _p(w), _p(typ), _p(ids), n)
        cnt = C.c_int64()
        self._chk(lib().mpic_batch_download(self.h, p, C.byref(s), C.byref(cnt)))
        k = cnt.value
        return dict(x=x[:, :k], xr=xr[:, :k], v=v[:, :k], elem=elem[:k], w=w[:k], type=typ[:k], id=ids[:k])

    def check_ownership(self):
        n = C.c_int64()
        self._chk(lib().mpic_batch_check_ownership(self.h, C.byref(n)))
        return n.value
