"""The reference's C++ class interface (include/minipic) and driver over the C ABI.

CPU: the driver and the check program are built and refuse cleanly without a
GPU.  GPU: tests/cpp/class_api_check drives DataWarehouse / PIC / ES /
ParticleList / ParticleFill / ParticleMove and its results are compared with the
FP64 oracle (bit for bit where the FP64 form is bit-identical), with the closed
form of a manufactured Poisson problem, and with the exact gradient of a linear
potential; minipic_main runs the reference's driver loop and its KE lines match
the oracle's time steps.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "minipic-crossroads_amd", "minipic_main")
CHECK = os.path.join(ROOT, "tests", "cpp", "class_api_check")
DENSITY = 2.0 / 3.0 * math.pi * math.pi


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def have_gpu_device():
    return os.path.exists("/dev/kfd")


def test_driver_built_and_usage():
    assert os.path.exists(DRIVER), "run build() first"
    r = subprocess.run([DRIVER, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage: minipic_main" in r.stdout


def test_driver_rejects_unknown_problem_and_bad_args():
    r = subprocess.run([DRIVER, "stray"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "unexpected argument" in r.stderr
    # the two-stream problem's periodic mesh needs an even number of hexes in x (mesh.cpp:91-94)
    r = subprocess.run([DRIVER, "--nx", "3", "--problem", "two-stream"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "even number of hexes" in r.stderr


@pytest.mark.skipif(have_gpu_device(), reason="checks the refusal without a GPU")
def test_classes_refuse_without_device():
    r = subprocess.run([DRIVER, "--nx", "2", "--nparts", "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "no HIP device" in r.stderr
    os.makedirs("/tmp/mpic_cpp_check", exist_ok=True)
    r = subprocess.run([CHECK, "/tmp/mpic_cpp_check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no HIP device" in r.stderr


def read_info(path):
    info = {}
    for line in open(path):
        k, v = line.split()
        info[k] = v
    return info


def read_particles(path):
    a = np.loadtxt(path, ndmin=2)
    order = np.argsort(a[:, 0], kind="stable")
    a = a[order]
    return dict(id=a[:, 0].astype(np.int64), elem=a[:, 1].astype(np.int32), type=a[:, 2].astype(np.int32),
                x=a[:, 3:6].T, xr=a[:, 6:9].T, v=a[:, 9:12].T, w=a[:, 12])


def assert_same_particles(g, p, what):
    ids = np.nonzero(p.elem[:p.n] >= 0)[0]
    assert np.array_equal(g["id"], ids), what + " ids"
    assert np.array_equal(g["elem"], p.elem[ids]), what + " elem"
    for k in ("x", "xr", "v"):
        assert np.array_equal(g[k], getattr(p, k)[:, ids]), what + " " + k
    assert np.array_equal(g["w"], p.w[ids]), what + " w"


@pytest.mark.gpu
def test_class_interface_against_oracle(oracle, tmp_path):
    say("C++ class interface: class_api_check on Kuhn 4^3 and 8^3")
    r = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    info = read_info(tmp_path / "info.txt")
    assert info["done"] == "1"
    O = oracle
    om = O.Mesh(4, 4, 4, 6)
    bg = om.background(DENSITY)
    p = O.Particles(om.ne * 32).fill(om, 16, 0, DENSITY, temp=0.02, seed_offset=1)
    S = int(info["charge_scale"])
    assert int(info["num_particles"]) == p.n
    # weight_charge + ES::solve(1e-10) from phi = 0
    split = np.loadtxt(tmp_path / "split.txt")
    rhs = O.deposit(om, p, bg, S)
    assert np.array_equal(split[:, 0], rhs), "ES::rhs after PIC::weight_charge"
    phi = np.zeros(om.nn)
    conv, iters, _ = O.cg(om, rhs, phi, tol=1e-10)
    assert conv and int(info["split_solve_iters"]) == iters
    assert np.array_equal(split[:, 1], phi), "ES::phi after ES::solve"
    # weight_Efield + ParticleMove::move
    st = O.push(om, p, phi, 0.1)
    assert int(info["move_crossings"]) == st["crossings"]
    assert_same_particles(read_particles(tmp_path / "moved.txt"), p, "after ParticleMove::move")
    counts = np.loadtxt(tmp_path / "counts.txt", dtype=np.int64)
    ref_counts = np.bincount(p.elem[:p.n][p.elem[:p.n] >= 0], minlength=om.ne)
    assert np.array_equal(counts[:, 1], ref_counts), "count_for_element"
    assert np.array_equal(counts[:, 0], np.concatenate([[0], np.cumsum(ref_counts)[:-1]])), "start"
    # two PIC::time_step calls (warm-started CG from the split solve's phi)
    for s in range(2):
        _, inf = O.step(om, p, phi, 0.1, bg=bg, S=S)
        assert int(info["step%d_cg_iters" % s]) == inf["cg_iters"]
        ke = float(info["step%d_ke" % s])
        assert abs(ke - inf["ke"]) <= 1e-13 * abs(inf["ke"]), (ke, inf["ke"])
    assert_same_particles(read_particles(tmp_path / "stepped.txt"), p, "after PIC::time_step")
    assert int(info["dump_lines"]) == p.n
    # ES::set_analytic_RHS: -f integrated against N_i; K phi = rhs is the weak form of
    # lap(phi) = f, so f = sin sin sin gives phi = -f / (3 pi^2) (phi = 0 on the boundary).
    an = np.loadtxt(tmp_path / "analytic.txt")
    x, rhs_a, phi_a = an[:, :3], an[:, 3], an[:, 4]
    f = np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]) * np.sin(np.pi * x[:, 2])
    exact = -f / (3 * np.pi ** 2)
    assert info["analytic_converged"] == "1"
    err = np.abs(phi_a - exact).max() / np.abs(exact).max()
    assert err < 0.03, err  # O(h^2) discretisation error at h = 1/8
    om8 = O.Mesh(8, 8, 8, 6)
    interior = np.nonzero(np.abs(rhs_a) > 0)[0]
    assert interior.size > 0 and np.all(rhs_a[interior] < 0)
    # the load vector of the constant 1 is the oracle's background of density 1
    assert abs(rhs_a.sum()) <= abs(om8.background(1.0).sum())
    # ES::write_matrix_market: the oracle's stiffness matrix, and the rhs / potential just compared
    import scipy.io
    K = scipy.io.mmread(str(tmp_path / "matrix.mm")).tocsr()
    K.sort_indices()
    assert np.array_equal(K.indptr, om8.rowptr) and np.array_equal(K.indices, om8.col)
    assert np.array_equal(K.data, om8.val)
    assert np.array_equal(scipy.io.mmread(str(tmp_path / "b.mm")).ravel(), rhs_a)
    assert np.array_equal(scipy.io.mmread(str(tmp_path / "x.mm")).ravel(), phi_a)
    # ES::set_analytic_phi with a linear potential: E = grad phi = (2, 3, -1)
    E = np.loadtxt(tmp_path / "efield.txt", ndmin=2)
    assert E.shape[0] == om8.ne * 4
    assert np.allclose(E, np.array([2.0, 3.0, -1.0])[None, :], rtol=0, atol=1e-12), np.abs(E - [2, 3, -1]).max()


@pytest.mark.gpu
def test_driver_loop_matches_oracle(oracle):
    say("minipic_main: plasma problem, Kuhn 4^3, 3 steps")
    r = subprocess.run([DRIVER, "--nx", "4", "--split", "6", "--unit-cube", "1", "--problem", "plasma",
                        "--nparts", "16", "--vth", "0.1", "--seed", "1", "--steps", "3", "--dt", "0.1",
                        "--capacity-factor", "1"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    kes = [float(l.split()[-1]) for l in r.stdout.splitlines() if l.startswith("Total KE is")]
    assert len(kes) == 3
    assert "Fill time" in r.stdout and "updates/second" in r.stdout and "Timers are:" in r.stdout
    O = oracle
    om = O.Mesh(4, 4, 4, 6)
    bg = om.background(DENSITY)
    p = O.Particles(om.ne * 16).fill(om, 16, 0, DENSITY, temp=2 * 0.1 ** 2, seed_offset=1)
    S = O.charge_scale(om, p, bg)
    phi = np.zeros(om.nn)
    for s in range(3):
        _, inf = O.step(om, p, phi, 0.1, bg=bg, S=S)
        assert abs(kes[s] - inf["ke"]) <= 1e-13 * abs(inf["ke"]), (s, kes[s], inf["ke"])


@pytest.mark.gpu
def test_driver_two_stream_matches_oracle(oracle):
    """minipic_main --problem two-stream: the reference's periodic problem
    (main.cpp:89-106) through the classes, against the oracle's steps."""
    say("minipic_main: two-stream problem, periodic 5-tet 6x3x3, 4 steps")
    r = subprocess.run([DRIVER, "--nx", "6", "--ny", "3", "--nz", "3", "--problem", "two-stream", "--nparts", "8",
                        "--steps", "4", "--dt", "0.2", "--capacity-factor", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    kes = [float(l.split()[-1]) for l in r.stdout.splitlines() if l.startswith("Total KE is")]
    assert len(kes) == 4
    O = oracle
    om = O.Mesh(6, 3, 3, 5, lo=(-1, -1, -1), hi=(1, 1, 1), periodic=True)
    p = O.Particles(om.ne * 16)
    p.fill(om, 4, 0, DENSITY / 2, vel=(0.5, 0, 0), temp=-0.05)
    p.fill(om, 4, 0, DENSITY / 2, vel=(-0.5, 0, 0), temp=-0.05)
    p.fill(om, 8, 2, DENSITY, vel=(0, 0, 0), temp=0.0)
    S = O.charge_scale(om, p, None)
    phi = np.zeros(om.nn)
    for s in range(4):
        rhs = O.deposit(om, p, None, S)
        O.bicgstab(om, rhs, phi, tol=1e-6)
        O.push(om, p, phi, 0.2)
        ke = O.ke(p)
        assert abs(kes[s] - ke) <= 1e-13 * abs(ke), (s, kes[s], ke)
