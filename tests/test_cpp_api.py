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


DUMP = os.path.join(ROOT, "tests", "cpp", "mesh_dump_check")


def expected_dump(om, periodic):
    """Mesh::dump_data's files (mesh.cpp:794-910) computed here from the oracle
    mesh: faces / edges keyed by sorted node tuples in lexicographic order
    (mesh.cpp:233-315, with Mesh::periodic_face, mesh.cpp:62-76), boundary by
    the centroid test of mesh.cpp:604-631, %g formatting (default precision)."""
    X, en = om.nodes, om.elem_nodes
    lo, hi = X.min(0), X.max(0)
    lo_plane = {(x[1], x[2]): i for i, x in enumerate(X) if x[0] == lo[0]}
    faces, edges = set(), set()
    for e in range(om.ne):
        for f in ((0, 1, 3), (1, 2, 3), (0, 3, 2), (0, 2, 1)):
            k = [int(en[e][j]) for j in f]
            if periodic and all(X[n][0] == hi[0] for n in k):
                k = [lo_plane[(X[n][1], X[n][2])] for n in k]
            faces.add(tuple(sorted(k)))
        for j in range(3):
            for m in range(j + 1, 4):
                edges.add(tuple(sorted((int(en[e][j]), int(en[e][m])))))
    faces, edges = sorted(faces), sorted(edges)

    def wall(c):
        s = (2 * c - (lo + hi)) / (hi - lo)
        return bool(np.any(s <= -0.99999999) or np.any(s >= 0.99999999))

    g = "%g"
    out = {
        "nodes-000.dat": "".join("%d %s %s %s \n" % (i, g % x[0], g % x[1], g % x[2]) for i, x in enumerate(X)),
        "elems-000.dat": "".join("%d %d %d %d %d \n" % (e, *en[e]) for e in range(om.ne)),
        "faces-000.dat": "".join("%d %d %d %d \n" % (i, *f) for i, f in enumerate(faces)),
        "edges-000.dat": "".join("%d %d %d \n" % (i, *ed) for i, ed in enumerate(edges)),
        "owned-faces-000.dat": "".join("%d\n" % i for i in range(len(faces))),
        "owned-edges-000.dat": "".join("%d\n" % i for i in range(len(edges))),
        "boundary-faces-000.dat": "".join("%d\n" % i for i, f in enumerate(faces)
                                          if wall((X[f[0]] + X[f[1]] + X[f[2]]) / 3)),
        "boundary-edges-000.dat": "".join("%d\n" % i for i, ed in enumerate(edges) if wall((X[ed[0]] + X[ed[1]]) / 2)),
    }
    return out, len(faces), len(edges)


@pytest.mark.parametrize("nx,ny,nz,mapping,periodic", [(2, 3, 2, 0, False), (3, 2, 2, 1, False), (2, 2, 3, 3, False),
                                                       (4, 2, 3, 0, True)])
def test_mesh_dump_reference_files(oracle, tmp_path, nx, ny, nz, mapping, periodic):
    """main.cpp:88-92's setup (create_mesh(meshfile, comm, periodic)) then the
    dump, host code only.  Pinned by an independent enumeration, by the
    reference's own boundary counts (mesh.cpp:641-647) and by Euler's formula
    V - E + F - T = 1 of a tetrahedral ball."""
    assert os.path.exists(DUMP), "run build() first"
    brick = tmp_path / "brick.txt"
    brick.write_text("%d %d %d %d\n" % (nx, ny, nz, mapping))
    r = subprocess.run([DUMP, str(brick), "1" if periodic else "0", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("rank 0 of 1")
    om = oracle.Mesh(nx, ny, nz, 5, lo=(-1, -1, -1), hi=(1, 1, 1), mapping=mapping, periodic=periodic)
    want, nf, ned = expected_dump(om, periodic)
    for name, text in want.items():
        assert (tmp_path / name).read_text() == text, name
    if not periodic:
        assert 1 == om.nn - ned + nf - om.ne
        nb_f = len(want["boundary-faces-000.dat"].split())
        nb_e = len(want["boundary-edges-000.dat"].split())
        assert nb_f == 2 * (2 * nx * ny + 2 * nx * nz + 2 * ny * nz)
        assert nb_e == (2 * (5 * nx * ny - nx * (ny - 1) - ny * (nx - 1) + 5 * nx * nz - nx * (nz - 1) - nz * (nx - 1)
                             + 5 * ny * nz - ny * (nz - 1) - nz * (ny - 1)) - 4 * (nx + ny + nz))
    else:
        # the 2 ny nz wall triangles on x = hi fold onto their x = lo images
        # (faces only: edges are not joined, mesh.cpp:252), and count once as
        # boundary faces (centroid on x = lo)
        assert nf == 1 - om.nn + ned + om.ne - 2 * ny * nz
        nb_f = len(want["boundary-faces-000.dat"].split())
        assert nb_f == 2 * (2 * nx * ny + 2 * nx * nz) + 2 * ny * nz


@pytest.mark.gpu
def test_driver_output_files_match_oracle(oracle, tmp_path):
    """minipic_main --output 1 writes main.cpp:125-136's output.<t> files (one
    line per sampled electron: x, v at default stream precision) before every
    step, and --dump_mesh 1 the mesh files: against the oracle's particles of
    fill index 0, nppe, 2 nppe, ... at the same times."""
    say("minipic_main: output files")
    nppe, dt, steps = 4, 0.1, 3
    r = subprocess.run([DRIVER, "--nx", "2", "--nparts", str(nppe), "--steps", str(steps), "--dt", str(dt),
                        "--output", "1", "--dump_mesh", "1", "--quiet", "1", "--capacity-factor", "2"],
                       capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    O = oracle
    om = O.Mesh(2, 2, 2, 5, lo=(-1, -1, -1), hi=(1, 1, 1))
    p = O.Particles(om.ne * nppe * 2)
    p.fill(om, nppe, 0, DENSITY)  # main.cpp:108-109
    p.fill(om, nppe, 1, DENSITY)
    S = O.charge_scale(om, p, None)
    phi = np.zeros(om.nn)
    t = 0.0
    for s in range(steps):
        want = "".join("%s %s %s %s %s %s\n" % tuple("%g" % a for a in (*p.x[:, i], *p.v[:, i]))
                       for i in range(0, p.n, nppe) if p.type[i] == 0)
        got = (tmp_path / ("output.%f" % t)).read_text()
        assert got == want, s
        O.step(om, p, phi, dt, bg=None, S=S)
        t += dt
    assert (tmp_path / "nodes-000.dat").exists() and (tmp_path / "boundary-edges-000.dat").exists()


MG_CHECK = os.path.join(ROOT, "tests", "cpp", "mg_check")


def test_multigrid_hierarchy_cpu():
    """The precision form's multigrid hierarchy on the CPU (tests/cpp/mg_check.cpp):
    Galerkin coarse operators equal to a dense P^T A P, V-cycle preconditioned
    CG iteration counts independent of the mesh size (Kuhn and 5-tet splits, a
    node mapping, jitter, even and odd cell counts), refusal of a direction
    with fewer than 3 cells and of periodic meshes."""
    if not os.path.exists(MG_CHECK):
        pytest.skip("tests/cpp not built (make -C tests/cpp)")
    out = subprocess.run([MG_CHECK], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mg_check ok" in out.stdout
    its = [int(l.split("multigrid ")[1].split()[0]) for l in out.stdout.splitlines() if l.startswith("pcg ")]
    assert len(its) == 9 and max(its) <= 25
