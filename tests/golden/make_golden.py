"""Writes the golden fixtures of tests/golden/ from independent computations.

  lcg_kat.json   dRandom (random.hpp:43-87) known answers from Python big-int
                 arithmetic: the state sequence and drand = float(state) * 2**-64.
  fill_kat.json  the first particles of ParticleFill (particle_fill.hpp:130-164)
                 on the reference's 2x2x2 brick of 5-tet hexes on [-1,1]^3
                 (mesh.cpp:83-224), restated in pure Python from the reference
                 text: seed formula, 10 burned draws, rejection, refToPhysical,
                 set_vel with math.sqrt/math.exp; "two_stream" holds the
                 reference's two-stream fills (main.cpp:98-106: beams of
                 type 0 at vel (+-0.5, 0, 0), temp -0.05, then type 2 at rest),
                 whose velocities come from set_vel's temp < 0 branch
                 (particle_fill.hpp:106-111) with the fill's sine (fill_sin
                 below: the build specifies the sine's last bit, which the
                 platform's sin leaves open).
  c1_solve.npz   the C1 problem's electron charge vector (oracle deposit) and the
                 Dirichlet Poisson solution from scipy.sparse.linalg.spsolve.

The reference ships no tests or fixtures, so these vectors come from the
mathematics and from independent restatements, not from running it.
Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

A = 6364136223846793005
Cc = 1442695040888963407
M64 = (1 << 64) - 1


def lcg_next(s):
    return (A * s + Cc) & M64


def drand(s):
    return float(s) * 2.0 ** -64


def lcg_kat():
    out = []
    for seed in (0, 1, 12345, 34030 * 3072 + 7, M64):
        s = seed
        seq = []
        for _ in range(20):
            s = lcg_next(s)
            seq.append([str(s), drand(s).hex()])
        out.append({"seed": str(seed), "seq": seq})
    return out


# Reference 5-tet split (mesh.cpp:200-223), corner c = 4*di + 2*dj + dk.
ODD = [[3, 1, 0, 5], [6, 4, 5, 0], [6, 2, 0, 3], [5, 7, 6, 3], [5, 6, 0, 3]]
EVEN = [[2, 3, 1, 7], [7, 6, 4, 2], [2, 0, 4, 1], [4, 5, 7, 1], [2, 4, 7, 1]]


def ref_brick_elements(n):
    """Element node coordinates of the reference brick n^3 on [-1,1]^3."""
    dx = 2.0 / n
    els = []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                corner = []
                for c in range(8):
                    di, dj, dk = (c >> 2) & 1, (c >> 1) & 1, c & 1
                    corner.append((-1 + (i + di) * dx, -1 + (j + dj) * dx, -1 + (k + dk) * dx))
                tab = ODD if (i + j + k) % 2 != 0 else EVEN
                for t in tab:
                    els.append([corner[v] for v in t])
    return els


def fill_sin(x):
    """The fill's sine, as specified for the build: Cody and Waite's reduction
    x = k pi/2 + r with fdlibm's two-part pi/2 (head 0x1.921fb544p+0), then the
    Taylor polynomials of sin r (to r^21) and cos r (to r^20) in Horner form in
    z = r^2, each operation an IEEE double operation (Python never fuses)."""
    k = float(round(x * float.fromhex("0x1.45f306dc9c883p-1")))
    r = (x - k * float.fromhex("0x1.921fb54400000p+0")) - k * float.fromhex("0x1.0b4611a626331p-34")
    z = r * r
    s = (-1) ** 10 / math.factorial(21)
    for j in range(9, 0, -1):
        s = s * z + (-1) ** j / math.factorial(2 * j + 1)
    c = (-1) ** 10 / math.factorial(20)
    for j in range(9, 0, -1):
        c = c * z + (-1) ** j / math.factorial(2 * j)
    sn = r + (r * z) * s
    cs = 1.0 + z * c
    return (sn, cs, -sn, -cs)[int(k) & 3]


def det3(J):
    return (J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1]
            - J[0][2] * J[1][1] * J[2][0] - J[0][0] * J[1][2] * J[2][1] - J[0][1] * J[1][0] * J[2][2])


def fill_particle(els, elem, ipart, ptype, nppe, density, vel, temp, mass):
    N = len(els)
    s = (elem + (34030 + ipart) * N + N * 12345 * ptype) & M64
    for _ in range(10):
        s = lcg_next(s)

    def dr():
        nonlocal s
        s = lcg_next(s)
        return drand(s)

    while True:
        ref = [dr(), dr(), dr()]
        if not (ref[0] + ref[1] + ref[2] > 1.0):
            break
    X = els[elem]
    wts = [1 - ref[0] - ref[1] - ref[2], ref[0], ref[1], ref[2]]
    ph = [0.0, 0.0, 0.0]
    for i in range(4):
        for d in range(3):
            ph[d] += wts[i] * X[i][d]
    if temp == 0:
        v = list(vel)
    elif temp < 0:
        # beam with a velocity perturbation; no random draws
        v = [vel[0] * (1. + temp * fill_sin(1. * math.pi * ph[0])), 0.0, 0.0]
    else:
        while True:
            r = dr()
            vmax = math.sqrt(6 * temp / mass)
            u = [vmax * 2 * (-.5 + dr()) for _ in range(3)]
            prob = math.exp(-(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / temp)
            if r < prob:
                v = [vel[d] + u[d] for d in range(3)]
                break
    J = [[X[c + 1][r] - X[0][r] for c in range(3)] for r in range(3)]
    vol = det3(J) / 6.
    w = density * vol / nppe
    return {"elem": elem, "ipart": ipart, "type": ptype, "x": [a.hex() for a in ph], "xr": [a.hex() for a in ref],
            "v": [a.hex() for a in v], "w": w.hex()}


def fill_kat():
    els = ref_brick_elements(2)
    density = 2. / 3. * math.pi * math.pi  # main.cpp:68
    out = {"mesh": "reference brick 2x2x2, 5 tets/hex, [-1,1]^3", "nppe": 3, "density": density.hex(),
           "particles": []}
    for ptype in (0, 1):
        for elem in (0, 1, 17, 39):
            for ip in range(3):
                out["particles"].append(fill_particle(els, elem, ip, ptype, 3, density, (0, 0, 0), 1.0, 1.0))
    # two-stream fills on the same brick: two beams of 2 per tet at half density,
    # then the ions (type 2, 4 per tet, at rest)
    out["two_stream"] = []
    for (nppe, ptype, dens, vel, temp) in ((2, 0, density / 2, (0.5, 0.0, 0.0), -0.05),
                                           (2, 0, density / 2, (-0.5, 0.0, 0.0), -0.05),
                                           (4, 2, density, (0.0, 0.0, 0.0), 0.0)):
        fill = {"nppe": nppe, "type": ptype, "density": dens.hex(), "vel": list(vel), "temp": temp, "particles": []}
        for elem in (0, 3, 17, 39):
            for ip in range(nppe):
                fill["particles"].append(fill_particle(els, elem, ip, ptype, nppe, dens, vel, temp, 1.0))
        out["two_stream"].append(fill)
    return out


def c1_solve():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import minipic_oracle as O
    import scipy.sparse.linalg as sla
    m = O.Mesh(8, 8, 8, 6)
    p = O.Particles(m.ne * 16).fill(m, 16, 0, 2. / 3. * math.pi * math.pi, temp=0.02, seed_offset=1)
    rhs = O.deposit(m, p)
    A = m.csr_scipy()
    phi = sla.spsolve(A.tocsc(), rhs)
    np.savez_compressed(os.path.join(HERE, "c1_solve.npz"), rhs=rhs, phi=phi)


if __name__ == "__main__":
    with open(os.path.join(HERE, "lcg_kat.json"), "w") as f:
        json.dump(lcg_kat(), f, indent=1)
    with open(os.path.join(HERE, "fill_kat.json"), "w") as f:
        json.dump(fill_kat(), f, indent=1)
    c1_solve()
    print("golden fixtures written to", HERE)
