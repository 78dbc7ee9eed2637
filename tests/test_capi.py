"""CPU checks of the C ABI library: it loads, exports every declared symbol, and
its host-side setup (mesh generator, Jacobians, gradients, weighted basis
values, stiffness CSR, background load) equals the oracle's bit for bit.
No GPU compute is called here."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "minipic_c.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpic_[a-z_]+)\s*\(", src)))


def test_exports_match_header(mp):
    decl = declared_symbols()
    assert sorted(mp.EXPORTS) == decl
    out = subprocess.run(["nm", "-D", "--defined-only", mp.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (mpic_\w+)", out))
    missing = set(decl) - exported
    assert not missing, missing
    lib = mp.lib()
    for name in decl:
        assert getattr(lib, name) is not None


def test_config_default(mp):
    cfg = mp.default_config()
    assert cfg.form == mp.FORM_FP64 and cfg.cg_tol == 1e-10 and cfg.cg_max_iter == 1000
    assert list(cfg.species_q)[:3] == [-1.0, 1.0, 1.0] and list(cfg.species_m)[:3] == [1.0, 1.0, 1837.0]


CASES = [
    dict(nx=3, ny=2, nz=4, split=5, lo=(-1, -1, -1), hi=(1, 1, 1), jitter=0.0),
    dict(nx=4, ny=4, nz=3, split=6, lo=(0, 0, 0), hi=(1, 1, 1), jitter=0.0),
    dict(nx=5, ny=3, nz=3, split=6, lo=(0, 0, 0), hi=(2, 1, 1.5), jitter=0.2),
    dict(nx=4, ny=4, nz=4, split=5, lo=(-1, -1, -1), hi=(1, 1, 1), jitter=0.0, mapping=1),
    dict(nx=4, ny=3, nz=4, split=5, lo=(-1, -1, -1), hi=(1, 1, 1), jitter=0.0, mapping=2),
    dict(nx=4, ny=4, nz=2, split=6, lo=(-1, -1, -1), hi=(1, 1, 1), jitter=0.0, mapping=3),
    dict(nx=6, ny=3, nz=4, split=5, lo=(-1, -1, -1), hi=(1, 1, 1), jitter=0.0, periodic=True),
    dict(nx=4, ny=4, nz=3, split=6, lo=(0, 0, 0), hi=(1, 1, 1), jitter=0.0, periodic=True),
]


@pytest.mark.parametrize("case", CASES)
def test_host_setup_matches_oracle(mp, oracle, case):
    c = dict(case)
    jit = c.pop("jitter")
    mapping = c.pop("mapping", 0)
    per = c.pop("periodic", False)
    o = oracle.Mesh(c["nx"], c["ny"], c["nz"], c["split"], lo=c["lo"], hi=c["hi"], jitter=jit, jitter_seed=11,
                    mapping=mapping, periodic=per)
    p = mp.Mesh(c["nx"], c["ny"], c["nz"], c["split"], lo=c["lo"], hi=c["hi"], jitter=jit, jitter_seed=11,
                background=1.75, mapping=mapping, periodic=per)
    pairs = [("nodes", o.nodes), ("elem_nodes", o.elem_nodes), ("elem_nbrs", o.neighbors), ("normals", o.normals),
             ("inv_jacobian", o.jinv), ("det_jacobian", o.detj), ("grad_basis", o.grad), ("wbv", o.wbv),
             ("boundary", o.boundary), ("row_ptr", o.rowptr), ("col_idx", o.col), ("values", o.val),
             ("background", o.background(1.75))]
    if per:
        pairs.append(("node_pair", o.pair))
    else:
        assert p.array("node_pair") is None
    for name, ref in pairs:
        got = p.array(name)
        assert got.shape == ref.shape, name
        assert np.array_equal(got, ref), name


def test_bad_arguments(mp):
    with pytest.raises(mp.MpicError):
        mp.Mesh(3, 2, 2, periodic=True)  # the periodic mesh needs an even nx (mesh.cpp:91-94)
    with pytest.raises(mp.MpicError):
        mp.Mesh(0, 1, 1)
    with pytest.raises(mp.MpicError):
        mp.Mesh(2, 2, 2, split=4)
