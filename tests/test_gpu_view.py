"""The device side of ParticleList (include/minipic/particle_list_view.hpp):
HIP kernels written against the reference's functor API (particle_list.hpp:54-213)
run on a context of the library.  tests/cpp/view_check.hip appends with
push_back, deletes with delete_particle, migrates and kills with
add_particles_to_migrate_list, reduces through the accessors, overflows the
capacity, and checks every result itself (counts, surviving ids, copies equal
to their originals field by field, the reduction against mpic_compute_ke, the
capacity error).  The library's own fill (push_back) and the FP64 move's wall
kills (add_particles_to_migrate_list) go through the same view: their bitwise
parity with the oracle is tested in test_gpu_parity.py and test_gpu_edge_cases.py.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VIEW = os.path.join(ROOT, "tests", "cpp", "view_check")


def test_device_particle_list_functor_api():
    print("view_check", flush=True)
    sys.stdout.flush()
    assert os.path.exists(VIEW), "run build() first"
    r = subprocess.run([VIEW], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "view_check ok" in r.stdout
