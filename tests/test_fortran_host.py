"""The Fortran host (iso_c_binding module cqmm_f + cq_gpu_mat_mult, built with
flang) drives the drop-in multiply the way CONQUEST's matrix_product would; its
results must equal the oracle's for the same seeded system."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from cases import Case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "conquest-release_amd", "fortran", "build", "cq_mult_driver")


def test_fortran_host_is_built():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() (flang builds the Fortran host)"
    src = open(os.path.join(ROOT, "conquest-release_amd", "fortran", "cqmm_f.f90")).read()
    assert "bind(c, name=\"cqmm_mat_mult\")" in src


@pytest.mark.gpu
@pytest.mark.parametrize("args", [(40, 4, 6.0, 8.0, 14.0, 1, 3), (36, 13, 8.0, 8.0, 6.0, 2, 4)])
def test_fortran_host_matches_oracle(tmp_path, args):
    natoms, nsf, ra, rb, rc, mt, seed = args
    out = tmp_path / "out.bin"
    print("fortran driver", args, flush=True)
    r = subprocess.run([DRIVER] + [str(x) for x in args] + [str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.uint8)
    ntrip = int(raw[:8].view(np.int64)[0])
    got = raw[8:].view(np.float64)
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=seed)
    a, b, c = case.values()
    oracle.mat_mult(case.desc, a, b, c)
    ref = c if mt == 1 else a
    assert ntrip == len(oracle.triplets(case.desc)[0])
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
