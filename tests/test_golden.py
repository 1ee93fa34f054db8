"""Fixtures in tests/golden/ (written by tests/golden/make_golden.py from the
pinned oracle): the oracle and index mirror reproduce them exactly on the CPU,
and the GPU path matches them."""
import os

import numpy as np
import pytest

import cqmm
import oracle
from cases import Case

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["t1_nsf4_full", "t1_mixed_trunc", "t2_nsf13"]


def load(name):
    f = np.load(os.path.join(HERE, name + ".npz"))
    nsf = f["nsf"]
    uniq = np.unique(nsf)
    spec_nsf = int(uniq[0]) if len(uniq) == 1 else ((4, .5), (9, .25), (13, .25))
    ra, rb, rc = f["ranges"]
    case = Case(int(f["natoms"]), spec_nsf, float(ra), float(rb), float(rc), int(f["mult_type"]), seed=int(f["seed"]))
    np.testing.assert_array_equal(case.nsf, nsf)
    return f, case


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    f, case = load(name)
    o, p, q = oracle.triplets(case.desc)
    np.testing.assert_array_equal(o, f["trip_out"])
    np.testing.assert_array_equal(p, f["trip_p"])
    np.testing.assert_array_equal(q, f["trip_q"])
    a, b, c = case.values()
    oracle.mat_mult(case.desc, a, b, c)
    np.testing.assert_array_equal(c if case.mult_type == 1 else a, f["result"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(name):
    f, case = load(name)
    plan = cqmm.Plan(case.desc)
    a, b, c = case.values()
    plan.mat_mult(a, b, c)
    got = c if case.mult_type == 1 else a
    ref = f["result"]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
