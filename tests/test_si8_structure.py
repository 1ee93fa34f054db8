"""CPU tests on the structure of the reference's own Order-N test
(testsuite/test_002_bulk_Si_1proc_OrderN: 8 Si atoms in a 10.36 bohr cubic
cell, single-zeta 4 support functions per atom, PAO radius 7.12 bohr so an
S range of 14.24 bohr, L range 16 bohr), read from
tests/golden/si8_test002.json (written from the test's input files by
make_si8.py).

The diamond lattice puts many atom pairs at exactly equal distances, and every
range here exceeds the cell, so each pattern holds many periodic images of
each atom. The index mirror's patterns, the oracle's triplet walk and its
values must still equal the brute-force enumeration and the independent
blockwise re-summation.
"""
import numpy as np
import pytest

import cqmm
import oracle
from oracle import brute
from cases import SI8, SI8_PRODUCTS, si8_case
from test_oracle import _trip_identities, ident, sorted_rows

EDGE = SI8["cell_bohr"][0][0]
XYZ = np.array(SI8["frac"]) * EDGE
NSF = SI8["nsf"]
R_S, R_L = SI8["s_range_bohr"], SI8["l_range_bohr"]


def test_fixture_is_the_reference_test_structure():
    assert np.allclose(SI8["cell_bohr"], np.diag([10.36] * 3))
    assert len(SI8["frac"]) == 8 and NSF == 4
    assert (R_S, R_L) == (14.24, 16.0)
    # Diamond: every atom has 4 nearest neighbours at a*sqrt(3)/4 (the first
    # atom is displaced by 0.001-0.003 of the cell in the input).
    d = np.linalg.norm(XYZ[4] - XYZ[0])
    assert abs(d - EDGE * np.sqrt(3) / 4) < 0.05


@pytest.mark.parametrize("rcut", [R_S, R_L, R_S + R_L])
def test_patterns_match_brute_force(rcut):
    s = cqmm.System(XYZ, EDGE, NSF, rcut, (2, 2, 2))
    rid = s.add_range(rcut)
    got = ident(s.blocks(rid))
    ref = brute.pattern(XYZ, EDGE, (2, 2, 2), rcut)
    assert len(got) == len(ref) > 8 * 8            # images of every atom
    assert len(np.unique(brute.keys(got, 8))) == len(got)
    np.testing.assert_array_equal(sorted_rows(got), sorted_rows(ref))


@pytest.mark.parametrize("ra,rb,rc,mt", SI8_PRODUCTS)
def test_triplets_and_values_match_brute_force(ra, rb, rc, mt):
    case = si8_case(ra, rb, rc, mt)
    o, p, q = oracle.triplets(case.desc)
    got = _trip_identities(case, o, p, q)
    A = brute.pattern(XYZ, EDGE, case.grid, ra)
    B = brute.pattern(XYZ, EDGE, case.grid, rb)
    Cp = brute.pattern(XYZ, EDGE, case.grid, rc)
    ref = brute.triplets(A, B, Cp, 8)
    assert len(got) == len(ref) > 0
    np.testing.assert_array_equal(sorted_rows(got), sorted_rows(ref))

    a, b, c = case.values()
    a0, c0 = a.copy(), c.copy()
    oracle.mat_mult(case.desc, a, b, c)
    s = case.sys
    idx = {r: brute.BlockIndex(s.blocks(r), 8) for r in {case.ida, case.idb, case.idc}}
    if mt == 1:
        want = brute.image_product(got, 1, idx[case.idc], idx[case.ida], idx[case.idb], a0, b, len(c))
        out = c
    else:
        want = brute.image_product(got, 2, idx[case.ida], idx[case.idc], idx[case.idb], c0, b, len(a))
        out = a
    assert np.abs(out - want).max() <= 1e-13 * np.abs(want).max()
