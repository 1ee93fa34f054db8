"""CPU tests that pin the oracle and the host index builder (no GPU).

The reference ships no unit-level golden vectors for mat_mult or its kernels
(its only check on this path is the end-to-end test 002 at rtol 1e-4), so the
oracle is pinned by what the geometry and the mathematics fix:
  * the index builder's patterns equal brute-force image enumeration, bit for bit;
  * the oracle's triplet walk (default-kernel loop over halo/cover-set indices)
    equals the brute-force triplet set, with no duplicates;
  * its values equal an independent blockwise re-summation, and for a full
    product the Bloch identity C(q) = A(q) B(q) at random q;
  * special cases: B = identity gives C = A on C's pattern;
  * the ompGemm restatement (CPU baseline) agrees with the default kernels.
"""
import numpy as np
import pytest

import cqmm
import oracle
from oracle import brute
from cases import Case


def ident(blocks):
    return np.stack([blocks["gid_i"], blocks["gid_j"], blocks["shift"][:, 0], blocks["shift"][:, 1],
                     blocks["shift"][:, 2]], axis=1).astype(np.int64)


def sorted_rows(a):
    a = np.asarray(a)
    return a[np.lexsort(a.T[::-1])]


@pytest.mark.parametrize("natoms,rcut,grid", [(27, 6.0, (2, 2, 2)), (40, 9.0, (3, 3, 3)), (64, 12.0, (2, 3, 4))])
def test_pattern_matches_brute_force(natoms, rcut, grid):
    xyz, L = cqmm.random_structure(natoms, 0.05, 2.0, seed=natoms)
    s = cqmm.System(xyz, L, 4, rcut, grid)
    rid = s.add_range(rcut)
    got = ident(s.blocks(rid))
    ref = brute.pattern(xyz, L, grid, rcut)
    assert len(got) == len(ref)
    assert len(np.unique(brute.keys(got, natoms))) == len(got)  # each image once
    np.testing.assert_array_equal(sorted_rows(got), sorted_rows(ref))
    # Rows are stored in primary order with contiguous offsets.
    b = s.blocks(rid)
    sizes = b["nd1"].astype(np.int64) * b["nd2"]
    np.testing.assert_array_equal(b["offset"], np.concatenate([[0], np.cumsum(sizes)[:-1]]))


def _trip_identities(case, o, p, q):
    """Map (out, p, q) offsets to (gi, gk, gj, s1, s2) rows."""
    s = case.sys
    N = s.natoms
    bo = s.blocks(case.idc if case.mult_type == 1 else case.ida)
    bp = s.blocks(case.ida if case.mult_type == 1 else case.idc)
    bq = s.blocks(case.idb)

    def lookup(blocks, off):
        pos = np.searchsorted(blocks["offset"], off - 1)
        assert np.all(blocks["offset"][pos] == off - 1)
        return ident(blocks)[pos]

    io, ip, iq = lookup(bo, o), lookup(bp, p), lookup(bq, q)
    # q = (k, j, s2); type 1: p = (i, k, s1), out = (i, j, s1+s2); type 2: out = (i,k,s1), p = (i,j,s1+s2)
    ik = ip if case.mult_type == 1 else io
    ij = io if case.mult_type == 1 else ip
    assert np.all(ik[:, 1] == iq[:, 0])
    np.testing.assert_array_equal(ij[:, 0], ik[:, 0])
    np.testing.assert_array_equal(ij[:, 1], iq[:, 1])
    np.testing.assert_array_equal(ij[:, 2:], ik[:, 2:] + iq[:, 2:])
    return np.concatenate([ik[:, :2], iq[:, 1:2], ik[:, 2:], iq[:, 2:]], axis=1)


CASES = [
    # natoms, nsf, ra, rb, rc, type
    (40, 4, 6.0, 8.0, 14.0, 1),                   # full product, images (box ~9.3)
    (40, 4, 6.0, 8.0, 9.0, 1),                    # C truncated
    (36, 13, 6.0, 8.0, 8.0, 2),                   # reduction, output truncated
    (48, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1),
    (48, ((4, .5), (9, .25), (13, .25)), 7.0, 9.0, 6.0, 2),
    # Edge cases: one atom (every neighbour is a periodic image of itself), two
    # atoms with ranges shorter than their separation images, a 4/22 mix
    # (blocks above 16x16).
    (1, 4, 4.0, 5.0, 6.0, 1),
    (1, 13, 4.0, 5.0, 3.0, 2),
    (2, 4, 3.0, 3.0, 3.0, 1),
    (20, ((4, .5), (22, .5)), 5.0, 6.0, 7.0, 1),
]


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", CASES)
def test_oracle_triplets_match_brute_force(natoms, nsf, ra, rb, rc, mt):
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=11)
    o, p, q = oracle.triplets(case.desc)
    got = _trip_identities(case, o, p, q)
    N = case.sys.natoms
    xyz, L, g = case.xyz, case.edge, case.grid
    A = brute.pattern(xyz, L, g, ra)
    B = brute.pattern(xyz, L, g, rb)
    Cp = brute.pattern(xyz, L, g, rc)
    # type 1: A(i,k) B(k,j) -> C(i,j); type 2: X(i,k)=A(ra) B(k,j) -> Y(i,j)=C(rc)
    ref = brute.triplets(A, B, Cp, N)
    assert len(got) == len(ref) > 0
    np.testing.assert_array_equal(sorted_rows(got), sorted_rows(ref))


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", CASES)
def test_oracle_values_match_image_product(natoms, nsf, ra, rb, rc, mt):
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=5)
    a, b, c = case.values()
    a0, c0 = a.copy(), c.copy()
    oracle.mat_mult(case.desc, a, b, c)
    s, N = case.sys, case.sys.natoms
    o, p, q = oracle.triplets(case.desc)
    trip = _trip_identities(case, o, p, q)
    idx = {r: brute.BlockIndex(s.blocks(r), N) for r in {case.ida, case.idb, case.idc}}
    if mt == 1:
        ref = brute.image_product(trip, 1, idx[case.idc], idx[case.ida], idx[case.idb], a0, b, len(c))
        got = c
        np.testing.assert_array_equal(a, a0)
    else:
        ref = brute.image_product(trip, 2, idx[case.ida], idx[case.idc], idx[case.idb], c0, b, len(a))
        got = a
        np.testing.assert_array_equal(c, c0)
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(got - ref).max() <= 1e-13 * scale


def test_oracle_bloch_identity_full_product():
    # rc = ra + rb: nothing is truncated, so C(q) = A(q) B(q) for every q.
    case = Case(30, 4, 5.0, 6.0, 11.0, 1, seed=3)
    a, b, c = case.values()
    oracle.mat_mult(case.desc, a, b, c)
    s = case.sys
    nsf = case.nsf
    rng = np.random.default_rng(0)
    for _ in range(3):
        qv = rng.random(3)
        Aq = brute.bloch(s.blocks(case.ida), a, s.natoms, nsf, qv)
        Bq = brute.bloch(s.blocks(case.idb), b, s.natoms, nsf, qv)
        Cq = brute.bloch(s.blocks(case.idc), c, s.natoms, nsf, qv)
        err = np.abs(Cq - Aq @ Bq).max() / np.abs(Cq).max()
        assert err < 1e-12


def test_oracle_bloch_detects_dropped_image():
    # The identity is sensitive: zeroing one C block breaks it.
    case = Case(30, 4, 5.0, 6.0, 11.0, 1, seed=3)
    a, b, c = case.values()
    oracle.mat_mult(case.desc, a, b, c)
    c[16:32] = 0.0
    s = case.sys
    qv = np.array([0.13, 0.71, 0.29])
    Aq = brute.bloch(s.blocks(case.ida), a, s.natoms, case.nsf, qv)
    Bq = brute.bloch(s.blocks(case.idb), b, s.natoms, case.nsf, qv)
    Cq = brute.bloch(s.blocks(case.idc), c, s.natoms, case.nsf, qv)
    assert np.abs(Cq - Aq @ Bq).max() / np.abs(Cq).max() > 1e-6


def test_oracle_identity_b_gives_a_on_c_pattern():
    case = Case(40, 4, 6.0, 8.0, 5.0, 1, seed=9)
    s = case.sys
    a = s.fill(case.ida, 0, 1, 1, (2.0,))
    b = s.fill(case.idb, 3, 0, 0, ())
    c = np.zeros(s.length(case.idc))
    oracle.mat_mult(case.desc, a, b, c)
    ia = brute.BlockIndex(s.blocks(case.ida), s.natoms)
    bc = s.blocks(case.idc)
    pos = ia.find(ident(bc))
    for t in range(len(pos)):
        oa, oc = ia.offset[pos[t]], bc["offset"][t]
        n = bc["nd1"][t] * bc["nd2"][t]
        np.testing.assert_array_equal(c[oc:oc + n], a[oa:oa + n])


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", CASES)
def test_ompgemm_baseline_matches_default(natoms, nsf, ra, rb, rc, mt):
    oracle.use_numpy_openblas()
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=2)
    a, b, c = case.values()
    a2, c2 = a.copy(), c.copy()
    oracle.mat_mult(case.desc, a, b, c)
    oracle.mat_mult_ompgemm(case.desc, a2, b, c2)
    ref, got = (c, c2) if mt == 1 else (a, a2)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_flops_count_matches_triplets():
    case = Case(40, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1, seed=4)
    o, p, q = oracle.triplets(case.desc)
    trip = _trip_identities(case, o, p, q)
    nsf = case.nsf
    fl = 2.0 * (nsf[trip[:, 0]].astype(float) * nsf[trip[:, 1]] * nsf[trip[:, 2]]).sum()
    assert oracle.flops(case.desc) == fl
