"""CPU tests that pin the LNV / DMM oracle (oracle/lnv.py).

The reference holds no fixture for LNV_matrix_multiply or the DMM helpers (its
only check on this path is the end-to-end test 001/002 energy), so they are
pinned by what the mathematics fixes:
  * M3 and phi are the gradients of E = Tr[H K] and Ne = Tr[S K] with respect
    to L (finite differences along random directions; a dropped term, a wrong
    sign or a missing transpose changes the directional derivative);
  * the S-idempotent projector P is a fixed point: K(P) = P, Ne = n_occ;
  * the sparse restatement (image-resolved blocks, oracle mat_mult, type-2
    products, transposes) equals the dense closed form on an isolated cluster;
  * SolveCubic returns a root of the cubic (the smallest-|z| one of three
    real roots), checked against numpy's polynomial roots;
  * DoPulay's coefficients are the Lagrange minimiser of a.A.a with sum a = 1.
"""
import numpy as np
import pytest

from cqmm.dm import Purifier
from cases import cluster_structure
from oracle import lnv
from oracle import mcweeny as mo


def random_spd(n, rng, scale=0.1):
    A = rng.normal(size=(n, n)) * scale
    return np.eye(n) + 0.5 * (A + A.T)


def sym(rng, n, scale=1.0):
    A = rng.normal(size=(n, n)) * scale
    return 0.5 * (A + A.T)


@pytest.mark.parametrize("seed", [0, 1])
def test_dense_lnv_gradients_are_finite_differences(seed):
    rng = np.random.default_rng(seed)
    n = 24
    S = random_spd(n, rng)
    H = sym(rng, n)
    L = 0.5 * np.linalg.inv(S) + sym(rng, n, 0.05)
    r = lnv.lnv_dense(L, S, H)
    h = 1e-5
    for D in (sym(rng, n), rng.normal(size=(n, n))):   # symmetric and general directions
        rp, rm = lnv.lnv_dense(L + h * D, S, H), lnv.lnv_dense(L - h * D, S, H)
        dE = (rp["energy"] - rm["energy"]) / (2 * h)
        dN = (rp["electrons"] - rm["electrons"]) / (2 * h)
        assert abs(dE - np.sum(r["M3"] * D)) < 1e-7 * max(1.0, abs(dE))
        assert abs(dN - np.sum(r["phi"] * D)) < 1e-7 * max(1.0, abs(dN))


def test_dense_lnv_projector_is_fixed_point():
    rng = np.random.default_rng(2)
    n = 30
    S = random_spd(n, rng)
    H = sym(rng, n)
    P, _ = mo.projector(H, S, 11)
    r = lnv.lnv_dense(P, S, H)
    np.testing.assert_allclose(r["K"], P, atol=1e-11)
    assert abs(r["electrons"] - 11) < 1e-10
    assert abs(r["energy"] - np.sum(H * P)) < 1e-10
    # At the projector Ne is stationary along the S-metric subspace rotations only;
    # phi = 6(SPS - SPSPS) vanishes there.
    np.testing.assert_allclose(r["phi"], 0.0, atol=1e-10)


def cluster_purifier(natoms=9, nsf=4, seed=3):
    xyz, box = cluster_structure(natoms, seed=seed)
    return Purifier(xyz, box, nsf, r_S=8.0, r_H=8.0, r_L=12.0, grid=(5, 5, 5), seed=seed, device=False, lnv=True)


def test_sparse_transpose_matches_dense():
    pu = cluster_purifier(natoms=7, nsf=[4, 9, 4, 1, 9, 4, 13][:7])
    s, R = pu.sys, pu.rid
    sp = mo.Sparse(s, R)
    for name in ("L", "LS", "LSL"):
        b = s.blocks(R[name])
        v = np.random.default_rng(4).normal(size=s.length(R[name]))
        M = mo.densify(b, v, pu.nsf)
        np.testing.assert_array_equal(mo.densify(b, lnv.transpose(sp, v, name), pu.nsf), M.T)


def test_sparse_lnv_equals_dense_on_cluster():
    pu = cluster_purifier()
    s, R = pu.sys, pu.rid
    sp = mo.Sparse(s, R)
    nsf = pu.nsf
    bS, bH, bL = s.blocks(R["S"]), s.blocks(R["H"]), s.blocks(R["L"])
    S = mo.densify(bS, pu.S_vals, nsf)
    H = mo.densify(bH, pu.H_vals, nsf)
    rng = np.random.default_rng(5)
    n = S.shape[0]
    A = rng.normal(size=(n, n)) * 0.05
    Ld = np.linalg.inv(S) * 0.5 + 0.5 * (A + A.T)
    L = mo.sparsify(Ld, bL, nsf, s.length(R["L"]))
    r = lnv.lnv_sparse(sp, L, pu.S_vals, pu.H_vals)
    d = lnv.lnv_dense(Ld, S, H)
    assert abs(r["energy"] - d["energy"]) < 1e-10 * max(1, abs(d["energy"]))
    assert abs(r["electrons"] - d["electrons"]) < 1e-10 * max(1, abs(d["electrons"]))
    np.testing.assert_allclose(mo.densify(bH, r["K"], nsf), d["K"], atol=1e-12)
    np.testing.assert_allclose(mo.densify(bL, r["M3"], nsf), d["M3"], atol=1e-11)
    np.testing.assert_allclose(mo.densify(bL, r["phi"], nsf), d["phi"], atol=1e-11)


@pytest.mark.parametrize("roots", [(-0.3, 0.1, 2.0), (0.05, -1.0, 3.0), (-2.0, -0.5, 0.7)])
def test_solve_cubic_three_real_roots(roots):
    a, b, c = np.poly(roots)[1:]
    z = lnv.solve_cubic(a, b, c, 1.0)
    assert abs(z - min(roots, key=abs)) < 1e-12


def test_solve_cubic_one_real_root():
    # (x - 0.2)(x^2 + x + 3): one real root 0.2
    a, b, c = np.polymul([1, -0.2], [1, 1, 3])[1:]
    assert abs(lnv.solve_cubic(a, b, c, 0.1) - 0.2) < 1e-12
    # a root far beyond ten times the linear guess is replaced by the guess
    a, b, c = np.polymul([1, -50.0], [1, 1, 3])[1:]
    assert lnv.solve_cubic(a, b, c, 0.1) == 0.1
    r = [x.real for x in np.roots([1, a, b, c]) if abs(x.imag) < 1e-9]
    assert abs(lnv.solve_cubic(a, b, c, 10.0) - r[0]) < 1e-9


@pytest.mark.parametrize("n", [1, 3, 5])
def test_do_pulay_is_constrained_minimiser(n):
    rng = np.random.default_rng(n)
    R = rng.normal(size=(n, 40))
    A = R @ R.T
    al = lnv.do_pulay(n, A)
    # Lagrange: [A 1; 1^T 0] [a; mu] = [0; 1]
    M = np.zeros((n + 1, n + 1))
    M[:n, :n] = A
    M[:n, n] = M[n, :n] = 1.0
    sol = np.linalg.solve(M, np.r_[np.zeros(n), 1.0])
    np.testing.assert_allclose(al, sol[:n], atol=1e-10)
    assert abs(al.sum() - 1.0) < 1e-12
