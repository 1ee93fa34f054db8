"""CPU tests that pin the McWeeny / InitMcW / Hotelling oracle.

No reference fixture covers these loops (the reference's only check on this path
is the end-to-end test 002), so they are pinned by closed forms on isolated
clusters, where no range truncates any product:
  * one step is the cubic map of McWeeny on the eigenvalues of S^1/2 L S^1/2;
  * the converged density matrix is the S-metric projector of the lowest Ne
    eigenvectors of H c = e S c;
  * Hotelling converges to S^-1;
  * the sparse restatement (image-resolved blocks, oracle mat_mult, range maps)
    equals the dense one step for step.
"""
import numpy as np
import pytest
import scipy.linalg

import cqmm
from cqmm.dm import Purifier
from cases import cluster_structure
from oracle import mcweeny as mo


def random_spd(n, rng, scale=0.1):
    A = rng.normal(size=(n, n)) * scale
    return np.eye(n) + 0.5 * (A + A.T)


def test_dense_step_is_the_eigenvalue_polynomial():
    rng = np.random.default_rng(0)
    n = 30
    S = random_spd(n, rng)
    H = rng.normal(size=(n, n)); H = 0.5 * (H + H.T)
    # A near-idempotent L in the S metric with eigenvalues spread in (0,1).
    P, _ = mo.projector(H, S, n // 2)
    Sh = scipy.linalg.sqrtm(S).real
    Shi = np.linalg.inv(Sh)
    X0 = Sh @ P @ Sh
    Q = np.linalg.eigh(X0)[1]
    x = np.clip(np.linspace(-0.05, 1.05, n) + rng.normal(scale=0.01, size=n), -0.1, 1.1)
    X = Q @ np.diag(x) @ Q.T
    L = Shi @ X @ Shi
    rho, E, Ne, c = mo.dense_step(L, S, H)
    xn, c_ref = mo.step_polynomial(np.linalg.eigvalsh(X))
    assert abs(c - c_ref) < 1e-12
    got = np.sort(np.linalg.eigvalsh(Sh @ rho @ Sh))
    np.testing.assert_allclose(got, np.sort(xn), atol=1e-11)
    assert abs(Ne - np.trace(rho @ S)) < 1e-10


def test_dense_loop_converges_to_projector():
    rng = np.random.default_rng(1)
    n = 40
    S = random_spd(n, rng)
    d = np.concatenate([rng.uniform(-1.5, -0.6, n // 2), rng.uniform(0.6, 1.5, n - n // 2)])
    V = np.linalg.qr(rng.normal(size=(n, n)))[0]
    H = V @ np.diag(d) @ V.T
    T = mo.hotelling_dense(S)
    np.testing.assert_allclose(T, np.linalg.inv(S), atol=1e-9)
    nsf = np.full(10, 4)
    L0, mubar, lam = mo.init_mcw_dense(H, S, T, n / 2, nsf)
    L, E, Ne, it = mo.dense_mcwmin(L0, S, H, 200)
    P, _ = mo.projector(H, S, n // 2)
    assert np.abs(L - P).max() < 1e-6
    assert abs(Ne - n / 2) < 1e-6
    assert abs(E - np.sum(P * H)) < 1e-6


def cluster_purifier(natoms=9, nsf=4, seed=3):
    xyz, box = cluster_structure(natoms, seed=seed)
    return Purifier(xyz, box, nsf, r_S=8.0, r_H=8.0, r_L=12.0, grid=(5, 5, 5), seed=seed, device=False)


def test_sparse_step_equals_dense_step_on_cluster():
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
    rho, E, Ne, c = sp.step(L, pu.S_vals, pu.H_vals)
    rho_d, E_d, Ne_d, c_d = mo.dense_step(Ld, S, H)
    assert abs(c - c_d) < 1e-12
    assert abs(E - E_d) < 1e-10 * max(1, abs(E_d))
    assert abs(Ne - Ne_d) < 1e-10 * max(1, abs(Ne_d))
    np.testing.assert_allclose(mo.densify(bL, rho, nsf), rho_d, atol=1e-12)


def test_sparse_symmetrise_matches_dense():
    pu = cluster_purifier(natoms=7)
    s, R = pu.sys, pu.rid
    sp = mo.Sparse(s, R)
    bL = s.blocks(R["L"])
    v = np.random.default_rng(2).normal(size=s.length(R["L"]))
    M = mo.densify(bL, v, pu.nsf)
    got = mo.densify(bL, sp.symmetrise(v, "L"), pu.nsf)
    np.testing.assert_allclose(got, 0.5 * (M + M.T), atol=0)


def test_hlimits_known_answer():
    """GetHLimits (src/McWeeny.f90:1020-1109) on a two-atom H worked through by
    hand from the reference loop. For primary atom a and each neighbour block,
    HMin(i) / HMax(i) collect |H(j of a, i of the neighbour)| summed over j —
    i is the neighbour's orbital (return_matrix_value(..., j, i) is element
    (i-1)*nsf_a + j of the block, src/mult_module.f90:3016), added to both
    for the onsite block, subtracted from HMin otherwise — and the extremes are
    taken over i = 1..nsf_a only. Atoms a (2 orbitals), b (3); rows/columns
    [a1 a2 b1 b2 b3]; the blocks are not symmetric so that a transposed block
    or a row sum gives other numbers.
      atom a: onsite columns |1|+|0.2| = 1.2, |0.5|+|-2| = 2.5; block ab
              columns 0.1, 0.3, 0.7 -> HMin = [1.1, 2.2, (-0.7 not counted)],
              HMax = [1.3, 2.8] -> min 1.1, max 2.8
      atom b: block ba columns 0, 0.9 -> HMin = [-0, -0.9, .], onsite 3, 4, 5
              -> HMin = [3, 3.1, 5], HMax = [3, 4.9, 5] -> min 3, max 5
      result (1.1, 5.0); with row sums instead, atom a gives 0.7; with signed
      onsite values, 2.2 becomes -1.8; with the extremes over all neighbour
      orbitals, -0.7."""
    H = np.zeros((5, 5))
    H[:2, :2] = [[1.0, 0.5], [0.2, -2.0]]
    H[2:, 2:] = np.diag([3.0, 4.0, 5.0])
    H[:2, 2:] = [[0.1, 0.0, 0.7], [0.0, 0.3, 0.0]]
    H[2:, :2] = [[0.0, 0.0], [0.0, 0.0], [0.0, 0.9]]
    hmin, hmax = mo.hlimits_dense_blocks(H, [2, 3])
    assert hmin == pytest.approx(1.1, abs=1e-15) and hmax == pytest.approx(5.0, abs=1e-15)


def test_hlimits_diagonal_closed_form():
    """A diagonal H: every orbital's sums hold only |H_ii|, so the limits are
    the smallest and largest |H_ii| (the onsite block adds absolute values)."""
    d = np.array([2.5, -0.75, 1.0, 4.0, -3.0, 0.5])
    hmin, hmax = mo.hlimits_dense_blocks(np.diag(d), [2, 1, 3])
    assert (hmin, hmax) == (0.5, 4.0)
