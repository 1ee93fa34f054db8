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
