"""CPU tests that pin the InitMcW restatement (oracle.mcweeny.init_mcw_dense).

InitMcW (src/McWeeny.f90:340-570, McWRho0 :961-1000) builds the starting
density matrix L0 = ((lambda*mubar + Ne)/No) T - (lambda/No) T H T from the
Gershgorin-type limits of H, choosing mubar between mu1 and mu2 by the test at
:487-505. McWeeny converges from many starting points, so convergence alone
does not pin this; what does:

  * hand-worked known answers (three levels, S = I, T = t I) for each branch,
    with T = S^-1 and with T != S^-1 (so the A = 1 - Tr[TS]/No terms of mu1 and
    mu2 are exercised), and the symmetric two-level case for which neither
    branch's strict inequality holds (the reference stops there);
  * with T = S^-1 exactly, Tr[L0 S] = Ne for any lambda, and the S-metric
    occupancies of L0 are the linear map f(e) = (lambda (mubar - e) + Ne)/No of
    the generalised eigenvalues e of H c = e S c;
  * on the mu1 branch f(hmax) = 0, on the mu2 branch f(hmin) = 1, and all
    occupancies lie in [0, 1] (the spectrum lies inside [hmin, hmax]).

A swapped branch, a wrong lambda denominator, a sign slip in mu2 or a
transposed L0 term fails at least one of these.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import mcweeny as mo


def levels(h, t):
    """Three (or two) levels on one atom: H diagonal, S = I, T = t I."""
    n = len(h)
    return np.diag(np.asarray(h, float)), np.eye(n), t * np.eye(n), np.array([n])


# (levels, Ne, t) -> (branch, mubar, lambda, diag(L0)), worked by hand:
#   No = 3; SX = 3t; A = 1 - t; SXHX = t^2 sum(h); hmin/hmax = min/max |h|
#   (GetHLimits on a diagonal onsite block).
KNOWN = [
    # h = (1,2,6), Ne = 1, t = 1: A = 0, mu1 = mu2 = 9/3 = 3;
    # 1/(6-3) = 1/3 < 2/(3-1) = 1 -> mu1, lambda = 1/3;
    # L0 = ((1 + 1)/3) - (1/9) h = (5/9, 4/9, 0).
    ((1.0, 2.0, 6.0), 1.0, 1.0, 1, 3.0, 1.0 / 3.0, (5.0 / 9.0, 4.0 / 9.0, 0.0)),
    # h = (1,2,6), Ne = 1, t = 1/2: A = 1/2, SXHX = 9/4; mu1 = 3/4 + 3 = 15/4;
    # 1/(6-15/4) = 4/9 < 2/(15/4-1) = 8/11 -> mu1, lambda = 4/9;
    # L0 = ((4/9*15/4 + 1)/3)(1/2) - (4/27)(1/4) h = 4/9 - h/27 = (11/27, 10/27, 2/9).
    ((1.0, 2.0, 6.0), 1.0, 0.5, 1, 15.0 / 4.0, 4.0 / 9.0, (11.0 / 27.0, 10.0 / 27.0, 2.0 / 9.0)),
    # h = (1,5,6), Ne = 2, t = 1: A = 0, mu1 = mu2 = 4; 2/(6-4) = 1 < 1/(4-1)? no;
    # 2/(6-4) = 1 > 1/3 -> mu2, lambda = 1/(4-1) = 1/3;
    # L0 = ((4/3 + 2)/3) - (1/9) h = (1, 5/9, 4/9).
    ((1.0, 5.0, 6.0), 2.0, 1.0, 2, 4.0, 1.0 / 3.0, (1.0, 5.0 / 9.0, 4.0 / 9.0)),
    # h = (1,5,6), Ne = 2, t = 4/5: A = 1/5, SXHX = 192/25 (/3 = 64/25);
    # mu1 = 64/25 + 6/5 = 94/25: 2/(56/25) = 25/28 < 1/(69/25) = 25/69? no;
    # mu2 = (1*(1/5)*2/1 - 64/25)/(3/5 - 1) = (-54/25)/(-2/5) = 27/5:
    # 2/(3/5) = 10/3 > 1/(22/5) = 5/22 -> mu2, lambda = 5/22;
    # L0 = ((27/22 + 2)/3)(4/5) - (5/66)(16/25) h = 142/165 - (8/165) h = (134/165, 34/55, 94/165).
    ((1.0, 5.0, 6.0), 2.0, 0.8, 2, 27.0 / 5.0, 5.0 / 22.0, (134.0 / 165.0, 34.0 / 55.0, 94.0 / 165.0)),
]


@pytest.mark.parametrize("h,ne,t,branch,mubar_ref,lam_ref,l0_ref", KNOWN)
def test_init_mcw_known_answers(h, ne, t, branch, mubar_ref, lam_ref, l0_ref):
    H, S, T, nsf = levels(h, t)
    hmin, hmax = mo.hlimits_dense_blocks(H, nsf)
    assert (hmin, hmax) == (min(h), max(h))
    L0, mubar, lam = mo.init_mcw_dense(H, S, T, ne, nsf)
    assert mubar == pytest.approx(mubar_ref, abs=1e-14)
    assert lam == pytest.approx(lam_ref, abs=1e-14)
    np.testing.assert_allclose(np.diag(L0), l0_ref, atol=1e-14)
    assert np.abs(L0 - np.diag(np.diag(L0))).max() == 0.0
    # The branch the answer came from: mu1 empties the top level, mu2 fills the bottom one.
    if branch == 1:
        assert (lam * (mubar - hmax) + ne) / 3.0 == pytest.approx(0.0, abs=1e-14)
    else:
        assert (lam * (mubar - hmin) + ne) / 3.0 == pytest.approx(1.0, abs=1e-14)


def test_init_mcw_symmetric_two_levels_has_no_mubar():
    # h = (1,3), Ne = 1, T = S^-1: mu1 = mu2 = 2 and both tests are equalities
    # (1/(3-2) = 1/(2-1)), so neither strict inequality holds.
    H, S, T, nsf = levels((1.0, 3.0), 1.0)
    with pytest.raises(RuntimeError, match="mubar"):
        mo.init_mcw_dense(H, S, T, 1.0, nsf)


def random_system(n_atoms, nsf, seed, gap_split):
    rng = np.random.default_rng(seed)
    n = n_atoms * nsf
    A = rng.normal(size=(n, n)) * 0.08
    S = np.eye(n) + 0.5 * (A + A.T)
    B = rng.normal(size=(n, n)) * 0.1
    d = np.concatenate([rng.uniform(-0.5, 0.5, n // 2) - gap_split, rng.uniform(-0.5, 0.5, n - n // 2) + gap_split])
    H = 0.5 * (B + B.T) + np.diag(d) + 3.0 * S   # onsite diagonal kept positive, as the workload recipe does
    return H, S, np.full(n_atoms, nsf)


@pytest.mark.parametrize("n_atoms,nsf,seed,ne_frac", [(6, 4, 0, 0.5), (5, 5, 1, 0.3), (4, 6, 2, 0.7)])
def test_init_mcw_with_exact_inverse_traces_ne_and_maps_eigenvalues_linearly(n_atoms, nsf, seed, ne_frac):
    H, S, nsfv = random_system(n_atoms, nsf, seed, 1.5)
    n = H.shape[0]
    ne = float(round(ne_frac * n))
    T = np.linalg.inv(S)
    L0, mubar, lam = mo.init_mcw_dense(H, S, T, ne, nsfv)
    # Tr[L0 S] = (lambda*mubar + Ne) Tr[TS]/No - lambda Tr[THTS]/No = Ne when T = S^-1.
    assert abs(np.trace(L0 @ S) - ne) < 1e-11 * n
    # Occupancies in the S metric: eigenvalues of S^1/2 L0 S^1/2 = f(e).
    e = scipy.linalg.eigh(H, S, eigvals_only=True)
    Sh = scipy.linalg.sqrtm(S).real
    occ = np.sort(np.linalg.eigvalsh(Sh @ L0 @ Sh))
    f = np.sort((lam * (mubar - e) + ne) / n)
    np.testing.assert_allclose(occ, f, atol=1e-11)
    hmin, hmax = mo.hlimits_dense_blocks(H, nsfv)
    assert hmin <= e.min() and e.max() <= hmax
    assert occ.min() >= -1e-12 and occ.max() <= 1.0 + 1e-12
    f_hmax, f_hmin = (lam * (mubar - hmax) + ne) / n, (lam * (mubar - hmin) + ne) / n
    # Exactly one end of [hmin, hmax] is pinned: 0 at hmax (mu1) or 1 at hmin (mu2).
    assert (abs(f_hmax) < 1e-12) != (abs(f_hmin - 1.0) < 1e-12)
    assert -1e-12 <= f_hmax and f_hmin <= 1.0 + 1e-12


def test_init_mcw_takes_both_branches_over_the_filling():
    H, S, nsfv = random_system(6, 4, 3, 1.5)
    n = H.shape[0]
    T = np.linalg.inv(S)
    hmin, hmax = mo.hlimits_dense_blocks(H, nsfv)
    seen = set()
    for ne in (2.0, 6.0, 12.0, 18.0, 22.0):
        L0, mubar, lam = mo.init_mcw_dense(H, S, T, ne, nsfv)
        if abs(lam * (mubar - hmax) + ne) < 1e-10:
            seen.add(1)
            assert lam == pytest.approx(ne / (hmax - mubar), rel=1e-14)
        else:
            seen.add(2)
            assert lam == pytest.approx((n - ne) / (mubar - hmin), rel=1e-14)
    assert seen == {1, 2}
