"""CPU tests that pin the spin-polarised McWeeny oracle (oracle.mcweeny.Sparse.mcwmin_spin).

McWMin with two spin channels sharing S (src/McWeeny.f90:144-256): with fixed
spin populations each spin takes McW_matrix_multiply_nonds steps and the loop
ends when both spins' stop tests have fired; with variable spin populations
McW_matrix_multiply_ds (:772-890) uses one c = (c1_up + c1_dn)/(c2_up + c2_dn)
and one stop test on the total energy. Pinned by:
  * identical channels reduce to the spin-unpolarised loop (bit for bit);
  * on an isolated cluster (no range truncates anything) the fixed-spin loop
    converges to each spin's S-metric projector onto its lowest N_s states, and
    the variable-spin loop to the projector onto the N_up + N_dn states of both
    spins with the largest starting occupations (the common polynomial map keeps
    the order of the eigenvalues of S^1/2 L S^1/2 across the two channels and the
    total trace), which with InitMcW's start is the lowest N_up + N_dn levels of
    the two spectra together when both spins share mubar and lambda;
  * a swapped c combination (per-spin c in the variable case) or a missing
    update of a finished spin changes these results.
"""
import numpy as np
import scipy.linalg

from cases import cluster_structure
from cqmm.dm import Purifier
from oracle import mcweeny as mo


def cluster(natoms=8, nsf=4, seed=3):
    xyz, box = cluster_structure(natoms, seed=seed)
    return Purifier(xyz, box, nsf, r_S=8.0, r_H=8.0, r_L=12.0, grid=(5, 5, 5), seed=seed, device=False)


def setup(natoms=8, seed=3, shift=0.25):
    pu = cluster(natoms=natoms, seed=seed)
    s, R = pu.sys, pu.rid
    sp = mo.Sparse(s, R)
    nsf = pu.nsf
    bS, bH, bL = s.blocks(R["S"]), s.blocks(R["H"]), s.blocks(R["L"])
    S = mo.densify(bS, pu.S_vals, nsf)
    Hu = mo.densify(bH, pu.H_vals, nsf)
    Hd = Hu + shift * S  # spin down: the same states, every level `shift` higher
    Hd_vals = mo.sparsify(Hd, bH, nsf, s.length(R["H"]))
    T = np.linalg.inv(S)
    return pu, sp, S, Hu, Hd, Hd_vals, T, bL


def l0(H, S, T, ne, pu, bL):
    L0, mubar, lam = mo.init_mcw_dense(H, S, T, ne, pu.nsf)
    return mo.sparsify(L0, bL, pu.nsf, pu.sys.length(pu.rid["L"])), mubar, lam


def test_identical_channels_reduce_to_the_unpolarised_loop():
    pu, sp, S, Hu, Hd, Hd_vals, T, bL = setup()
    ne = pu.n_orb / 2
    L0, _, _ = l0(Hu, S, T, ne, pu, bL)
    L1, E1, Ne1, n1 = sp.mcwmin(L0, pu.S_vals, pu.H_vals, 100)
    Ls, E, Ne, n = sp.mcwmin_spin([L0, L0], pu.S_vals, [pu.H_vals, pu.H_vals], 100, fixed=True)
    assert n == n1 and E == [E1, E1] and Ne == [Ne1, Ne1]
    np.testing.assert_array_equal(Ls[0], L1)
    np.testing.assert_array_equal(Ls[1], L1)
    # Variable spin: c1 and c2 double, so c and every step are the same while
    # no stop test fires (its threshold then applies to twice the energy change).
    for k in (1, 3, 5):
        Lk, _, _, _ = sp.mcwmin(L0, pu.S_vals, pu.H_vals, k)
        Ls, E, Ne, n = sp.mcwmin_spin([L0, L0], pu.S_vals, [pu.H_vals, pu.H_vals], k, fixed=False)
        if k >= 3:
            np.testing.assert_array_equal(Ls[0], Lk)
            np.testing.assert_array_equal(Ls[1], Lk)


def occupied_projectors(Xs, S, n_tot):
    """Projectors (S metric) onto the n_tot states with the largest eigenvalues
    of S^1/2 L S^1/2 across the spin channels."""
    Sh = scipy.linalg.sqrtm(S).real
    Shi = np.linalg.inv(Sh)
    cand = []
    vecs = []
    for s, X in enumerate(Xs):
        x, V = np.linalg.eigh(Sh @ X @ Sh)
        vecs.append(V)
        cand += [(x[i], s, i) for i in range(len(x))]
    cand.sort(key=lambda t: -t[0])
    occ = cand[:n_tot]
    Ps = []
    for s in range(len(Xs)):
        idx = [i for (_, ss, i) in occ if ss == s]
        V = vecs[s][:, idx]
        Ps.append(Shi @ V @ V.T @ Shi)
    return Ps, cand[n_tot - 1][0] - cand[n_tot][0]


def test_fixed_spin_converges_to_each_spin_projector():
    pu, sp, S, Hu, Hd, Hd_vals, T, bL = setup()
    no = pu.n_orb
    ne_u, ne_d = no // 2 + 2, no // 2 - 3
    Lu, _, _ = l0(Hu, S, T, ne_u, pu, bL)
    Ld, _, _ = l0(Hd, S, T, ne_d, pu, bL)
    Ls, E, Ne, n = sp.mcwmin_spin([Lu, Ld], pu.S_vals, [pu.H_vals, Hd_vals], 200, fixed=True)
    Pu, _ = mo.projector(Hu, S, ne_u)
    Pd, _ = mo.projector(Hd, S, ne_d)
    np.testing.assert_allclose(mo.densify(bL, Ls[0], pu.nsf), Pu, atol=2e-5)
    np.testing.assert_allclose(mo.densify(bL, Ls[1], pu.nsf), Pd, atol=2e-5)
    assert abs(Ne[0] - ne_u) < 1e-6 and abs(Ne[1] - ne_d) < 1e-6
    assert abs(E[0] - np.sum(Pu * Hu)) < 1e-6 and abs(E[1] - np.sum(Pd * Hd)) < 1e-6


def test_variable_spin_converges_to_the_common_occupation():
    """Per-spin InitMcW starts (each spin's own mubar, lambda): the common map
    keeps the total and orders the states by their starting occupation."""
    pu, sp, S, Hu, Hd, Hd_vals, T, bL = setup()
    no = pu.n_orb
    ne_u, ne_d = no // 2 + 2, no // 2 - 2
    Lu, _, _ = l0(Hu, S, T, ne_u, pu, bL)
    Ld, _, _ = l0(Hd, S, T, ne_d, pu, bL)
    X0 = [mo.densify(bL, Lu, pu.nsf), mo.densify(bL, Ld, pu.nsf)]
    Ps, gap = occupied_projectors(X0, S, ne_u + ne_d)
    assert gap > 1e-3
    Ls, E, Ne, n = sp.mcwmin_spin([Lu, Ld], pu.S_vals, [pu.H_vals, Hd_vals], 300, fixed=False)
    got = [mo.densify(bL, L, pu.nsf) for L in Ls]
    np.testing.assert_allclose(got[0], Ps[0], atol=2e-5)
    np.testing.assert_allclose(got[1], Ps[1], atol=2e-5)
    assert abs(Ne[0] + Ne[1] - (ne_u + ne_d)) < 1e-6


def test_variable_spin_init_and_loop_give_the_lowest_levels_of_both_spins():
    """The direct-sum InitMcW (one mubar and lambda) makes the starting
    occupation one decreasing function of the level energy for both spins, so
    the variable-spin loop ends on the lowest ne_up + ne_dn levels of the two
    spectra together: with spin down shifted up by more than the gap, spin up
    takes electrons from spin down."""
    pu, sp, S, Hu, Hd, Hd_vals, T, bL = setup(shift=4.0)
    no = pu.n_orb
    ne_u, ne_d = no // 2, no // 2
    L0s, mubar, lam = mo.init_mcw_dense_spin([Hu, Hd], S, T, (ne_u, ne_d), pu.nsf)
    # Tr[L0_up S] + Tr[L0_dn S] = n_e_ds with T = S^-1.
    assert abs(sum(np.trace(L @ S) for L in L0s) - (ne_u + ne_d)) < 1e-9
    e_u = scipy.linalg.eigh(Hu, S, eigvals_only=True)
    e_d = scipy.linalg.eigh(Hd, S, eigvals_only=True)
    allv = np.sort(np.concatenate([e_u, e_d]))
    n_tot = ne_u + ne_d
    occ_u = int(np.sum(e_u <= allv[n_tot - 1]))
    assert occ_u > ne_u  # the shift moves electrons to spin up
    Lu = mo.sparsify(L0s[0], bL, pu.nsf, pu.sys.length(pu.rid["L"]))
    Ld = mo.sparsify(L0s[1], bL, pu.nsf, pu.sys.length(pu.rid["L"]))
    Ls, E, Ne, n = sp.mcwmin_spin([Lu, Ld], pu.S_vals, [pu.H_vals, Hd_vals], 400, fixed=False)
    Pu, _ = mo.projector(Hu, S, occ_u)
    Pd, _ = mo.projector(Hd, S, n_tot - occ_u)
    np.testing.assert_allclose(mo.densify(bL, Ls[0], pu.nsf), Pu, atol=2e-5)
    np.testing.assert_allclose(mo.densify(bL, Ls[1], pu.nsf), Pd, atol=2e-5)
    assert abs(Ne[0] - occ_u) < 1e-6 and abs(Ne[1] - (n_tot - occ_u)) < 1e-6
    # The fixed-spin rule (per-spin c) would keep each spin's population.
    Lf, Ef, Nef, nf = sp.mcwmin_spin([Lu, Ld], pu.S_vals, [pu.H_vals, Hd_vals], 400, fixed=True)
    assert abs(Nef[0] + Nef[1] - n_tot) < 1e-6 and abs(Nef[0] - occ_u) > 0.5


def test_variable_spin_init_known_answer():
    """Two levels per spin on one atom (S = I, T = I): hand-worked direct-sum
    values. H_up = diag(1, 3), H_dn = diag(2, 6), ne = (1, 1): n_o = 2,
    n_o_ds = 4, n_e_ds = 2, SX = 2, A_ds = 0, SXHX_ds = 12, hmin_ds = 1,
    hmax_ds = 6; mu1 = mu2 = 12/4 = 3; n_e/(hmax - mu1) = 2/3 < 2/(3 - 1) = 1,
    so mubar = 3, lambda = 2/3 and L0_s = ((2 + 2)/4) I - (1/6) H_s:
    diag(5/6, 1/2) and diag(2/3, 0)."""
    S = np.eye(2)
    L0s, mubar, lam = mo.init_mcw_dense_spin([np.diag([1.0, 3.0]), np.diag([2.0, 6.0])], S, np.eye(2), (1, 1),
                                             np.array([2]))
    assert abs(mubar - 3.0) < 1e-14 and abs(lam - 2.0 / 3.0) < 1e-14
    np.testing.assert_allclose(np.diag(L0s[0]), [5 / 6, 1 / 2], atol=1e-14)
    np.testing.assert_allclose(np.diag(L0s[1]), [2 / 3, 0.0], atol=1e-14)


def test_spin_dependent_overlaps():
    """S given per spin: the same S twice is the shared-S loop bit for bit; with
    a different (well-conditioned) overlap for spin down, fixed populations
    converge to each spin's projector in its own metric."""
    pu, sp, S, Hu, Hd, Hd_vals, T, bL = setup()
    ne = pu.n_orb // 2
    L0, _, _ = l0(Hu, S, T, ne, pu, bL)
    a = sp.mcwmin_spin([L0, L0], pu.S_vals, [pu.H_vals, Hd_vals], 60, fixed=True)
    b = sp.mcwmin_spin([L0, L0], (pu.S_vals, pu.S_vals), [pu.H_vals, Hd_vals], 60, fixed=True)
    assert a[3] == b[3] and a[1] == b[1]
    np.testing.assert_array_equal(a[0][1], b[0][1])
    # Spin down with S' = S + 0.02 (S - I) (same pattern, still positive definite).
    bS = pu.sys.blocks(pu.rid["S"])
    S2 = S + 0.02 * (S - np.eye(S.shape[0]))
    S2_vals = mo.sparsify(S2, bS, pu.nsf, pu.sys.length(pu.rid["S"]))
    L0d, _, _ = mo.init_mcw_dense(Hd, S2, np.linalg.inv(S2), ne, pu.nsf)
    L0d = mo.sparsify(L0d, bL, pu.nsf, pu.sys.length(pu.rid["L"]))
    Ls, E, Ne, n = sp.mcwmin_spin([L0, L0d], (pu.S_vals, S2_vals), [pu.H_vals, Hd_vals], 300, fixed=True)
    Pu, _ = mo.projector(Hu, S, ne)
    Pd, _ = mo.projector(Hd, S2, ne)
    np.testing.assert_allclose(mo.densify(bL, Ls[0], pu.nsf), Pu, atol=2e-5)
    np.testing.assert_allclose(mo.densify(bL, Ls[1], pu.nsf), Pd, atol=2e-5)
    assert abs(Ne[1] - ne) < 1e-6
