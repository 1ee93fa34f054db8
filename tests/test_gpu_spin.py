"""GPU tests of the spin-polarised McWeeny path (cqmm_init_mcweeny_spin,
cqmm_mcweeny_spin) against the oracle.

Bars: InitMcW for both spin channels matches the dense oracle to 1e-12 (per-spin
for fixed populations, the direct sum for variable ones); the McWMin loop over
both channels matches the oracle's two-channel loop (stop step within one,
energies to 1e-8 relative, L to 1e-8); identical channels give the
spin-unpolarised device loop bit for bit (fixed populations: every step;
variable populations: c1 and c2 both double, so c is the same); the same
against the oracle with symmetric-result plans for LSL and LSLSL; with spin down
shifted above the gap the variable-population loop moves electrons to spin up
and ends on the lowest levels of both spectra; spin-dependent overlaps (one S
per spin) match the oracle's per-spin-metric loop; mixed precision on a
variable-population pair stays within 1e-6 Ha/atom of the FP64 oracle.
"""
import numpy as np
import pytest

from cqmm import CqmmError
from cqmm.dm import Purifier
from cases import cluster_structure
from oracle import mcweeny as mo

pytestmark = pytest.mark.gpu


def spin_cluster(natoms=8, nsf=4, seed=3, shift=0.25, symmetric=False):
    xyz, box = cluster_structure(natoms, seed=seed)
    pu = Purifier(xyz, box, nsf, r_S=8.0, r_H=8.0, r_L=12.0, grid=(5, 5, 5), seed=seed, symmetric=symmetric)
    s, R = pu.sys, pu.rid
    bS, bH = s.blocks(R["S"]), s.blocks(R["H"])
    S = mo.densify(bS, pu.S_vals, pu.nsf)
    Hu = mo.densify(bH, pu.H_vals, pu.nsf)
    Hd = Hu + shift * S  # spin down: the same states, every level `shift` higher
    pu.add_spin_down(mo.sparsify(Hd, bH, pu.nsf, s.length(R["H"])))
    pu.hotelling(max_iter=50, tol=1e-10)
    return pu, S, Hu, Hd


def dense(pu, name, vals):
    return mo.densify(pu.sys.blocks(pu.rid[name]), vals, pu.nsf)


@pytest.mark.parametrize("fixed", [True, False], ids=["fixed", "variable"])
def test_spin_init_matches_dense_oracle(fixed):
    pu, S, Hu, Hd = spin_cluster(shift=4.0)
    T = dense(pu, "T", pu.ctx.get(pu.matT))
    ne_up, ne_dn = pu.n_orb // 2 + 1, pu.n_orb // 2 - 1
    print(f"spin InitMcW ({'fixed' if fixed else 'variable'}): {pu.natoms} atoms, ne {ne_up}/{ne_dn}", flush=True)
    (mu_u, la_u), (mu_d, la_d) = pu.init_mcweeny_spin(ne_up, ne_dn, fixed)
    L0u = dense(pu, "L", pu.ctx.get(pu.matL))
    L0d = dense(pu, "L", pu.ctx.get(pu.matL_dn))
    if fixed:
        ref_u, mu_ru, la_ru = mo.init_mcw_dense(Hu, S, T, ne_up, pu.nsf)
        ref_d, mu_rd, la_rd = mo.init_mcw_dense(Hd, S, T, ne_dn, pu.nsf)
    else:
        (ref_u, ref_d), mu_ru, la_ru = mo.init_mcw_dense_spin([Hu, Hd], S, T, (ne_up, ne_dn), pu.nsf)
        mu_rd, la_rd = mu_ru, la_ru
        assert mu_u == mu_d and la_u == la_d
    for got, ref in ((mu_u, mu_ru), (la_u, la_ru), (mu_d, mu_rd), (la_d, la_rd)):
        assert abs(got - ref) < 1e-10 * max(1.0, abs(ref))
    np.testing.assert_allclose(L0u, ref_u, atol=1e-12)
    np.testing.assert_allclose(L0d, ref_d, atol=1e-12)


@pytest.mark.parametrize("symmetric", [False, True], ids=["full", "symmetric_plans"])
@pytest.mark.parametrize("fixed", [True, False], ids=["fixed", "variable"])
def test_spin_mcweeny_matches_sparse_oracle(fixed, symmetric):
    pu, S, Hu, Hd = spin_cluster(shift=4.0, symmetric=symmetric)
    ne = pu.n_orb // 2
    pu.init_mcweeny_spin(ne, ne, fixed)
    L0 = [pu.ctx.get(pu.matL), pu.ctx.get(pu.matL_dn)]
    print(f"spin McWeeny ({'fixed' if fixed else 'variable'}) vs oracle: {pu.natoms} atoms", flush=True)
    res = pu.mcweeny_spin(fixed, precision=0, max_iter=200)
    L = [pu.ctx.get(pu.matL), pu.ctx.get(pu.matL_dn)]
    sp = mo.Sparse(pu.sys, pu.rid)
    L_ref, E_ref, Ne_ref, n_ref = sp.mcwmin_spin(L0, pu.S_vals, [pu.H_vals, pu.H_dn_vals], 200, fixed=fixed)
    assert abs(res[0]["iterations"] - n_ref) <= 1 and res[0]["iterations"] == res[1]["iterations"]
    for k in range(2):
        assert abs(res[k]["energy"] - E_ref[k]) < 1e-8 * max(1.0, abs(E_ref[k]))
        assert abs(res[k]["electrons"] - Ne_ref[k]) < 1e-8
        assert np.abs(L[k] - L_ref[k]).max() < 1e-8
    if fixed:
        assert abs(res[0]["electrons"] - ne) < 1e-6 and abs(res[1]["electrons"] - ne) < 1e-6
        assert res[0]["stop_reason"] in (1, 2, 3) and res[1]["stop_reason"] in (1, 2, 3)
    else:
        # Spin down lies 4 Ha higher than spin up (above the gap): the lowest
        # 2 ne levels of both spectra together hold more spin-up states.
        occ_up = round(res[0]["electrons"])
        assert occ_up > ne and abs(res[0]["electrons"] + res[1]["electrons"] - 2 * ne) < 1e-6
        Pu, _ = mo.projector(Hu, S, occ_up)
        Pd, _ = mo.projector(Hd, S, 2 * ne - occ_up)
        assert np.abs(dense(pu, "L", L[0]) - Pu).max() < 1e-6
        assert np.abs(dense(pu, "L", L[1]) - Pd).max() < 1e-6


@pytest.mark.parametrize("fixed", [True, False], ids=["fixed", "variable"])
def test_identical_spin_channels_match_the_unpolarised_device_loop(fixed):
    """H_dn = H and L_dn = L: fixed populations repeat the spin-unpolarised loop
    in each channel; variable populations double c1 and c2 (exact), so each of
    the first steps is the unpolarised step, bit for bit."""
    pu, S, Hu, Hd = spin_cluster(shift=0.0)
    ne = pu.n_orb // 2
    pu.init_mcweeny(ne)
    L0 = pu.ctx.get(pu.matL)
    max_iter = 200 if fixed else 4
    single = pu.mcweeny(precision=0, max_iter=max_iter)
    L1 = pu.ctx.get(pu.matL)
    pu.ctx.set(pu.matL, L0)
    pu.ctx.set(pu.matL_dn, L0)
    print(f"identical channels ({'fixed' if fixed else 'variable'}), {single['iterations']} steps", flush=True)
    res = pu.mcweeny_spin(fixed, precision=0, max_iter=max_iter)
    if fixed or single["iterations"] == max_iter:
        assert res[0]["iterations"] == single["iterations"]
        np.testing.assert_array_equal(pu.ctx.get(pu.matL), L1)
        np.testing.assert_array_equal(pu.ctx.get(pu.matL_dn), L1)
        assert res[0]["energy"] == single["energy"] and res[1]["energy"] == single["energy"]


def test_spin_dependent_overlaps_against_oracle():
    """Spin down with its own overlap S' (matS[1] != matS[0], spin-dependent
    support functions): fixed populations against the oracle's two-channel
    McWMin with per-spin overlaps."""
    pu, S, Hu, Hd = spin_cluster(shift=4.0)
    ne = pu.n_orb // 2
    bS, bL = pu.sys.blocks(pu.rid["S"]), pu.sys.blocks(pu.rid["L"])
    S2 = S + 0.02 * (S - np.eye(S.shape[0]))
    S2_vals = mo.sparsify(S2, bS, pu.nsf, pu.sys.length(pu.rid["S"]))
    matS2 = pu.ctx.allocate(pu.rid["S"])
    pu.ctx.set(matS2, S2_vals)
    T = dense(pu, "T", pu.ctx.get(pu.matT))
    L0u, _, _ = mo.init_mcw_dense(Hu, S, T, ne, pu.nsf)
    L0d, _, _ = mo.init_mcw_dense(Hd, S2, np.linalg.inv(S2), ne, pu.nsf)
    L0 = [mo.sparsify(L0u, bL, pu.nsf, pu.sys.length(pu.rid["L"])),
          mo.sparsify(L0d, bL, pu.nsf, pu.sys.length(pu.rid["L"]))]
    pu.ctx.set(pu.matL, L0[0])
    pu.ctx.set(pu.matL_dn, L0[1])
    print(f"spin-dependent overlaps: {pu.natoms} atoms", flush=True)
    res = pu.mcweeny_spin(True, precision=0, max_iter=300, matS_dn=matS2)
    sp = mo.Sparse(pu.sys, pu.rid)
    L_ref, E_ref, Ne_ref, n_ref = sp.mcwmin_spin(L0, (pu.S_vals, S2_vals), [pu.H_vals, pu.H_dn_vals], 300,
                                                 fixed=True)
    assert abs(res[0]["iterations"] - n_ref) <= 1
    for k, h in enumerate((pu.matL, pu.matL_dn)):
        assert abs(res[k]["energy"] - E_ref[k]) < 1e-8 * max(1.0, abs(E_ref[k]))
        assert abs(res[k]["electrons"] - Ne_ref[k]) < 1e-8
        assert np.abs(pu.ctx.get(h) - L_ref[k]).max() < 1e-8


def test_variable_spin_mixed_precision_against_fp64_oracle():
    pu, S, Hu, Hd = spin_cluster(natoms=8, nsf=13, shift=0.25)
    ne = pu.n_orb // 2
    pu.init_mcweeny_spin(ne, ne, False)
    L0 = [pu.ctx.get(pu.matL), pu.ctx.get(pu.matL_dn)]
    print(f"variable spin, precision 1 vs FP64 oracle: {pu.natoms} atoms x 13", flush=True)
    res = pu.mcweeny_spin(False, precision=1, max_iter=200, fp32_tol=0.0)
    assert res[0]["fp32_iterations"] > 0 and res[0]["fp32_iterations"] == res[1]["fp32_iterations"]
    sp = mo.Sparse(pu.sys, pu.rid, ompgemm=True)
    L_ref, E_ref, Ne_ref, n_ref = sp.mcwmin_spin(L0, pu.S_vals, [pu.H_vals, pu.H_dn_vals], 200, fixed=False)
    dE = abs(res[0]["energy"] + res[1]["energy"] - E_ref[0] - E_ref[1]) / pu.natoms
    dN = abs(res[0]["electrons"] + res[1]["electrons"] - Ne_ref[0] - Ne_ref[1])
    print(f"dE/atom {dE:.2e}, dNe {dN:.2e}", flush=True)
    assert dE <= 1e-6 and dN <= 1e-8


def test_fixed_spin_pair_refuses_reduced_precision_and_keeps_the_stack():
    pu, S, Hu, Hd = spin_cluster(shift=0.25)
    ne = pu.n_orb // 2
    pu.init_mcweeny_spin(ne, ne, True)
    with pytest.raises(CqmmError):
        pu.mcweeny_spin(True, precision=1)
    # The temporary-matrix stack is as it was: a further allocation and free work.
    h = pu.ctx.allocate(pu.rid["L"])
    pu.ctx.free(h)
    res = pu.mcweeny_spin(True, precision=0)
    assert res[0]["iterations"] > 2
