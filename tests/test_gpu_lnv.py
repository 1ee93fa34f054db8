"""GPU tests of the LNV functional, the electron-number correction and the
FindMinDM loop (early + late stages) on device matrices.

Bars: the transpose is exact (a permutation); LNV's K, M3, phi, energy and
electron count match the sparse oracle (same range truncations) to 1e-11
relative on a periodic cell, and the dense closed form on an isolated cluster;
the correction brings Tr[SK] to the target within 1e-4 per spin (its own stop
test); FindMinDM on an isolated cluster converges to the eigen-projector of
(H, S): band energy within 1e-6 Ha/atom, electron count within 1e-6; on a
periodic cell with truncated ranges it follows the oracle's loop (same stage
iteration counts within one or two, energy within 1e-8 Ha/atom).
"""
import numpy as np
import pytest

import cqmm
from cqmm.dm import Purifier
from cases import SI8, cluster_structure
from oracle import dmm as od
from oracle import lnv
from oracle import mcweeny as mo

pytestmark = pytest.mark.gpu


def cluster(natoms=9, nsf=4, seed=3):
    xyz, box = cluster_structure(natoms, seed=seed)
    return Purifier(xyz, box, nsf, r_S=8.0, r_H=8.0, r_L=12.0, grid=(5, 5, 5), seed=seed, lnv=True)


def periodic(natoms=24, nsf=4, seed=5, symmetric=False):
    xyz, edge = cqmm.random_structure(natoms, 0.05, 2.0, seed)
    if isinstance(nsf, list):  # a species pattern repeated over the atoms (mixed block sizes)
        nsf = (nsf * natoms)[:natoms]
    return Purifier(xyz, edge, nsf, r_S=4.5, r_H=4.5, r_L=6.0, grid=(2, 2, 2), seed=seed, lnv=True,
                    symmetric=symmetric)


def dense_of(pu, name, vals):
    return mo.densify(pu.sys.blocks(pu.rid[name]), vals, pu.nsf)


def rel(a, b):
    return np.abs(a - b).max() / max(1.0, np.abs(b).max())


@pytest.mark.parametrize("name", ["L", "LS", "LSL", "T"])
def test_transpose_matches_oracle(name):
    pu = periodic(nsf=[4, 9, 1, 13])
    print(f"transpose {name}", flush=True)
    sp = mo.Sparse(pu.sys, pu.rid)
    c = pu.ctx
    a, b = c.allocate(pu.rid[name]), c.allocate(pu.rid[name])
    v = np.random.default_rng(1).normal(size=c.length(a))
    c.set(a, v)
    c.transpose(a, b)
    np.testing.assert_array_equal(c.get(b), lnv.transpose(sp, v, name))
    c.free(b)
    c.free(a)


def random_L(pu, seed, scale=0.05):
    """A symmetric L: half of S on the L pattern plus symmetrised noise."""
    rng = np.random.default_rng(seed)
    n_L = pu.sys.length(pu.rid["L"])
    v = rng.normal(size=n_L) * scale
    sp = mo.Sparse(pu.sys, pu.rid)
    v = sp.symmetrise(v, "L")
    onsite = pu._map_values(pu.rid["S"], pu.rid["L"], pu.S_vals)
    return 0.5 * onsite + v


@pytest.mark.parametrize("nsf", [4, 13, [4, 9, 13, 4], [4, 20]])
def test_lnv_matches_sparse_oracle_periodic(nsf):
    pu = periodic(nsf=nsf)
    print(f"lnv periodic nsf={nsf} natoms={pu.natoms}", flush=True)
    sp = mo.Sparse(pu.sys, pu.rid)
    L = random_L(pu, 7)
    pu.ctx.set(pu.matL, L)
    got = pu.lnv(15)
    ref = lnv.lnv_sparse(sp, L, pu.S_vals, pu.H_vals)
    for k in ("K", "M3", "phi"):
        assert rel(got[k], ref[k]) < 1e-11, k
    assert abs(got["energy"] - ref["energy"]) < 1e-11 * max(1.0, abs(ref["energy"]))
    assert abs(got["electrons"] - ref["electrons"]) < 1e-11 * max(1.0, abs(ref["electrons"]))
    # The phi flag alone gives the same electron count and phi.
    pu.ctx.set(pu.matL, L)
    only_phi = pu.lnv(8)
    assert abs(only_phi["electrons"] - got["electrons"]) < 1e-12 * abs(got["electrons"])
    np.testing.assert_array_equal(only_phi["phi"], got["phi"])


def test_lnv_matches_sparse_oracle_on_si8_reference_structure():
    """The structure and ranges of the reference's Order-N test 002 (bohr):
    L.S.L reaches 46 bohr, over four cells, on 8 atoms."""
    edge = SI8["cell_bohr"][0][0]
    xyz = np.array(SI8["frac"]) * edge
    pu = Purifier(xyz, edge, SI8["nsf"], r_S=SI8["s_range_bohr"], r_H=SI8["s_range_bohr"],
                  r_L=SI8["l_range_bohr"], grid=(2, 2, 2), seed=5, lnv=True)
    print("lnv si8 test-002 structure", flush=True)
    sp = mo.Sparse(pu.sys, pu.rid)
    L = random_L(pu, 7)
    pu.ctx.set(pu.matL, L)
    got = pu.lnv(15)
    ref = lnv.lnv_sparse(sp, L, pu.S_vals, pu.H_vals)
    for k in ("K", "M3", "phi"):
        assert rel(got[k], ref[k]) < 1e-11, k
    assert abs(got["energy"] - ref["energy"]) < 1e-11 * max(1.0, abs(ref["energy"]))
    assert abs(got["electrons"] - ref["electrons"]) < 1e-11 * max(1.0, abs(ref["electrons"]))


def test_lnv_matches_dense_closed_form_on_cluster():
    pu = cluster()
    print("lnv cluster", flush=True)
    S = dense_of(pu, "S", pu.S_vals)
    H = dense_of(pu, "H", pu.H_vals)
    rng = np.random.default_rng(5)
    n = S.shape[0]
    A = rng.normal(size=(n, n)) * 0.05
    Ld = 0.5 * np.linalg.inv(S) + 0.5 * (A + A.T)
    pu.ctx.set(pu.matL, mo.sparsify(Ld, pu.sys.blocks(pu.rid["L"]), pu.nsf, pu.sys.length(pu.rid["L"])))
    got = pu.lnv(15)
    d = lnv.lnv_dense(Ld, S, H)
    np.testing.assert_allclose(dense_of(pu, "H", got["K"]), d["K"], atol=1e-11)
    np.testing.assert_allclose(dense_of(pu, "L", got["M3"]), d["M3"], atol=1e-10)
    np.testing.assert_allclose(dense_of(pu, "L", got["phi"]), d["phi"], atol=1e-10)
    assert abs(got["energy"] - d["energy"]) < 1e-10 * max(1.0, abs(d["energy"]))
    assert abs(got["electrons"] - d["electrons"]) < 1e-10 * max(1.0, abs(d["electrons"]))


def test_correct_electrons_reaches_target():
    pu = periodic(nsf=4)
    print("correct_electrons", flush=True)
    pu.hotelling()
    pu.init_mcweeny()
    pu.ctx.scale(1.04, pu.matL)
    target = float(pu.n_orb // 2)
    before = pu.lnv(8)["electrons"]
    assert abs(before - target) > 0.1
    after = pu.correct_electrons(n_electrons=target)
    assert abs(after - target) < 1e-4
    assert abs(pu.lnv(8)["electrons"] - after) < 1e-12 * target


@pytest.mark.parametrize("nsf", [4, 13])
def test_findmindm_cluster_converges_to_projector(nsf):
    pu = cluster(natoms=9, nsf=nsf)
    print(f"findmindm cluster nsf={nsf}", flush=True)
    S = dense_of(pu, "S", pu.S_vals)
    H = dense_of(pu, "H", pu.H_vals)
    ne = pu.n_orb // 2
    P, e = mo.projector(H, S, ne)
    assert e[ne] - e[ne - 1] > 0.1
    pu.hotelling(max_iter=50, tol=1e-10)
    pu.init_mcweeny(float(ne))
    res = pu.dmm(n_electrons=float(ne), tolerance=1e-10, n_L_iterations=100)
    print(f"  early {res['early_iterations']} late {res['late_iterations']} inflexions {res['inflexions']} "
          f"residual {res['residual']:.3e} E {res['band_energy']:.12f} Ne {res['n_electrons']:.12f} "
          f"({res['seconds']:.2f} s)", flush=True)
    assert res["done"] == 1
    assert abs(res["band_energy"] - 2.0 * np.sum(P * H)) / pu.natoms < 1e-6
    assert abs(res["n_electrons"] - 2.0 * ne) < 1e-6
    K = dense_of(pu, "H", res["K"])
    assert np.abs(K - P).max() < 1e-4


@pytest.mark.parametrize("symmetric", [False, True], ids=["full", "symmetric_plans"])
def test_findmindm_periodic_matches_oracle(symmetric):
    """Truncated ranges on a periodic cell: the device loop and the oracle's,
    started from the same L and T, take the same path to the residual bar
    (also with symmetric-result plans for the symmetric products of McWeeny
    and LNV: LSL, LSLSL, SLS, SLSLS, which change results by rounding only)."""
    pu = periodic(natoms=24, nsf=4, symmetric=symmetric)
    print(f"findmindm periodic vs oracle (symmetric plans: {symmetric})", flush=True)
    ne = float(pu.n_orb // 2)
    pu.hotelling()
    pu.init_mcweeny(ne)
    pu.mcweeny(precision=0, max_iter=100)
    L0 = pu.ctx.get(pu.matL)
    T = pu.ctx.get(pu.matT)
    res = pu.dmm(n_electrons=ne, tolerance=1e-7, n_L_iterations=100)
    sp = mo.Sparse(pu.sys, pu.rid)
    d = od.Dmm(sp, pu.S_vals, pu.H_vals, T, ne, pu.natoms, tolerance=1e-7, n_L_iterations=100)
    ref = d.run(L0)
    print(f"  device: early {res['early_iterations']} late {res['late_iterations']} residual {res['residual']:.3e} "
          f"E {res['band_energy']:.12f} Ne {res['n_electrons']:.12f}\n"
          f"  oracle: early {ref['early_iterations']} late {ref['late_iterations']} residual {ref['residual']:.3e} "
          f"E {ref['energy']:.12f} Ne {ref['electrons']:.12f}", flush=True)
    assert res["done"] == 1 and ref["done"]
    assert res["residual"] < 1e-7
    assert abs(res["early_iterations"] - ref["early_iterations"]) <= 1
    assert abs(res["late_iterations"] - ref["late_iterations"]) <= 2
    # Both stop below the same residual; the energies agree to far below the
    # 1e-6 Ha/atom bar.
    assert abs(res["band_energy"] - ref["energy"]) / pu.natoms < 1e-8
    assert abs(res["n_electrons"] - ref["electrons"]) < 1e-8
    assert rel(res["K"], ref["K"]) < 1e-5


@pytest.mark.parametrize("precision,nsf", [(2, 4), (3, 13)])
def test_mixed_precision_findmindm_matches_fp64_oracle(precision, nsf):
    """FindMinDM with FP32 bulk products (FP32 MFMA; block sums in FP64 for
    precision 2, chunk runs in FP32 for 3) and FP64 electron-number correction
    and final K / 2Tr[KH] / N_e, on a periodic cell, against the oracle's
    all-FP64 FindMinDM from the same L and T: band energy within 1e-6 Ha/atom,
    electron count within 1e-8 of the oracle's, K within 1e-5."""
    natoms = 24 if nsf == 4 else 12
    pu = periodic(natoms=natoms, nsf=nsf)
    print(f"mixed findmindm (precision {precision}) vs oracle: {natoms} atoms x {nsf} SF", flush=True)
    ne = float(pu.n_orb // 2)
    pu.hotelling()
    pu.init_mcweeny(ne)
    pu.mcweeny(precision=0, max_iter=100)
    L0 = pu.ctx.get(pu.matL)
    T = pu.ctx.get(pu.matT)
    res = pu.dmm(product_precision=precision, n_electrons=ne, tolerance=1e-7, n_L_iterations=100)
    assert pu.ctx.product_precision == 0
    sp = mo.Sparse(pu.sys, pu.rid)
    d = od.Dmm(sp, pu.S_vals, pu.H_vals, T, ne, pu.natoms, tolerance=1e-7, n_L_iterations=100)
    ref = d.run(L0)
    dE = abs(res["band_energy"] - ref["energy"]) / pu.natoms
    dN = abs(res["n_electrons"] - ref["electrons"])
    print(f"  device mixed: early {res['early_iterations']} late {res['late_iterations']} residual "
          f"{res['residual']:.3e} E {res['band_energy']:.12f} Ne {res['n_electrons']:.12f}\n"
          f"  oracle FP64:  early {ref['early_iterations']} late {ref['late_iterations']} residual "
          f"{ref['residual']:.3e} E {ref['energy']:.12f} Ne {ref['electrons']:.12f}; dE/atom {dE:.2e} dNe {dN:.2e}",
          flush=True)
    assert res["done"] == 1 and ref["done"]
    assert dE <= 1e-6
    assert dN <= 1e-8
    assert rel(res["K"], ref["K"]) < 1e-5
