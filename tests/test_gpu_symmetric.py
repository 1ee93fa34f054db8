"""GPU tests of symmetric-result plans (cqmm_plan_create_symmetric,
cqmm_sys_plan_create_symmetric): only the first block of each transpose pair of
the result is computed and the other written as its transpose.

Bars: for symmetric L and S, LSL = LS.L (type 1) and LSLSL = LSL.(LS)^T (type
2) through symmetric plans equal the full products within 1e-13 of the largest
element, are symmetric bit for bit (blocks (i, i, 0), their own partners, are computed), and the computed half equals the full
product's blocks bit for bit (same kernels, same products in the same order);
the same through cqmm_mat_mult and in the FP32 and split-BF16 product modes;
McWeeny with symmetric plans matches the oracle's McWMin as the full plans do
(stop step within one, L to 1e-8, energy to 1e-8 relative); pairs that are not
blocks of the result are refused.
"""
import numpy as np
import pytest

import cqmm
from cqmm import CqmmError, Plan
from cqmm.dm import Context, Purifier, LS_L_LSL, LSL_SL_L, L_S_LS
from oracle import mcweeny as mo

pytestmark = pytest.mark.gpu

SYM_LSL, SYM_LSLSL = 40, 41


def system(nsf, natoms=48, seed=7):
    xyz, edge = cqmm.random_structure(natoms, 0.05, 2.0, seed)
    return Purifier(xyz, edge, nsf, r_S=4.5, r_H=4.5, r_L=6.0, grid=(2, 2, 2), seed=seed)


def symmetric_L(pu):
    c = pu.ctx
    L = c.allocate(pu.rid["L"])
    c.set(L, pu.sys.fill(pu.rid["L"], 1, pu.seed + 3, 21, (0.1, 1.5)))
    c.symmetrise(L)
    return L


@pytest.mark.parametrize("nsf", [4, 13, (4, 9, 13)], ids=["4", "13", "mix"])
def test_symmetric_plans_match_full_products(nsf):
    pu = system(nsf if not isinstance(nsf, tuple) else
                np.random.default_rng(5).choice(nsf, size=48).astype(np.int32))
    R, c = pu.rid, pu.ctx
    c.add_mult(SYM_LSL, 1, R["LS"], R["L"], R["LSL"], symmetric=True)
    c.add_mult(SYM_LSLSL, 2, R["L"], R["LS"], R["LSL"], symmetric=True)
    L = symmetric_L(pu)
    LS, LSL_f, LSL_s = c.allocate(R["LS"]), c.allocate(R["LSL"]), c.allocate(R["LSL"])
    X_f, X_s = c.allocate(R["L"]), c.allocate(R["L"])
    print(f"symmetric plans, nsf {nsf}: {pu.natoms} atoms", flush=True)
    c.product(L, pu.matS, LS, L_S_LS)
    c.product(LS, L, LSL_f, LS_L_LSL)
    c.product(LS, L, LSL_s, SYM_LSL)
    # matrix_product(LSL, LS, LSLSL, mult(LSL_SL_L)): LSLSL = LSL.(LS)^T at L range.
    c.product(LSL_f, LS, X_f, LSL_SL_L)
    c.product(LSL_f, LS, X_s, SYM_LSLSL)
    for rid, full, sym in ((R["LSL"], LSL_f, LSL_s), (R["L"], X_f, X_s)):
        f, s = c.get(full), c.get(sym)
        assert np.abs(f - s).max() <= 1e-13 * np.abs(f).max()
        p, q, n1, n2 = pu.sys.transpose_pairs(rid)
        computed = 0
        for a, b, m, n in zip(p - 1, q - 1, n1, n2):
            if a == b:  # a block (i, i, 0) is its own partner and is computed
                np.testing.assert_array_equal(s[a:a + m * n], f[a:a + m * n])
                continue
            blk = s[a:a + m * n].reshape(n, m)  # column-major m x n
            np.testing.assert_array_equal(s[b:b + m * n].reshape(m, n), blk.T)
            # One block of the pair is computed: the full product's, bit for bit.
            if np.array_equal(s[a:a + m * n], f[a:a + m * n]):
                computed += 1
            else:
                np.testing.assert_array_equal(s[b:b + m * n], f[b:b + m * n])
        assert computed > 0


def test_symmetric_plan_through_mat_mult_and_reduced_precision():
    pu = system(13)
    R, c = pu.rid, pu.ctx
    c.add_mult(SYM_LSL, 1, R["LS"], R["L"], R["LSL"], symmetric=True)
    L = symmetric_L(pu)
    LS = c.allocate(R["LS"])
    c.product(L, pu.matS, LS, L_S_LS)
    ls, l_vals = c.get(LS), c.get(L)
    desc = pu.sys.mult_desc(1, R["LS"], R["L"], R["LSL"])
    full = Plan(desc)
    sym = Plan(desc, pairs=pu.sys.transpose_pairs(R["LSL"]))
    assert sym.info().n_mirror > 0 and full.info().n_mirror == 0
    assert sym.info().flops < 0.6 * full.info().flops
    out_f = np.zeros(pu.sys.length(R["LSL"]))
    out_s = np.full(pu.sys.length(R["LSL"]), np.nan)
    full.mat_mult(ls, l_vals, out_f)
    sym.mat_mult(ls, l_vals, out_s)
    assert np.abs(out_f - out_s).max() <= 1e-13 * np.abs(out_f).max()
    for precision, tol in ((1, 1e-6), (4, 1e-4)):
        c.product_precision = precision
        A, B = c.allocate(R["LSL"]), c.allocate(R["LSL"])
        c.product(LS, L, A, LS_L_LSL)
        c.product(LS, L, B, SYM_LSL)
        a, b = c.get(A), c.get(B)
        print(f"precision {precision}: max diff {np.abs(a - b).max() / np.abs(a).max():.2e}", flush=True)
        assert np.abs(a - b).max() <= tol * np.abs(a).max() and np.isfinite(b).all()
        c.free(B)
        c.free(A)
    c.product_precision = 0


def test_symmetric_plan_refuses_pairs_outside_the_result():
    """Offsets that are not the start of a result block, or reach past the
    result, are refused (the mirror then never reads or writes outside the
    result's blocks)."""
    pu = system(4, natoms=24)
    R = pu.rid
    desc = pu.sys.mult_desc(1, R["LS"], R["L"], R["LSL"])
    p, q, n1, n2 = pu.sys.transpose_pairs(R["LSL"])
    off = p != q
    with pytest.raises(CqmmError):
        Plan(desc, pairs=(p[off] + 1, q[off], n1[off], n2[off]))
    with pytest.raises(CqmmError):
        Plan(desc, pairs=(p, q + pu.sys.length(R["LSL"]), n1, n2))
    k = int(np.nonzero(off)[0][0])
    with pytest.raises(CqmmError):  # one block computed by one pair and mirrored by another
        Plan(desc, pairs=(np.append(p, q[k]), np.append(q, p[k]), np.append(n1, n2[k]), np.append(n2, n1[k])))
    Plan(desc, pairs=(q, p, n2, n1))  # either block of a pair may be the computed one


@pytest.mark.parametrize("structure", ["random24", "mix48"])
def test_symmetric_mcweeny_matches_oracle(structure):
    if structure == "random24":
        xyz, edge = cqmm.random_structure(24, 0.05, 2.0, 5)
        nsf, seed = 4, 5
    else:
        xyz, edge = cqmm.random_structure(48, 0.05, 2.0, 9)
        nsf, seed = np.random.default_rng(9).choice([4, 9, 13], size=48).astype(np.int32), 9
    pu = Purifier(xyz, edge, nsf, r_S=4.5, r_H=4.5, r_L=6.0, grid=(2, 2, 2), seed=seed, symmetric=True)
    print(f"symmetric McWeeny {structure}", flush=True)
    sp = mo.Sparse(pu.sys, pu.rid)
    pu.hotelling()
    pu.init_mcweeny()
    L0 = pu.ctx.get(pu.matL)
    res = pu.mcweeny(precision=0, max_iter=60)
    L = pu.ctx.get(pu.matL)
    L_ref, E_ref, Ne_ref, n_ref = sp.mcwmin(L0, pu.S_vals, pu.H_vals, 60)
    assert abs(res["iterations"] - n_ref) <= 1
    assert abs(res["energy"] - E_ref) < 1e-8 * max(1.0, abs(E_ref))
    assert np.abs(L - L_ref).max() < 1e-8


def test_symmetric_split_bf16_mcweeny_against_fp64_oracle():
    xyz, edge = cqmm.random_structure(4, 0.05, 2.0, 11)
    pu = Purifier(xyz, edge, 13, r_S=8.0, r_H=8.0, r_L=12.0, seed=11, symmetric=True)
    sp = mo.Sparse(pu.sys, pu.rid, ompgemm=True)
    pu.hotelling()
    pu.init_mcweeny()
    L0 = pu.ctx.get(pu.matL)
    r = pu.mcweeny(precision=4, max_iter=100, fp32_tol=0.0)
    L_ref, E_ref, Ne_ref, n_ref = sp.mcwmin(L0, pu.S_vals, pu.H_vals, 100)
    dE = abs(2.0 * r["energy"] - 2.0 * E_ref) / pu.natoms
    dN = abs(2.0 * r["electrons"] - 2.0 * Ne_ref)
    print(f"symmetric split-BF16 McWeeny: dE/atom {dE:.2e}, dNe {dN:.2e}", flush=True)
    assert r["fp32_iterations"] > 0 and dE <= 1e-6 and dN <= 1e-8


def test_morton_tile_rows_give_the_same_products(monkeypatch):
    """Full plans with tile rows in the columns' Morton order
    (CQMM_TILE_ROWS=morton): other tiles, the same products in the same order
    for every output block, so the result is the primary-order plan's bit for
    bit (both product types, species mix)."""
    nsf = np.random.default_rng(5).choice([4, 9, 13], size=48).astype(np.int32)
    pu = system(nsf)
    R, c = pu.rid, pu.ctx
    monkeypatch.setenv("CQMM_TILE_ROWS", "morton")
    c.add_mult(42, 1, R["LS"], R["L"], R["LSL"])
    c.add_mult(43, 2, R["L"], R["LS"], R["LSL"])
    monkeypatch.delenv("CQMM_TILE_ROWS")
    L = symmetric_L(pu)
    LS, A, B = c.allocate(R["LS"]), c.allocate(R["LSL"]), c.allocate(R["LSL"])
    X, Y = c.allocate(R["L"]), c.allocate(R["L"])
    c.product(L, pu.matS, LS, L_S_LS)
    c.product(LS, L, A, LS_L_LSL)
    c.product(LS, L, B, 42)
    np.testing.assert_array_equal(c.get(A), c.get(B))
    c.product(A, LS, X, LSL_SL_L)
    c.product(A, LS, Y, 43)
    np.testing.assert_array_equal(c.get(X), c.get(Y))
