"""CPU test that pins the FindMinDM oracle (oracle/dmm.py) by its fixed point.

On an isolated cluster no range truncates a product, so the constrained
minimum of the LNV functional at fixed electron number is the S-metric
eigen-projector of (H, S): the oracle, started from InitMcW's L, must reach it
(band energy within 1e-6 Ha/atom, electron count exact, K = P). A sign error in
a search direction, a wrong Pulay coefficient or a dropped projection would
stop it short of that point.
"""
import numpy as np

from cqmm.dm import Purifier
from cases import cluster_structure
from oracle import dmm as od
from oracle import mcweeny as mo


def test_oracle_findmindm_reaches_projector_on_cluster():
    xyz, box = cluster_structure(7, seed=3)
    pu = Purifier(xyz, box, 4, r_S=8.0, r_H=8.0, r_L=12.0, grid=(5, 5, 5), seed=3, device=False, lnv=True)
    s, R = pu.sys, pu.rid
    sp = mo.Sparse(s, R)
    nsf = pu.nsf
    S = mo.densify(s.blocks(R["S"]), pu.S_vals, nsf)
    H = mo.densify(s.blocks(R["H"]), pu.H_vals, nsf)
    T = mo.hotelling_dense(S)
    ne = pu.n_orb // 2
    L0, _, _ = mo.init_mcw_dense(H, S, T, float(ne), nsf)
    Ts = mo.sparsify(T, s.blocks(R["T"]), nsf, s.length(R["T"]))
    Ls = mo.sparsify(L0, s.blocks(R["L"]), nsf, s.length(R["L"]))
    d = od.Dmm(sp, pu.S_vals, pu.H_vals, Ts, float(ne), pu.natoms, tolerance=1e-10, n_L_iterations=100)
    res = d.run(Ls)
    P, e = mo.projector(H, S, ne)
    assert e[ne] - e[ne - 1] > 0.1
    assert res["done"]
    assert abs(res["energy"] - 2.0 * np.sum(P * H)) / pu.natoms < 1e-6
    assert abs(res["electrons"] - 2.0 * ne) < 1e-6
    assert np.abs(mo.densify(s.blocks(R["H"]), res["K"], nsf) - P).max() < 1e-4
    # The late stage ends on the residual test, below the tolerance.
    assert d.trace[-1][0] == "late" and d.trace[-1][2] < 1e-10
