"""Test infrastructure only (the oracle): CPU restatements of the LNV
density-matrix functional and the DMM helpers, used by tests to check the
device path. Nothing in the library imports this module.

* ``transpose`` — matrix_transpose on one process (src/mult_module.f90:2609-2673):
  block (i,j,s) of the result is block (j,i,-s) of the input, transposed.
* ``lnv_sparse`` — LNV_matrix_multiply (src/mult_module.f90:1478-1823) with
  doK, doE, doM3, dophi, step by step on host arrays with the C oracle's
  products (so the range truncations are those of the device path).
* ``lnv_dense`` — the same functional in closed form on dense matrices
  (K = 3LSL - 2LSLSL, E = Tr[HK], Ne = 3Tr[LSLS] - 2Tr[LSLSLS],
  M3 = dE/dL = 3(HLS + SLH) - 2(HLSLS + SLHLS + SLSLH), phi = dNe/dL =
  6(SLS - SLSLS)); equal to ``lnv_sparse`` when no range truncates.
* ``solve_cubic`` — SolveCubic (src/DMMinModule.f90:2072-2143).
* ``do_pulay`` — DoPulay2D_nospin (src/Pulay.f90:188-262).

How each is pinned (tests/test_oracle_lnv.py): ``lnv_dense`` by finite
differences of E(L) and Ne(L) (the gradients) and by the projector (K = L
when L is the S-idempotent projector); ``lnv_sparse`` by ``lnv_dense`` on an
isolated cluster; ``solve_cubic`` by numpy's polynomial roots; ``do_pulay`` by
the Lagrange solution of min a.A.a with sum(a) = 1 (numpy.linalg.solve).
"""
from __future__ import annotations

import numpy as np

from . import mcweeny as mo


def transpose(sp, a, name):
    r = sp.rid[name]
    A = sp.idx[r]
    ident = A.ident
    tr = np.stack([ident[:, 1], ident[:, 0], -ident[:, 2], -ident[:, 3], -ident[:, 4]], axis=1)
    partner = A.find(tr)
    out = np.zeros_like(a)
    for x in range(len(A.offset)):
        y = partner[x]
        n1, n2 = A.nd1[x], A.nd2[x]            # block x is n1 x n2 (column-major)
        blk = a[A.offset[x]:A.offset[x] + n1 * n2].reshape(n2, n1).T
        out[A.offset[y]:A.offset[y] + n1 * n2] = blk.reshape(-1)   # column-major blk^T
    return out


def lnv_sparse(sp, L, S, H):
    """Returns dict(K, M3, phi, energy, electrons) per spin."""
    LS = sp.zeros("LS")
    sp.product(1, "L", "S", "LS", L, S, LS)
    SL = transpose(sp, LS, "LS")
    LH = sp.zeros("LH")
    sp.product(1, "L", "H", "LH", L, H, LH)
    HL = transpose(sp, LH, "LH")
    HLS = sp.zeros("LSL")
    sp.product(1, "LH", "S", "LSL", HL, S, HLS)
    SLH = transpose(sp, HLS, "LSL")
    A = sp.zeros("LSL")
    A = sp.sum(0.0, A, "LSL", 2.0, HLS, "LSL")
    A = sp.sum(1.0, A, "LSL", 1.0, SLH, "LSL")
    B = sp.zeros("L")
    sp.product(2, "L", "LS", "LSL", B, SL, A)          # B = A . SL^T
    Cm = transpose(sp, B, "L")
    SLS = sp.zeros("SLS")
    sp.product(1, "LS", "S", "SLS", SL, S, SLS)
    SLSLS = sp.zeros("L")
    sp.product(2, "L", "LS", "SLS", SLSLS, SL, SLS)    # SLSLS = SLS . SL^T
    LSL = sp.zeros("LSL")
    sp.product(1, "LS", "L", "LSL", LS, L, LSL)
    LSLSL = sp.zeros("H")
    sp.product(2, "H", "LS", "LSL", LSLSL, LS, LSL)    # LSLSL = LSL . LS^T at H range
    K = sp.zeros("H")
    K = sp.sum(0.0, K, "H", 3.0, LSL, "LSL")
    K = sp.sum(1.0, K, "H", -2.0, LSLSL, "H")
    energy = float(np.dot(H, K))
    M3 = sp.zeros("L")
    M3 = sp.sum(0.0, M3, "L", 3.0, HLS, "LSL")
    M3 = sp.sum(1.0, M3, "L", 3.0, SLH, "LSL")
    M3 = sp.sum(1.0, M3, "L", -1.0, B, "L")
    M3 = sp.sum(1.0, M3, "L", -1.0, Cm, "L")
    phi = sp.zeros("L")
    phi = sp.sum(0.0, phi, "L", 1.0, SLS, "SLS")
    e1 = float(np.dot(L, phi))
    e2 = float(np.dot(L, SLSLS))
    phi = sp.sum(6.0, phi, "L", -6.0, SLSLS, "L")
    return dict(K=K, M3=M3, phi=phi, energy=energy, electrons=3.0 * e1 - 2.0 * e2)


def lnv_dense(L, S, H):
    LS = L @ S
    LSL = LS @ L
    LSLSL = LSL @ S @ L
    K = 3.0 * LSL - 2.0 * LSLSL
    SLS = S @ L @ S
    HLS = H @ L @ S
    SLH = HLS.T
    M3 = 3.0 * (HLS + SLH) - 2.0 * (HLS @ L @ S + S @ L @ H @ L @ S + S @ L @ S @ L @ H)
    phi = 6.0 * (SLS - SLS @ L @ S)
    return dict(K=K, M3=M3, phi=phi, energy=float(np.sum(H * K)), electrons=float(np.sum(S * K)))


def solve_cubic(a, b, c, guess):
    Q = (a * a - 3.0 * b) / 9.0
    R = (2.0 * a ** 3 - 9.0 * a * b + 27.0 * c) / 54.0
    if R * R < Q ** 3:
        theta = np.arccos(R / np.sqrt(Q ** 3))
        T = -2.0 * np.sqrt(Q)
        z = [T * np.cos(theta / 3.0) - a / 3.0, T * np.cos((theta + 2 * np.pi) / 3.0) - a / 3.0,
             T * np.cos((theta - 2 * np.pi) / 3.0) - a / 3.0]
        if abs(z[0]) <= abs(z[1]) and abs(z[0]) <= abs(z[2]):
            return z[0]
        return z[1] if abs(z[1]) <= abs(z[2]) else z[2]
    Q, R = -Q, -R
    S = R + np.sqrt(Q ** 3 + R * R)
    T = R - np.sqrt(Q ** 3 + R * R)
    S = np.cbrt(S)
    T = np.cbrt(T)
    z = S + T - a / 3.0
    return guess if abs(z) > 10.0 * abs(guess) else z


def do_pulay(ipulay, A):
    """alpha_i = sum_j (A^-1)_ji / sum (A^-1); e_ipulay (1-based) when A is singular."""
    n = A.shape[0]
    try:
        inv = np.linalg.inv(A)
    except np.linalg.LinAlgError:
        out = np.zeros(n)
        out[ipulay - 1] = 1.0
        return out
    col = inv.sum(axis=0)
    return col / inv.sum()


__all__ = ["transpose", "lnv_sparse", "lnv_dense", "solve_cubic", "do_pulay", "mo"]
