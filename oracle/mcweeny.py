"""McWeeny purification oracle (test infrastructure only).

Two restatements of McW_matrix_multiply_nonds + McWMin (src/McWeeny.f90:620-740,
107-272), InitMcW (:340-570, GetHLimits :1020-1109) and the Hotelling S^-1
iteration (src/S_matrix_module.f90:746-1034):

  * dense_*  — on dense matrices (an isolated cluster, where no range truncates
    anything), pinned by closed forms: the step is the polynomial map
    x -> ((1+c)x^2 - x^3 + (1-2c)x)/(1-c) (c < 1/2) or ((1+c)x^2 - x^3)/c on the
    eigenvalues x of S^{1/2} L S^{1/2}, with c = sum(x^2-x^3)/sum(x-x^2); the
    converged L is the S-metric projector onto the lowest Ne eigenvectors of
    H c = e S c (scipy.linalg.eigh).
  * Sparse    — the same loop on image-resolved block matrices in the reference
    layout, products through the C oracle's mat_mult, range changes through
    block-identity maps (matrix_add's semantics: only blocks common to both
    patterns change).
  * hlimits_dense_blocks (GetHLimits) — pinned by a two-atom known answer
    worked by hand from the reference loop (non-symmetric blocks, mixed sizes)
    and by the diagonal case (min/max of |H_ii|).
  * init_mcw_dense (InitMcW, McWRho0) — pinned by hand-worked three-level
    known answers on both mubar branches (with T = S^-1 and with T != S^-1),
    the no-mubar case, and, with T = S^-1, Tr[L0 S] = Ne and the S-metric
    occupancies of L0 being the linear map (lambda (mubar - e) + Ne)/No of the
    generalised eigenvalues (tests/test_oracle_initmcw.py).
  * Sparse.mcwmin_spin (McWMin for two spin channels, fixed and variable
    populations, src/McWeeny.f90:144-256, 772-890) and init_mcw_dense_spin
    (the direct-sum InitMcW for variable populations, :380-395, 460-565) —
    pinned by: identical channels reducing to the unpolarised loop; fixed
    populations converging to each spin's own S-metric projector; variable
    populations converging to the lowest ne_up + ne_dn levels of both spectra
    together (electrons move to the lower spin); a hand-worked two-level known
    answer of the direct-sum mubar and lambda; per-spin overlaps (spin-dependent
    support functions) reducing to the shared one and converging to each
    spin's projector in its own metric (tests/test_oracle_mcweeny_spin.py).
"""
from __future__ import annotations

import numpy as np

from . import brute
from . import mat_mult as oracle_mat_mult


# ---------------------------------------------------------------- dense forms
def dense_step(L, S, H):
    """One McWeeny step on dense matrices: returns (rho_new, E, Ne, c)."""
    LS = L @ S
    LSL = LS @ L
    LSLSL = LSL @ S @ L
    c1 = np.sum(S * (LSL - LSLSL))
    c2 = np.sum(S * (L - LSL))
    c = c1 / c2
    if c < 0.5:
        rho = ((1 + c) * LSL - LSLSL + (1 - 2 * c) * L) / (1 - c)
    else:
        rho = ((1 + c) * LSL - LSLSL) / c
    E = np.sum(rho * H)
    Ne = np.sum(rho * S)
    # A nearly idempotent L makes c2 vanish; the reference then returns c = -0.01,
    # which ends McWMin (src/McWeeny.f90:737).
    if abs(c2) < 1e-6:
        c = -0.01
    return rho, E, Ne, c


def step_polynomial(x):
    """Closed form of the step on eigenvalues x of S^1/2 L S^1/2."""
    c = np.sum(x ** 2 - x ** 3) / np.sum(x - x ** 2)
    if c < 0.5:
        return ((1 + c) * x ** 2 - x ** 3 + (1 - 2 * c) * x) / (1 - c), c
    return ((1 + c) * x ** 2 - x ** 3) / c, c


def dense_mcwmin(L, S, H, max_iter=100):
    """McWMin on dense matrices (same stop tests). Returns (L, E, Ne, iterations)."""
    oldE = 1e30
    Lold = L.copy()
    E = Ne = None
    n = 0
    for n in range(1, max_iter + 1):
        rho, omega, ne, c = dense_step(L, S, H)
        if omega > oldE or abs(omega - oldE) < 1e-8 or c > 1 or c < 0:
            break
        L = 0.5 * (rho + rho.T)
        oldE, E, Ne = omega, omega, ne
    else:
        n = max_iter + 1
    if n < 3:
        L = Lold
    return L, E, Ne, n


def projector(H, S, n_occ):
    """S-metric density matrix of the lowest n_occ states of H c = e S c."""
    import scipy.linalg
    e, V = scipy.linalg.eigh(H, S)
    Vo = V[:, :n_occ]
    return Vo @ Vo.T, e


def hotelling_dense(S, tol=1e-10, max_iter=50):
    """Iter_Hott_InvS on dense matrices; returns T."""
    n = S.shape[0]
    I = np.eye(n)
    T = S / np.sum(S * S)
    deltaomega = oldomega = 0.0
    Told = T.copy()
    for it in range(1, max_iter + 1):
        deltaomega *= 0.5
        if not (it < 3 or abs(deltaomega) > tol):
            break
        TS = T @ S
        A = I - TS
        T1 = 2 * T - TS @ T
        omega = np.sum(A * A)
        deltaomega = omega - oldomega
        if omega > oldomega and oldomega != 0.0:
            T = Told
            break
        oldomega = omega
        Told = T.copy()
        T = T1
    return T


def hlimits_dense_blocks(H, nsf):
    """GetHLimits' sums for a dense H of an isolated cluster (onsite blocks add |H|,
    others subtract), indexed per column orbital of each neighbour block."""
    off = np.concatenate([[0], np.cumsum(nsf)])
    hmin, hmax = 1e30, -1e30
    for a in range(len(nsf)):
        lo = np.zeros(64)
        hi = np.zeros(64)
        for b in range(len(nsf)):
            blk = np.abs(H[off[a]:off[a + 1], off[b]:off[b + 1]])
            col = blk.sum(axis=0)
            lo[:nsf[b]] += col if a == b else -col
            hi[:nsf[b]] += col
        hmin = min(hmin, lo[:nsf[a]].min())
        hmax = max(hmax, hi[:nsf[a]].max())
    return hmin, hmax


def init_mcw_dense(H, S, T, ne, nsf, r_L_mask=None):
    """InitMcW on dense matrices: L0 = ((lambda*mubar + Ne)/No) T - (lambda/No) T H T."""
    n_o = float(np.sum(nsf))
    SX = np.trace(T @ S)
    THT = T @ H @ T
    SXHX = np.sum(S * THT)
    hmin, hmax = hlimits_dense_blocks(H, nsf)
    A = 1.0 - SX / n_o
    mu1 = SXHX / n_o + hmax * A
    mu2 = (hmin * A * ne / (n_o - ne) - SXHX / n_o) / (n_o * A / (n_o - ne) - 1.0)
    if ne / (hmax - mu1) < (n_o - ne) / (mu1 - hmin):
        mubar, lam = mu1, ne / (hmax - mu1)
    elif ne / (hmax - mu2) > (n_o - ne) / (mu2 - hmin):
        mubar, lam = mu2, (n_o - ne) / (mu2 - hmin)
    else:
        raise RuntimeError("InitMcW: cannot find mubar")
    return (lam * mubar + ne) / n_o * T - lam / n_o * THT, mubar, lam


def init_mcw_dense_spin(Hs, S, T, nes, nsf):
    """InitMcW for variable spin populations (src/McWeeny.f90:380-395,460-565) on
    dense matrices: the direct sum of both spin channels, with n_o_ds = 2 n_o,
    n_e_ds = ne_up + ne_dn (spin_factor 1), SXHX_ds the sum over spins, the
    limits hmin_ds / hmax_ds the extremes over spins, A_ds = 1 - 2 SX / n_o_ds,
    one mubar and lambda for both spins, and L0_s = ((lambda mubar + n_e_ds) /
    n_o_ds) T - (lambda / n_o_ds) T H_s T. The second branch's test is the
    reference's as written there: n_e/(hmax - mu2) < (n_o - n_e)/(mu2 - hmin)."""
    n_o = float(np.sum(nsf))
    n_o_ds = 2.0 * n_o
    n_e = float(nes[0] + nes[1])
    SX = np.trace(T @ S)
    THTs = [T @ H @ T for H in Hs]
    SXHX = sum(float(np.sum(S * THT)) for THT in THTs)
    lims = [hlimits_dense_blocks(H, nsf) for H in Hs]
    hmin, hmax = min(l[0] for l in lims), max(l[1] for l in lims)
    A = 1.0 - (SX + SX) / n_o_ds
    if n_o_ds == n_e:
        mubar, lam = 0.0, 0.0
    else:
        mu1 = SXHX / n_o_ds + hmax * A
        mu2 = (hmin * A * n_e / (n_o_ds - n_e) - SXHX / n_o_ds) / (n_o_ds * A / (n_o_ds - n_e) - 1.0)
        if n_e / (hmax - mu1) < (n_o_ds - n_e) / (mu1 - hmin):
            mubar, lam = mu1, n_e / (hmax - mu1)
        elif n_e / (hmax - mu2) < (n_o_ds - n_e) / (mu2 - hmin):
            mubar, lam = mu2, (n_o_ds - n_e) / (mu2 - hmin)
        else:
            raise RuntimeError("InitMcW: cannot find mubar")
    Ls = [(lam * mubar + n_e) / n_o_ds * T - lam / n_o_ds * THT for THT in THTs]
    return Ls, mubar, lam


def densify(blocks, vals, nsf):
    """Dense matrix of an image-free pattern (all shifts zero)."""
    if np.any(blocks["shift"] != 0):
        raise ValueError("densify needs a pattern without periodic images")
    off = np.concatenate([[0], np.cumsum(nsf)])
    n = int(off[-1])
    M = np.zeros((n, n))
    for b in range(len(blocks["offset"])):
        i, j = blocks["gid_i"][b], blocks["gid_j"][b]
        n1, n2 = blocks["nd1"][b], blocks["nd2"][b]
        o = blocks["offset"][b]
        M[off[i]:off[i] + n1, off[j]:off[j] + n2] = vals[o:o + n1 * n2].reshape(n2, n1).T
    return M


def sparsify(M, blocks, nsf, length):
    off = np.concatenate([[0], np.cumsum(nsf)])
    v = np.zeros(length)
    for b in range(len(blocks["offset"])):
        i, j = blocks["gid_i"][b], blocks["gid_j"][b]
        n1, n2 = blocks["nd1"][b], blocks["nd2"][b]
        o = blocks["offset"][b]
        v[o:o + n1 * n2] = M[off[i]:off[i] + n1, off[j]:off[j] + n2].T.reshape(-1)
    return v


# --------------------------------------------------------------- sparse form
class Sparse:
    """mult_module operations on host arrays of a cqmm System (oracle products)."""

    def __init__(self, system, rid, ompgemm=False):
        """ompgemm: products through the ompGemm restatement (OpenMP + numpy's
        OpenBLAS dgemm), which equals the default kernels' to 1e-13
        (tests/test_oracle.py); for workloads the single-threaded default
        restatement would take minutes on."""
        self.ompgemm = ompgemm
        if ompgemm:
            from . import use_numpy_openblas
            use_numpy_openblas()
        self.sys = system
        self.rid = dict(rid)
        self.blocks = {r: system.blocks(r) for r in set(rid.values())}
        self.idx = {r: brute.BlockIndex(self.blocks[r], system.natoms) for r in set(rid.values())}
        self.len = {r: system.length(r) for r in set(rid.values())}
        self._maps = {}
        self._desc = {}

    def zeros(self, name):
        return np.zeros(self.len[self.rid[name]])

    def sum(self, alpha, a, ra, beta, b, rb):
        """matrix_sum(alpha, A, beta, B): only blocks common to both patterns change."""
        ra, rb = self.rid[ra], self.rid[rb]
        if ra == rb:
            a *= alpha
            a += beta * b
            return a
        key = (ra, rb)
        if key not in self._maps:
            A, B = self.idx[ra], self.idx[rb]
            ka = brute.keys(A.ident, self.sys.natoms)
            pos = np.searchsorted(B.ks, ka)
            pos = np.minimum(pos, len(B.ks) - 1)
            ok = B.ks[pos] == ka
            ia = np.nonzero(ok)[0]
            ib = B.order[pos[ok]]
            da = np.concatenate([np.arange(A.offset[x], A.offset[x] + A.nd1[x] * A.nd2[x]) for x in ia])
            db = np.concatenate([np.arange(B.offset[y], B.offset[y] + B.nd1[y] * B.nd2[y]) for y in ib])
            self._maps[key] = (da, db)
        da, db = self._maps[key]
        a[da] = alpha * a[da] + beta * b[db]
        return a

    def product(self, mult_type, ra, rb, rc, a, b, c):
        key = (mult_type, self.rid[ra], self.rid[rb], self.rid[rc])
        if key not in self._desc:
            self._desc[key] = self.sys.mult_desc(mult_type, *key[1:])
        if self.ompgemm:
            from . import mat_mult_ompgemm
            mat_mult_ompgemm(self._desc[key], a, b, c)
        else:
            oracle_mat_mult(self._desc[key], a, b, c)

    def symmetrise(self, a, name):
        r = self.rid[name]
        A = self.idx[r]
        ident = A.ident.copy()
        tr = np.stack([ident[:, 1], ident[:, 0], -ident[:, 2], -ident[:, 3], -ident[:, 4]], axis=1)
        partner = A.find(tr)
        out = a.copy()
        for x in range(len(A.offset)):
            y = partner[x]
            n1, n2 = A.nd1[x], A.nd2[x]
            P = a[A.offset[x]:A.offset[x] + n1 * n2].reshape(n2, n1).T
            Q = a[A.offset[y]:A.offset[y] + n1 * n2].reshape(n1, n2).T
            out[A.offset[x]:A.offset[x] + n1 * n2] = (0.5 * P + 0.5 * Q.T).T.reshape(-1)
        return out

    def step(self, L, S, H):
        """McW_matrix_multiply_nonds on host arrays: returns (rho, E, Ne, c)."""
        LS = self.zeros("LS")
        self.product(1, "L", "S", "LS", L, S, LS)
        LSL = self.zeros("LSL")
        self.product(1, "LS", "L", "LSL", LS, L, LSL)
        LSLSL = self.zeros("L")
        self.product(2, "L", "LS", "LSL", LSLSL, LS, LSL)
        top = self.zeros("S")
        top = self.sum(0.0, top, "S", 1.0, LSL, "LSL")
        top = self.sum(1.0, top, "S", -1.0, LSLSL, "L")
        c1 = float(np.dot(S, top))
        bottom = self.zeros("S")
        bottom = self.sum(0.0, bottom, "S", 1.0, L, "L")
        bottom = self.sum(1.0, bottom, "S", -1.0, LSL, "LSL")
        c2 = float(np.dot(S, bottom))
        c = c1 / c2
        rho = self.zeros("L")
        rho = self.sum(0.0, rho, "L", -1.0, LSLSL, "L")
        rho = self.sum(1.0, rho, "L", 1.0 + c, LSL, "LSL")
        if c < 0.5:
            rho = self.sum(1.0, rho, "L", 1.0 - 2.0 * c, L, "L")
            rho *= 1.0 / (1.0 - c)
        else:
            rho *= 1.0 / c
        botH = self.zeros("H")
        botH = self.sum(0.0, botH, "H", 1.0, rho, "L")
        E = float(np.dot(botH, H))
        top = self.sum(0.0, top, "S", 1.0, rho, "L")
        Ne = float(np.dot(top, S))
        if abs(c2) < 1e-6:
            c = -0.01
        return rho, E, Ne, c

    def step_traces(self, L, S):
        """The products and c1, c2 of McW_matrix_multiply_ds for one spin
        (src/McWeeny.f90:806-840): (LSL, LSLSL, c1, c2)."""
        LS = self.zeros("LS")
        self.product(1, "L", "S", "LS", L, S, LS)
        LSL = self.zeros("LSL")
        self.product(1, "LS", "L", "LSL", LS, L, LSL)
        LSLSL = self.zeros("L")
        self.product(2, "L", "LS", "LSL", LSLSL, LS, LSL)
        top = self.zeros("S")
        top = self.sum(0.0, top, "S", 1.0, LSL, "LSL")
        top = self.sum(1.0, top, "S", -1.0, LSLSL, "L")
        bottom = self.zeros("S")
        bottom = self.sum(0.0, bottom, "S", 1.0, L, "L")
        bottom = self.sum(1.0, bottom, "S", -1.0, LSL, "LSL")
        return LSL, LSLSL, float(np.dot(S, top)), float(np.dot(S, bottom))

    def rho_of(self, L, LSL, LSLSL, c):
        """rho' = ((1+c)LSL - LSLSL + (1-2c)L)/(1-c) for c < 1/2, else ((1+c)LSL - LSLSL)/c."""
        rho = self.zeros("L")
        rho = self.sum(0.0, rho, "L", -1.0, LSLSL, "L")
        rho = self.sum(1.0, rho, "L", 1.0 + c, LSL, "LSL")
        if c < 0.5:
            rho = self.sum(1.0, rho, "L", 1.0 - 2.0 * c, L, "L")
            rho *= 1.0 / (1.0 - c)
        else:
            rho *= 1.0 / c
        return rho

    def traces_of(self, rho, S, H):
        botH = self.zeros("H")
        botH = self.sum(0.0, botH, "H", 1.0, rho, "L")
        top = self.zeros("S")
        top = self.sum(0.0, top, "S", 1.0, rho, "L")
        return float(np.dot(botH, H)), float(np.dot(top, S))

    def mcwmin_spin(self, Ls, S, Hs, max_iter=100, fixed=False):
        """McWMin for two spin channels (src/McWeeny.f90:144-256). S is one
        overlap for both spins, or a pair (S_up, S_dn) when the support
        functions depend on spin (flag_SpinDependentSF: matS(spin)).

        fixed=True (fixed spin populations): McW_matrix_multiply_nonds per
        spin, each spin's stop test marks it done, the loop ends when both are
        done, and until then both spins take every step. fixed=False (variable
        spin, McW_matrix_multiply_ds :772-890): one c = (c1_up + c1_dn) /
        (c2_up + c2_dn) for both spins (spin_factor 1), one stop test on the
        total energy. Fewer than 3 steps restore both L. Returns
        (Ls, energies per spin, electrons per spin, steps)."""
        Ss = list(S) if isinstance(S, (list, tuple)) else [S, S]
        Ls = [L.copy() for L in Ls]
        Lold = [L.copy() for L in Ls]
        E = [None, None]
        Ne = [None, None]
        n = 0
        if fixed:
            oldE = [1e30, 1e30]
            done = [False, False]
            for n in range(1, max_iter + 1):
                res = [self.step(Ls[s], Ss[s], Hs[s]) for s in range(2)]
                for s in range(2):
                    _, om, _, c = res[s]
                    if om > oldE[s] or abs(om - oldE[s]) < 1e-8 or c > 1 or c < 0:
                        done[s] = True
                if done[0] and done[1]:
                    break
                for s in range(2):
                    rho, om, ne, _ = res[s]
                    Ls[s] = self.symmetrise(rho, "L")
                    oldE[s], E[s], Ne[s] = om, om, ne
            else:
                n = max_iter + 1
        else:
            oldE = 2e30
            for n in range(1, max_iter + 1):
                parts = [self.step_traces(Ls[s], Ss[s]) for s in range(2)]
                c1 = parts[0][2] + parts[1][2]
                c2 = parts[0][3] + parts[1][3]
                if not abs(c2) > 1e-15:
                    raise ZeroDivisionError("McW_matrix_multiply_ds: c2 is zero")
                c = c1 / c2
                rhos = [self.rho_of(Ls[s], parts[s][0], parts[s][1], c) for s in range(2)]
                tr = [self.traces_of(rhos[s], Ss[s], Hs[s]) for s in range(2)]
                om = tr[0][0] + tr[1][0]
                if om > oldE or abs(om - oldE) < 1e-8 or c > 1 or c < 0:
                    break
                for s in range(2):
                    Ls[s] = self.symmetrise(rhos[s], "L")
                    E[s], Ne[s] = tr[s]
                oldE = om
            else:
                n = max_iter + 1
        if n < 3:
            Ls = Lold
        return Ls, E, Ne, n

    def mcwmin(self, L, S, H, max_iter=100):
        oldE = 1e30
        Lold = L.copy()
        E = Ne = None
        n = 0
        for n in range(1, max_iter + 1):
            rho, omega, ne, c = self.step(L, S, H)
            if omega > oldE or abs(omega - oldE) < 1e-8 or c > 1 or c < 0:
                break
            L = self.symmetrise(rho, "L")
            oldE, E, Ne = omega, omega, ne
        else:
            n = max_iter + 1
        if n < 3:
            L = Lold
        return L, E, Ne, n
