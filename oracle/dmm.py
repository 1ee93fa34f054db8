"""Test infrastructure only (the oracle): a CPU restatement of FindMinDM on
host arrays with the sparse oracle's products (oracle/mcweeny.py Sparse and
oracle/lnv.py lnv_sparse), one spin, spin factor 2, electron number fixed
(vary_mu). Nothing in the library imports this module.

Follows, call for call:
  correct_electron_number_fixspin  src/DMMinModule.f90:1676-1817
  lineMinL                         src/DMMinModule.f90:1395-1614
  earlyDM                          src/DMMinModule.f90:366-723
  lateDM                           src/DMMinModule.f90:809-1339
  FindMinDM (early/late rounds)    src/DMMinModule.f90:172-312
with SolveCubic and DoPulay from oracle/lnv.py. FindMinDM's reset to
InitMcW/McWMin after an inflexion is the caller's (as on the device); the
rounds are capped at ten.

Pinned by: its pieces (lnv_sparse, solve_cubic, do_pulay) are pinned in
tests/test_oracle_lnv.py; the loop itself by its fixed point on an isolated
cluster, the S-metric eigen-projector of (H, S) (tests/test_oracle_dmm.py).
"""
from __future__ import annotations

import numpy as np

from . import lnv as lo

RD_ERR = 1.0e-8
SPIN = 2.0


class Dmm:
    def __init__(self, sp, S, H, T, n_electrons, n_atoms, tolerance=1e-7, n_L_iterations=50, maxpulay=5,
                 min_pulay_step=1e-3, max_pulay_step=0.1, lin_tol=0.1, global_tolerance=False):
        self.sp, self.S, self.H, self.T = sp, S, H, T
        self.Tt = lo.transpose(sp, T, "T")
        self.ne = n_electrons
        self.n_atoms = n_atoms
        self.tol = tolerance
        self.n_L = n_L_iterations
        self.mp = maxpulay
        self.min_step, self.max_step = min_pulay_step, max_pulay_step
        self.lin_tol = lin_tol
        self.glob = global_tolerance
        self.delta_e = 0.0
        self.trace = []           # (stage, energy, residual) per iteration

    # -- helpers ---------------------------------------------------------------
    def lnv(self, L):
        return lo.lnv_sparse(self.sp, L, self.S, self.H)

    def precondition(self, X, right):
        sp = self.sp
        TL = sp.zeros("TL")
        sp.product(1, "T", "L", "TL", self.T, X, TL)
        out = sp.zeros("L")
        sp.product(2, "L", "T", "TL", out, right, TL)   # out = TL . right^T
        return out

    def residual(self, tr):
        return SPIN * tr if self.glob else SPIN * tr / self.n_atoms

    # -- correct_electron_number_fixspin ---------------------------------------
    def correct(self, L):
        e0 = self.ne
        done, it = False, 0
        while not done and it < 20:
            r = self.lnv(L)
            el, phi = r["electrons"], r["phi"]
            Sphi = self.precondition(phi, self.Tt)
            g0 = float(np.dot(Sphi, phi))
            step1 = (e0 - el) / g0
            if abs(e0 - el) < 1.0e-9:
                L = L + step1 * Sphi
                done = True
            else:
                step = step1
                L = L + step1 * Sphi
                r2 = self.lnv(L)
                el2, phi2 = r2["electrons"], r2["phi"]
                Sphi2 = self.precondition(phi2, self.Tt)
                g1 = float(np.dot(phi, Sphi2))
                D, C = el - e0, g0
                B = 3.0 * (el2 - el) / (step * step) - (g1 + 2.0 * g0) / step
                recA = step ** 3 / (2.0 * (el - el2) + (g0 + g1) * step)
                B, C, D = B * recA, C * recA, D * recA
                truestep = lo.solve_cubic(B, C, D, step)
                L = L + (truestep - step) * Sphi
                if truestep == step:
                    el2 = self.lnv(L)["electrons"]
                    if abs(el2 - e0) < 1.0e-4:
                        done = True
                    it += 1
                else:
                    done = True
        return L

    def project(self, SM3, M3, phi):
        Sphi = self.precondition(phi, self.T)
        en = float(np.dot(SM3, phi))
        nn = float(np.dot(Sphi, phi))
        return Sphi, en, nn

    # -- lineMinL ----------------------------------------------------------------
    def line_min(self, L, M3, D, SM3, e_in):
        M3old = M3.copy()
        g0 = SPIN * float(np.dot(D, M3))
        ig0 = g0
        step = self.delta_e / g0
        if abs(step) < 1.0e-2:
            step = 0.01
        if abs(step) > 0.1:
            step = 0.1
        L = L + step * D
        r = self.lnv(L)
        M3 = r["M3"]
        energy1 = SPIN * r["energy"]
        g1 = SPIN * float(np.dot(D, M3))
        igcross = SPIN * float(np.dot(SM3, M3))
        SM3 = self.precondition(M3, self.T)
        ig1 = SPIN * float(np.dot(SM3, M3))
        ig1_step = ig1
        Dc, Cc = e_in, g0
        Bc = 3.0 * (energy1 - Dc) / (step * step) - (g1 - 2.0 * Cc) / step
        Ac = (g1 - 2.0 * Bc * step - Cc) / (3.0 * step * step)
        SQ = 4.0 * Bc * Bc - 12.0 * Ac * Cc
        if SQ < 0:
            return L, M3, SM3, None, True, None
        truestep = (-2.0 * Bc + np.sqrt(SQ)) / (6.0 * Ac) if abs(Ac) > RD_ERR else -Cc / (2.0 * Bc)
        if truestep < 1.0e-4:
            truestep = 1.0e-4
        L = L + (truestep - step) * D
        L = self.sp.symmetrise(L, "L")
        r = self.lnv(L)
        M3 = r["M3"]
        e_out = SPIN * r["energy"]
        self.delta_e = e_out - e_in
        if truestep < step:
            lamG = truestep / step
            iG = ig0 * (1 - lamG) ** 2 + ig1 * lamG ** 2 + 2 * igcross * lamG * (1 - lamG)
            zeta = (iG - ig1) / (ig0 - ig1)
        else:
            SM3 = self.precondition(M3, self.T)
            ig1 = SPIN * float(np.dot(SM3, M3))
            igcross = SPIN * float(np.dot(SM3, M3old))
            lamG = step / truestep
            iG = ig0 * (1 - lamG) ** 2 + ig1 * lamG ** 2 + 2 * igcross * lamG * (1 - lamG)
            zeta = (iG - ig1_step) / (ig0 - ig1_step)
        return L, M3, SM3, e_out, False, zeta

    # -- earlyDM -----------------------------------------------------------------
    def early(self, L):
        """Returns (L, iterations, inflexion, energy, residual)."""
        L = self.correct(L)
        r = self.lnv(L)
        M3, phi = r["M3"], r["phi"]
        energy0 = SPIN * r["energy"]
        SM3 = self.precondition(M3, self.T)
        search = -SM3
        Sphi, en, nn = self.project(SM3, M3, phi)
        if abs(nn) > RD_ERR:
            search = search + (en / nn) * Sphi
        self.delta_e = 0.0
        g0 = 0.0
        for n in range(1, self.n_L + 1):
            g0 = self.residual(float(np.dot(M3, SM3)))
            L, M3, SM3, energy1, inflex, zeta = self.line_min(L, M3, search, SM3, energy0)
            if inflex:
                return L, n, True, energy0, g0
            L = self.correct(L)
            r = self.lnv(L)
            M3, phi = r["M3"], r["phi"]
            energy1 = SPIN * r["energy"]
            SM3 = self.precondition(M3, self.T)
            energy0 = energy1
            search = -SM3
            Sphi, en, nn = self.project(SM3, M3, phi)
            if abs(nn) > RD_ERR:
                search = search + (en / nn) * Sphi
            g0 = self.residual(float(np.dot(M3, SM3)))
            self.trace.append(("early", energy0, g0))
            if abs(zeta) < self.lin_tol:
                return L, n, False, energy0, g0
        return L, self.n_L, False, energy0, g0

    # -- lateDM ------------------------------------------------------------------
    def _projected(self, L):
        r = self.lnv(L)
        M3, phi = r["M3"], r["phi"]
        energy = SPIN * r["energy"]
        SM3 = self.precondition(M3, self.T)
        Sphi, en, nn = self.project(SM3, M3, phi)
        g_raw = self.residual(float(np.dot(M3, SM3)))
        if abs(nn) > RD_ERR:
            SM3 = SM3 - (en / nn) * Sphi
            M3 = M3 - (en / nn) * phi
        return M3, SM3, energy, g_raw

    def late(self, L):
        """Returns (L, iterations, done, energy, residual)."""
        mp = self.mp
        Ls, Gs, SGs = [None] * mp, [None] * mp, [None] * mp
        M3, SM3, energy0, _ = self._projected(L)
        SGs[0], Gs[0], Ls[0] = -SM3, -M3, L.copy()
        g0 = self.residual(float(np.dot(M3, SM3)))
        done = self.n_L == 0
        energy1, g1 = energy0, g0
        stuck = 0
        for n in range(1, self.n_L + 1):
            npmod = n % mp + 1
            pul_mx = min(n + 1, mp)
            step = self.delta_e / (g0 + RD_ERR) if self.glob else self.delta_e / (self.n_atoms * g0 + RD_ERR)
            step = min(max(abs(step), self.min_step), self.max_step)
            prev = npmod - 1 if npmod > 1 else pul_mx
            L = Ls[prev - 1] + step * SGs[prev - 1]
            L = self.correct(L)
            M3, SM3, energy1, _ = self._projected(L)
            SGs[npmod - 1], Gs[npmod - 1], Ls[npmod - 1] = -SM3, -M3, L.copy()
            A = np.zeros((pul_mx, pul_mx))
            for i in range(pul_mx):
                for j in range(pul_mx):
                    A[j, i] = SPIN * float(np.dot(Gs[j], SGs[i]))
            al = lo.do_pulay(npmod, A)
            L = np.zeros_like(L)
            for i in range(pul_mx):
                L = L + al[i] * Ls[i]
            L = self.correct(L)
            M3, SM3, energy1, _ = self._projected(L)
            self.delta_e = energy1 - energy0
            g1 = self.residual(float(np.dot(M3, SM3)))
            self.trace.append(("late", energy1, g1))
            if g1 < self.tol:
                return L, n, True, energy1, g1
            if g1 > 2.0 * g0:
                return L, n, False, energy1, g1
            if g1 > 0.99 * g0:
                stuck += 1
                if stuck > 5:
                    return L, n, False, energy1, g1
            SGs[npmod - 1], Gs[npmod - 1], Ls[npmod - 1] = -SM3, -M3, L.copy()
            g0 = g1
            energy0 = energy1
        return L, self.n_L, done, energy1, g1

    # -- FindMinDM -----------------------------------------------------------------
    def run(self, L, delta_e=0.0, max_rounds=10):
        self.delta_e = delta_e
        res = dict(done=False, early_iterations=0, late_iterations=0, inflexions=0)
        g = 0.0
        for _ in range(max_rounds):
            L, n, inflex, energy, g = self.early(L)
            res["early_iterations"] += n
            if inflex:
                res["inflexions"] += 1
                continue
            L, n, done, energy, g = self.late(L)
            res["late_iterations"] += n
            if done:
                res["done"] = True
                break
            if n == self.n_L:
                break
        r = self.lnv(L)
        res.update(L=L, K=r["K"], energy=SPIN * r["energy"], electrons=SPIN * r["electrons"], residual=g)
        return res
