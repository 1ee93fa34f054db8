"""Device-resident matrix API and density-matrix solvers (host side).

``Context`` mirrors CONQUEST's mult_module (src/mult_module.f90): matrices are
integer handles into HBM, temporaries form a LIFO stack, and each product runs
through the plan registered for its mult id. ``Purifier`` sets up the McWeeny
workload (ranges, products, synthetic S and H) on a System and runs
Hotelling S^-1 -> InitMcW -> McWMin on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (CqmmError, DmmResult, DmmSetup, HotellingSetup, InitMcwSetup, LnvSetup, MatrixIndex, McwResult,
               McwSetup, System, check, init, lib)

# mult ids (names as in src/mult_module.f90:128-160)
L_S_LS, LS_L_LSL, LSL_SL_L, T_S_TS, TS_T_T, T_H_TH, TH_T_L = 1, 2, 3, 4, 5, 6, 7
L_H_LH, HL_S_LSL, HLS_LS_L, SL_S_SLS, SLS_LS_L, LSL_SL_H, T_L_TL, TL_T_L = 8, 9, 10, 11, 12, 13, 14, 15


class Context:
    """mult_module mirror over one System: ranges by System range id."""

    def __init__(self, system: System):
        init()
        self.sys = system
        h = C.c_void_p()
        check(lib().cqmm_ctx_create(C.byref(h)), "cqmm_ctx_create")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib().cqmm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up ----------------------------------------------------------
    def add_range(self, rid, transpose=False):
        idx = MatrixIndex()
        check(lib().cqmm_sys_matrix_index(self.sys.h, rid, C.byref(idx)), "cqmm_sys_matrix_index")
        check(lib().cqmm_ctx_add_range(self.h, rid, C.byref(idx)), "cqmm_ctx_add_range")
        if transpose:
            p, q, a, b = self.sys.transpose_pairs(rid)
            ptr = lambda x, t: x.ctypes.data_as(C.POINTER(t))
            check(lib().cqmm_ctx_set_transpose(self.h, rid, len(p), ptr(p, C.c_int64), ptr(q, C.c_int64),
                                               ptr(a, C.c_int32), ptr(b, C.c_int32)), "cqmm_ctx_set_transpose")

    def add_mult(self, mult_id, mult_type, ra, rb, rc, symmetric=False):
        """mult(mult_id): type 1 C(rc) = A(ra).B(rb); type 2 A(ra) = C(rc).B(rb)^T.
        symmetric: the result is symmetric; half of its blocks are computed and
        the others written as their transposes (cqmm_sys_plan_create_symmetric)."""
        plan = C.c_void_p()
        fn = lib().cqmm_sys_plan_create_symmetric if symmetric else lib().cqmm_sys_plan_create
        check(fn(self.sys.h, mult_type, ra, rb, rc, C.byref(plan)), "cqmm_sys_plan_create")
        check(lib().cqmm_ctx_add_mult(self.h, mult_id, ra, rb, rc, plan), "cqmm_ctx_add_mult")

    # -- matrices ----------------------------------------------------------
    def allocate(self, rid):
        h = C.c_int32()
        check(lib().cqmm_matrix_allocate(self.h, rid, C.byref(h)), "cqmm_matrix_allocate")
        return h.value

    def free(self, h):
        check(lib().cqmm_matrix_free(self.h, h), "cqmm_matrix_free")

    def length(self, h):
        n = C.c_int64()
        check(lib().cqmm_matrix_length(self.h, h, C.byref(n)), "cqmm_matrix_length")
        return n.value

    def set(self, h, values):
        v = np.ascontiguousarray(values, np.float64)
        if v.size != self.length(h):
            raise CqmmError("matrix_set: length mismatch")
        check(lib().cqmm_matrix_set(self.h, h, v.ctypes.data), "cqmm_matrix_set")

    def get(self, h):
        v = np.empty(self.length(h), np.float64)
        check(lib().cqmm_matrix_get(self.h, h, v.ctypes.data), "cqmm_matrix_get")
        return v

    def product(self, A, B, Cm, mult_id):
        check(lib().cqmm_matrix_product(self.h, A, B, Cm, mult_id), "cqmm_matrix_product")

    def sum(self, alpha, A, beta, B):
        check(lib().cqmm_matrix_sum(self.h, alpha, A, beta, B), "cqmm_matrix_sum")

    def scale(self, alpha, A):
        check(lib().cqmm_matrix_scale(self.h, alpha, A), "cqmm_matrix_scale")

    def product_trace(self, A, B):
        r = C.c_double()
        check(lib().cqmm_matrix_product_trace(self.h, A, B, C.byref(r)), "cqmm_matrix_product_trace")
        return r.value

    def trace(self, A):
        r = C.c_double()
        check(lib().cqmm_matrix_trace(self.h, A, C.byref(r)), "cqmm_matrix_trace")
        return r.value

    def symmetrise(self, A):
        check(lib().cqmm_matrix_symmetrise(self.h, A), "cqmm_matrix_symmetrise")

    def transpose(self, A, B):
        check(lib().cqmm_matrix_transpose(self.h, A, B), "cqmm_matrix_transpose")

    @property
    def product_precision(self):
        p = C.c_int32()
        check(lib().cqmm_ctx_get_product_precision(self.h, C.byref(p)), "cqmm_ctx_get_product_precision")
        return p.value

    @product_precision.setter
    def product_precision(self, precision):
        """0 FP64; 1, 2, 3 FP32 MFMA products (FP32 sums, FP64 block sums, FP32 chunk runs); 4 split-BF16 x3
        products where the plan's kernel has them (else as 2)."""
        check(lib().cqmm_ctx_set_product_precision(self.h, int(precision)), "cqmm_ctx_set_product_precision")


def mcw_result_dict(res):
    return dict(iterations=res.iterations, fp32_iterations=res.fp32_iterations, stop_reason=res.stop_reason,
                restored=res.restored, energy=res.energy, electrons=res.electrons, c=res.c,
                seconds=res.seconds, energy_trace=list(res.energy_trace[:min(res.iterations, 256)]))


BATCH_NO_GRAPH, BATCH_NO_GROUPING = 1, 2


def mcweeny_batch(purifiers, precision=0, max_iter=100, n_final_fp64=2, fp32_tol=1e-6, flags=0):
    """cqmm_mcweeny_batch over several Purifiers (one context each), in lockstep.
    Returns a list of (status, result dict)."""
    n = len(purifiers)
    ctxs = (C.c_void_p * n)(*[p.ctx.h.value for p in purifiers])
    setups = (McwSetup * n)(*[p.mcw_setup(precision, max_iter, n_final_fp64, fp32_tol) for p in purifiers])
    H = (C.c_int32 * n)(*[p.matH for p in purifiers])
    S = (C.c_int32 * n)(*[p.matS for p in purifiers])
    Lm = (C.c_int32 * n)(*[p.matL for p in purifiers])
    res = (McwResult * n)()
    status = (C.c_int32 * n)()
    check(lib().cqmm_mcweeny_batch(n, ctxs, setups, H, S, Lm, flags, res, status), "cqmm_mcweeny_batch")
    return [(status[i], mcw_result_dict(res[i])) for i in range(n)]


class Purifier:
    """McWeeny purification workload on one System.

    Ranges (Angstrom): S, H, L (the density-matrix range), T = S^-1 range;
    derived LS = L+S, LSL = 2L+S, TS = T+S, TH = T+H. S and H follow the
    synthetic recipe: S = I + symmetrised s_amp*N(0,1)*exp(-r/s_decay) off the
    onsite blocks; H onsite diagonal U(-h_onsite, h_onsite) split by -/+ h_split
    between the lower and upper half of the orbitals, symmetrised
    h_amp*N(0,1)*exp(-r/h_decay) elsewhere, plus h_shift*S. The h_shift*S term
    moves every eigenvalue of H c = e S c by h_shift and leaves the eigenvectors
    (so the gap and the density matrix) unchanged; it keeps the onsite diagonal
    positive, which GetHLimits' bounds assume (they add |H| over onsite blocks,
    src/McWeeny.f90:1066-1076).
    """

    def __init__(self, xyz, box, nsf, r_S=8.0, r_H=8.0, r_L=12.0, r_T=None, grid=None, seed=1,
                 s_amp=0.1, s_decay=1.5, h_amp=-0.2, h_decay=2.0, h_onsite=0.5, h_split=1.5, h_shift=3.0,
                 device=True, lnv=False, symmetric=False):
        r_T = r_S if r_T is None else r_T
        self.r = dict(S=r_S, H=r_H, L=r_L, T=r_T, LS=r_L + r_S, LSL=2 * r_L + r_S, TS=r_T + r_S, TH=r_T + r_H)
        self.lnv_ranges = lnv
        # LSL = LS.L and LSLSL = LSL.(LS)^T are symmetric: with `symmetric`
        # their plans compute half of the blocks and mirror the rest.
        self.symmetric = symmetric
        if lnv:  # LNV / DMM ranges (src/dimens_module.f90:342-355)
            self.r.update(LH=r_L + r_H, SLS=2 * r_S + r_L, TL=r_T + r_L)
        rcover = max(self.r["LSL"], self.r["L"] + self.r["LS"], self.r["LS"] + self.r["L"], self.r["TH"] + r_T,
                     2 * r_T, r_T + r_S, r_T + r_H)
        if lnv:
            rcover = max(rcover, self.r["LH"] + r_S, r_H + self.r["LS"], self.r["LS"] + r_S, r_T + r_L)
        self.sys = System(xyz, box, nsf, rcover, grid)
        self.natoms = self.sys.natoms
        self.nsf = np.broadcast_to(np.asarray(nsf, np.int32), (self.natoms,))
        self.n_orb = int(self.nsf.sum())
        self.rid = {}
        for name in ("S", "H", "L", "T", "LS", "LSL", "TS", "TH") + (("LH", "SLS", "TL") if lnv else ()):
            rc = self.r[name]
            same = [k for k, v in self.rid.items() if self.r[k] == rc]
            self.rid[name] = self.rid[same[0]] if same else self.sys.add_range(rc)
        self.seed = seed
        R = self.rid
        self.S_vals = self.sys.fill(R["S"], 1, seed, 11, (s_amp, s_decay))
        self.H_vals = self.sys.fill(R["H"], 2, seed, 12, (h_amp, h_decay, h_onsite, h_split))
        if h_shift:
            self.H_vals += h_shift * self._map_values(R["S"], R["H"], self.S_vals)
        self.ctx = None
        if device:
            self.to_device()

    def _map_values(self, r_from, r_to, vals):
        """Values of range r_from placed on the blocks of range r_to (zero elsewhere)."""
        if r_from == r_to:
            return vals.copy()
        bf, bt = self.sys.blocks(r_from), self.sys.blocks(r_to)
        key = lambda b: ((b["gid_i"].astype(np.int64) * self.natoms + b["gid_j"]) * 4096 +
                         (b["shift"][:, 0] + 8) * 256 + (b["shift"][:, 1] + 8) * 16 + (b["shift"][:, 2] + 8))
        kf, kt = key(bf), key(bt)
        order = np.argsort(kf)
        pos = np.minimum(np.searchsorted(kf[order], kt), len(kf) - 1)
        hit = kf[order][pos] == kt
        out = np.zeros(self.sys.length(r_to))
        for t in np.nonzero(hit)[0]:
            f = order[pos[t]]
            n = bt["nd1"][t] * bt["nd2"][t]
            out[bt["offset"][t]:bt["offset"][t] + n] = vals[bf["offset"][f]:bf["offset"][f] + n]
        return out

    def to_device(self):
        """Create the device context: ranges, plans, and HBM copies of S and H."""
        self.ctx = Context(self.sys)
        R = self.rid
        trans = {R["L"]} | ({R["LS"], R["LH"], R["LSL"], R["T"]} if self.lnv_ranges else set())
        for name in sorted(set(self.rid.values())):
            self.ctx.add_range(name, transpose=(name in trans))
        R = self.rid
        self.ctx.add_mult(L_S_LS, 1, R["L"], R["S"], R["LS"])
        self.ctx.add_mult(LS_L_LSL, 1, R["LS"], R["L"], R["LSL"], symmetric=self.symmetric)
        self.ctx.add_mult(LSL_SL_L, 2, R["L"], R["LS"], R["LSL"], symmetric=self.symmetric)
        self.ctx.add_mult(T_S_TS, 1, R["T"], R["S"], R["TS"])
        self.ctx.add_mult(TS_T_T, 2, R["T"], R["T"], R["TS"])
        self.ctx.add_mult(T_H_TH, 1, R["T"], R["H"], R["TH"])
        self.ctx.add_mult(TH_T_L, 1, R["TH"], R["T"], R["L"])
        if self.lnv_ranges:  # mult table, src/mult_module.f90:583-875
            self.ctx.add_mult(L_H_LH, 1, R["L"], R["H"], R["LH"])
            self.ctx.add_mult(HL_S_LSL, 1, R["LH"], R["S"], R["LSL"])
            self.ctx.add_mult(HLS_LS_L, 2, R["L"], R["LS"], R["LSL"])
            # SLS = SL.S, SLSLS = SLS.(SL)^T and LSLSL = LSL.(LS)^T are
            # symmetric too; HLS and (2HLS + SLH).(SL)^T are not.
            self.ctx.add_mult(SL_S_SLS, 1, R["LS"], R["S"], R["SLS"], symmetric=self.symmetric)
            self.ctx.add_mult(SLS_LS_L, 2, R["L"], R["LS"], R["SLS"], symmetric=self.symmetric)
            self.ctx.add_mult(LSL_SL_H, 2, R["H"], R["LS"], R["LSL"], symmetric=self.symmetric)
            self.ctx.add_mult(T_L_TL, 1, R["T"], R["L"], R["TL"])
            self.ctx.add_mult(TL_T_L, 2, R["L"], R["T"], R["TL"])
        self.matS = self.ctx.allocate(R["S"])
        self.matH = self.ctx.allocate(R["H"])
        self.matT = self.ctx.allocate(R["T"])
        self.matL = self.ctx.allocate(R["L"])
        self.ctx.set(self.matS, self.S_vals)
        self.ctx.set(self.matH, self.H_vals)

    def hotelling(self, max_iter=50, tol=1e-10):
        su = HotellingSetup(self.rid["T"], self.rid["S"], self.rid["TS"], T_S_TS, TS_T_T, max_iter, tol)
        it = C.c_int32()
        om = C.c_double()
        check(lib().cqmm_hotelling(self.ctx.h, C.byref(su), self.matS, self.matT, C.byref(it), C.byref(om)),
              "cqmm_hotelling")
        return it.value, om.value

    def init_mcweeny(self, n_electrons=None):
        ne = float(self.n_orb // 2) if n_electrons is None else n_electrons
        self.ctx.scale(0.0, self.matL)
        su = InitMcwSetup(self.rid["T"], self.rid["S"], self.rid["H"], self.rid["L"], self.rid["TS"], self.rid["TH"],
                          T_S_TS, T_H_TH, TH_T_L, ne)
        ml = (C.c_double * 2)()
        check(lib().cqmm_init_mcweeny(self.ctx.h, C.byref(su), self.matH, self.matS, self.matT, self.matL, ml),
              "cqmm_init_mcweeny")
        return ml[0], ml[1]

    def mcw_setup(self, precision=0, max_iter=100, n_final_fp64=2, fp32_tol=1e-6):
        return McwSetup(self.rid["L"], self.rid["S"], self.rid["H"], self.rid["LS"], self.rid["LSL"], L_S_LS,
                        LS_L_LSL, LSL_SL_L, max_iter, precision, n_final_fp64, fp32_tol)

    def mcweeny(self, precision=0, max_iter=100, n_final_fp64=2, fp32_tol=1e-6):
        """McWMin from the current L (cqmm_mcweeny). precision: 0 all FP64; 1 FP32 MFMA with FP64 block
        sums, 2 FP32 sums, 3 FP32 chunk runs, 4 split-BF16 x3, 5 split-BF16 x3 then FP32 MFMA with FP64
        sums, each followed by FP64 steps (at least n_final_fp64)."""
        su = McwSetup(self.rid["L"], self.rid["S"], self.rid["H"], self.rid["LS"], self.rid["LSL"], L_S_LS, LS_L_LSL,
                      LSL_SL_L, max_iter, precision, n_final_fp64, fp32_tol)
        res = McwResult()
        check(lib().cqmm_mcweeny(self.ctx.h, C.byref(su), self.matH, self.matS, self.matL, C.byref(res)),
              "cqmm_mcweeny")
        return mcw_result_dict(res)

    # -- two spin channels (src/McWeeny.f90, nspin = 2) ----------------------
    def add_spin_down(self, H_dn_vals):
        """Second spin channel: its own H (values on the H range) and L; both
        channels share S and T (support functions that do not depend on spin)."""
        self.H_dn_vals = np.ascontiguousarray(H_dn_vals, np.float64)
        self.matH_dn = self.ctx.allocate(self.rid["H"])
        self.matL_dn = self.ctx.allocate(self.rid["L"])
        self.ctx.set(self.matH_dn, self.H_dn_vals)

    def _pair(self, a, b):
        return (C.c_int32 * 2)(a, b)

    def init_mcweeny_spin(self, ne_up, ne_dn, fixed):
        """InitMcW for both channels: ((mubar_up, lambda_up), (mubar_dn, lambda_dn))."""
        su = InitMcwSetup(self.rid["T"], self.rid["S"], self.rid["H"], self.rid["L"], self.rid["TS"], self.rid["TH"],
                          T_S_TS, T_H_TH, TH_T_L, 0.0)
        ml = (C.c_double * 4)()
        ne = (C.c_double * 2)(ne_up, ne_dn)
        check(lib().cqmm_init_mcweeny_spin(self.ctx.h, C.byref(su), int(bool(fixed)), self._pair(self.matH, self.matH_dn),
                                           self._pair(self.matS, self.matS), self._pair(self.matT, self.matT),
                                           self._pair(self.matL, self.matL_dn), ne, ml), "cqmm_init_mcweeny_spin")
        return (ml[0], ml[1]), (ml[2], ml[3])

    def mcweeny_spin(self, fixed, precision=0, max_iter=100, n_final_fp64=2, fp32_tol=1e-6, matS_dn=None):
        """McWMin over both channels: [result of spin up, result of spin down].
        matS_dn: spin down's own overlap (spin-dependent support functions)."""
        su = self.mcw_setup(precision, max_iter, n_final_fp64, fp32_tol)
        res = (McwResult * 2)()
        s_dn = self.matS if matS_dn is None else matS_dn
        check(lib().cqmm_mcweeny_spin(self.ctx.h, C.byref(su), int(bool(fixed)), self._pair(self.matH, self.matH_dn),
                                      self._pair(self.matS, s_dn), self._pair(self.matL, self.matL_dn), res),
              "cqmm_mcweeny_spin")
        return [mcw_result_dict(res[0]), mcw_result_dict(res[1])]

    # -- LNV / DMM (src/mult_module.f90:1478-1823, src/DMMinModule.f90) ----
    def lnv_setup(self):
        R = self.rid
        return LnvSetup(R["L"], R["S"], R["H"], R["LS"], R["LH"], R["LSL"], R["SLS"], L_S_LS, L_H_LH, HL_S_LSL,
                        HLS_LS_L, SL_S_SLS, SLS_LS_L, LS_L_LSL, LSL_SL_H)

    def lnv(self, flags=15):
        """LNV_matrix_multiply on the current L: dict(K, M3, phi, energy, electrons) (per spin)."""
        c = self.ctx
        K, M3, phi = c.allocate(self.rid["H"]), c.allocate(self.rid["L"]), c.allocate(self.rid["L"])
        el, en = C.c_double(), C.c_double()
        su = self.lnv_setup()
        check(lib().cqmm_lnv(c.h, C.byref(su), flags, self.matH, self.matS, self.matL, K, M3, phi, C.byref(el),
                             C.byref(en)), "cqmm_lnv")
        out = dict(K=c.get(K), M3=c.get(M3), phi=c.get(phi), energy=en.value, electrons=el.value)
        for h in (phi, M3, K):
            c.free(h)
        return out

    def dmm_setup(self, n_electrons=None, tolerance=1e-7, n_L_iterations=50, vary_mu=True, maxpulay=5,
                  delta_e=0.0):
        ne = float(self.n_orb // 2) if n_electrons is None else n_electrons
        R = self.rid
        return DmmSetup(self.lnv_setup(), R["T"], R["TL"], T_L_TL, TL_T_L, n_L_iterations, int(vary_mu), maxpulay, 0,
                        self.natoms, tolerance, 1e-3, 0.1, 0.1, ne, delta_e)

    def correct_electrons(self, **kw):
        su = self.dmm_setup(**kw)
        el = C.c_double()
        check(lib().cqmm_correct_electrons(self.ctx.h, C.byref(su), self.matH, self.matS, self.matT, self.matL,
                                           C.byref(el)), "cqmm_correct_electrons")
        return el.value

    def dmm(self, product_precision=0, **kw):
        """FindMinDM from the current L; K receives 3LSL - 2LSLSL. product_precision
        (0 FP64; 1, 2, 3 FP32 modes; 4 split-BF16 x3) applies to the bulk products of the
        minimisation; the electron-number correction and the final K, energy and
        electron count use FP64 products."""
        su = self.dmm_setup(**kw)
        res = DmmResult()
        K = self.ctx.allocate(self.rid["H"])
        saved = self.ctx.product_precision
        self.ctx.product_precision = product_precision
        try:
            check(lib().cqmm_dmm(self.ctx.h, C.byref(su), self.matH, self.matS, self.matT, self.matL, K,
                                 C.byref(res)), "cqmm_dmm")
        finally:
            self.ctx.product_precision = saved
        out = dict(done=res.done, early_iterations=res.early_iterations, late_iterations=res.late_iterations,
                   inflexions=res.inflexions, band_energy=res.energy, n_electrons=res.electrons,
                   residual=res.residual, seconds=res.seconds, K=self.ctx.get(K))
        self.ctx.free(K)
        return out

    def solve(self, precision=0, max_iter=100, n_electrons=None, n_final_fp64=2, fp32_tol=1e-6):
        """Hotelling -> InitMcW -> McWMin; band energy = 2 Tr[rho H] (spin-unpolarised)."""
        hot = self.hotelling()
        mu_l = self.init_mcweeny(n_electrons)
        r = self.mcweeny(precision, max_iter, n_final_fp64, fp32_tol)
        r["band_energy"] = 2.0 * r["energy"]
        r["n_electrons"] = 2.0 * r["electrons"]
        r["hotelling_iterations"], r["hotelling_omega"] = hot
        r["mubar"], r["lambda"] = mu_l
        return r
