"""One-call purification solves and the request loop (include/cqmm_serve.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import check, init, lib


class PurifyParams(C.Structure):
    _fields_ = [("r_S", C.c_double), ("r_H", C.c_double), ("r_L", C.c_double), ("density", C.c_double),
                ("dmin", C.c_double), ("s_amp", C.c_double), ("s_decay", C.c_double), ("h_amp", C.c_double),
                ("h_decay", C.c_double), ("h_onsite", C.c_double), ("h_split", C.c_double), ("h_shift", C.c_double),
                ("precision", C.c_int32), ("max_iter", C.c_int32), ("fp32_tol", C.c_double),
                ("symmetric", C.c_int32)]


class PurifyResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("fp32_iterations", C.c_int32),
                ("stop_reason", C.c_int32), ("hotelling_iterations", C.c_int32), ("band_energy", C.c_double),
                ("n_electrons", C.c_double), ("ne_error", C.c_double), ("setup_s", C.c_double),
                ("solve_s", C.c_double), ("hbm_estimate", C.c_int64), ("natoms", C.c_int32), ("n_orb", C.c_int32),
                ("hbm_peak", C.c_int64), ("mcw_s", C.c_double), ("mcw_flops", C.c_double)]


class ServeRequest(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("nsf", C.c_int32), ("seed", C.c_uint64), ("arrival_s", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("hbm_budget", C.c_int64), ("params", PurifyParams), ("log_progress", C.c_int32),
                ("batch", C.c_int32), ("max_batch", C.c_int32)]


class ServeResult(C.Structure):
    _fields_ = [("r", PurifyResult), ("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
                ("slot", C.c_int32)]


def _bind():
    L = lib()
    if getattr(L, "_serve_bound", False):
        return L
    L.cqmm_purify_default_params.argtypes = [C.POINTER(PurifyParams)]
    L.cqmm_purify_default_params.restype = None
    L.cqmm_purify_hbm_estimate.argtypes = [C.c_int32, C.c_int32, C.POINTER(PurifyParams)]
    L.cqmm_purify_hbm_estimate.restype = C.c_int64
    L.cqmm_purify.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.POINTER(PurifyParams), C.POINTER(PurifyResult)]
    L.cqmm_serve.argtypes = [C.POINTER(ServeRequest), C.c_int32, C.POINTER(ServeConfig), C.POINTER(ServeResult),
                             C.POINTER(C.c_double)]
    L._serve_bound = True
    return L


def default_params(**kw):
    p = PurifyParams()
    _bind().cqmm_purify_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def hbm_estimate(natoms, nsf, params=None):
    params = params or default_params()
    return _bind().cqmm_purify_hbm_estimate(natoms, nsf, C.byref(params))


def result_dict(r):
    return {k: getattr(r, k) for k, _ in PurifyResult._fields_}


def purify(natoms, nsf, seed, params=None):
    init()
    params = params or default_params()
    r = PurifyResult()
    check(_bind().cqmm_purify(natoms, nsf, seed, C.byref(params), C.byref(r)), "cqmm_purify")
    return result_dict(r)


def poisson_stream(n, rate, seed, nmin=64, nmax=2048, nsf_choices=(4, 13)):
    """n requests: atoms log-uniform in [nmin, nmax], nsf drawn from nsf_choices,
    exponential inter-arrival times at `rate` per second (seeded)."""
    rng = np.random.default_rng(seed)
    natoms = np.exp(rng.uniform(np.log(nmin), np.log(nmax), n)).astype(np.int64)
    nsf = rng.choice(np.asarray(nsf_choices), n)
    arrival = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    seeds = rng.integers(1, 2 ** 31, n)
    return [dict(natoms=int(a), nsf=int(b), seed=int(s), arrival_s=float(t))
            for a, b, s, t in zip(natoms, nsf, seeds, arrival)]


def serve(requests, slots=4, hbm_budget=200 << 30, params=None, log_progress=False, batch=0, max_batch=16):
    """Run the request loop; returns (list of per-request dicts, wall seconds).
    batch 0: each slot solves its request alone; 1: McWeeny steps of all
    prepared requests batched (grouped launches, hipGraph); 2: as 1 without graphs."""
    init()
    params = params or default_params()
    n = len(requests)
    reqs = (ServeRequest * max(n, 1))()
    for i, q in enumerate(requests):
        reqs[i] = ServeRequest(q["natoms"], q["nsf"], q["seed"], q["arrival_s"])
    cfg = ServeConfig(slots, int(hbm_budget), params, int(bool(log_progress)), int(batch), int(max_batch))
    out = (ServeResult * max(n, 1))()
    wall = C.c_double()
    check(_bind().cqmm_serve(reqs, n, C.byref(cfg), out, C.byref(wall)), "cqmm_serve")
    res = []
    for i in range(n):
        d = result_dict(out[i].r)
        d.update(arrival_s=out[i].arrival_s, admit_s=out[i].admit_s, done_s=out[i].done_s, slot=out[i].slot,
                 latency_s=out[i].done_s - out[i].arrival_s)
        res.append(d)
    return res, wall.value
