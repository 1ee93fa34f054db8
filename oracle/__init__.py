"""CPU ORACLE — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only as the checker (or the timed CPU baseline). The product path
(libcqmm.so and conquest-release_amd/cqmm) never imports it.

Contents
  * cq_oracle.c  — C restatement of mat_mult + the default and ompGemm kernels
                   (bound here through ctypes).
  * brute.py     — numpy brute force over periodic images: neighbour patterns,
                   the triplet set, and an image-resolved block product.
  * mcweeny.py   — numpy restatement of the McWeeny step and its closed forms.
"""
from __future__ import annotations

import ctypes as C
import glob
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libcqoracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing at {LIB_PATH}; run make -C oracle")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.oracle_mat_mult.argtypes = [vp, vp, vp, vp]
        L.oracle_mat_mult_ompgemm.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.oracle_triplets.argtypes = [vp, vp, vp, vp, C.c_int64]
        L.oracle_triplets.restype = C.c_int64
        L.oracle_flops.argtypes = [vp, C.c_int, C.c_int]
        L.oracle_flops.restype = C.c_double
        L.oracle_set_blas.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        L.oracle_blas_name.restype = C.c_char_p
        L.oracle_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _addr(desc):
    return C.addressof(desc)


def mat_mult(desc, a, b, c):
    """Default-kernel mat_mult; type 1 overwrites c, type 2 overwrites a."""
    r = lib().oracle_mat_mult(_addr(desc), a.ctypes.data, b.ctypes.data, c.ctypes.data)
    if r != 0:
        raise RuntimeError("oracle_mat_mult failed")


def mat_mult_ompgemm(desc, a, b, c, kp_begin=1, kp_end=None):
    """ompGemm mat_mult over halo partitions [kp_begin, kp_end); returns the loop's seconds."""
    if kp_end is None:
        kp_end = desc.ahalo.np_in_halo + 1
    t = C.c_double(0.0)
    r = lib().oracle_mat_mult_ompgemm(_addr(desc), a.ctypes.data, b.ctypes.data, c.ctypes.data, kp_begin, kp_end,
                                      C.byref(t))
    if r != 0:
        raise RuntimeError("oracle_mat_mult_ompgemm failed")
    return t.value


def triplets(desc):
    """(out, p, q) 1-based offsets in the default kernels' loop order."""
    n = lib().oracle_triplets(_addr(desc), None, None, None, 0)
    o = np.zeros(n, np.int64)
    p = np.zeros(n, np.int64)
    q = np.zeros(n, np.int64)
    lib().oracle_triplets(_addr(desc), o.ctypes.data, p.ctypes.data, q.ctypes.data, n)
    return o, p, q


def flops(desc, kp_begin=1, kp_end=None):
    if kp_end is None:
        kp_end = desc.ahalo.np_in_halo + 1
    return lib().oracle_flops(_addr(desc), kp_begin, kp_end)


def use_numpy_openblas():
    """Point the ompGemm baseline's dgemm at numpy's bundled OpenBLAS (ILP64).

    Returns the provider name, or the built-in loop's name if no BLAS is found."""
    import numpy
    libdir = os.path.join(os.path.dirname(os.path.dirname(numpy.__file__)), "numpy.libs")
    cands = sorted(glob.glob(os.path.join(libdir, "libscipy_openblas64_*.so")))
    for path in cands:
        r = lib().oracle_set_blas(path.encode(), b"scipy_dgemm_64_", b"scipy_openblas_set_num_threads64_")
        if r == 0:
            return "OpenBLAS (numpy-bundled, ILP64): " + os.path.basename(path)
    return lib().oracle_blas_name().decode()


def max_threads():
    return lib().oracle_max_threads()
