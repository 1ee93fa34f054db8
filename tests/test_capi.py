"""The C-ABI library loads and exports every function include/*.h declares (CPU only)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import cqmm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(cqmm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = cqmm.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_layer_binds_every_export():
    assert declared_functions() <= set(cqmm.EXPORTS)


def test_errors_are_reported_not_raised_in_c():
    lib = cqmm.lib()
    h = ctypes.c_void_p()
    # Bad arguments return a status and a message instead of crashing.
    st = lib.cqmm_sys_create(0, None, None, None, None, 1.0, ctypes.byref(h))
    assert st != 0
    assert b"cqmm_sys_create" in lib.cqmm_last_error()


def test_random_structure_small_dmin_and_bad_dmin():
    # dmin 0 (no exclusion) must not size the rejection cell list by 1/dmin.
    xyz, edge = cqmm.random_structure(100, 0.05, 0.0, 4)
    assert xyz.shape == (100, 3) and np.all((xyz >= 0) & (xyz < edge))
    with pytest.raises(cqmm.CqmmError, match="bad argument"):
        cqmm.random_structure(10, 0.05, -1.0, 4)


def _overcommit_refuses_huge():
    try:
        return open("/proc/sys/vm/overcommit_memory").read().strip() in ("0", "2")
    except OSError:
        return False


@pytest.mark.skipif(not _overcommit_refuses_huge(), reason="needs an allocator that refuses a 400 GB request")
def test_host_allocation_failure_is_a_status_not_an_abort():
    # 2^31-1 atoms: the rejection cell list alone asks for ~400 GB. The entry
    # point must return CQMM_ERR_NOMEM (the output buffer is never written:
    # the failure comes before any atom is placed).
    lib = cqmm.lib()
    xyz = (ctypes.c_double * 3)()
    edge = ctypes.c_double()
    st = lib.cqmm_random_structure(2**31 - 1, ctypes.c_double(0.05), ctypes.c_double(0.0), ctypes.c_uint64(1),
                                   xyz, ctypes.byref(edge))
    assert st == -3
    assert b"cqmm_random_structure: host allocation failed" in lib.cqmm_last_error()
