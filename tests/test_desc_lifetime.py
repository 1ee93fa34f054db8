"""CPU test: a product descriptor from System.mult_desc stays valid after the
System object goes out of scope (its arrays live in the system's memory; the
descriptor holds a reference), so the oracle can read it later."""
import gc

import numpy as np

import cqmm
import oracle


def make_desc():
    xyz, edge = cqmm.random_structure(24, 0.05, 2.0, 3)
    sysm = cqmm.System(xyz, edge, 4, 10.0, cqmm.default_grid(edge, 4.0))
    ra, rb = sysm.add_range(4.0), sysm.add_range(6.0)
    desc = sysm.mult_desc(1, ra, rb, rb)
    a, b = sysm.fill(ra, 0, 3, 1, (2.0,)), sysm.fill(rb, 0, 3, 2, (2.0,))
    c = np.zeros(sysm.length(rb))
    ref = c.copy()
    oracle.mat_mult(desc, a, b, ref)
    return desc, a, b, c, ref


def test_descriptor_outlives_its_system_object():
    desc, a, b, c, ref = make_desc()
    gc.collect()
    oracle.mat_mult(desc, a, b, c)
    np.testing.assert_array_equal(c, ref)
