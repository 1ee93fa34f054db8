"""Writes the small fixtures of tests/golden/ from the pinned oracle.

Each fixture holds the seeded inputs of one product (structure seed, sizes,
ranges, mult type) and the oracle's outputs: the triplet list in the default
kernels' loop order and the result array. They guard the oracle and the host
index mirror against drift and give GPU tests a reference that needs no CPU
kernel run. Regenerate with:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "conquest-release_amd"), os.path.join(ROOT, "tests")]

import oracle  # noqa: E402
from cases import Case  # noqa: E402

FIXTURES = {
    # name: (natoms, nsf, ra, rb, rc, mult_type, seed)
    "t1_nsf4_full": (20, 4, 4.0, 5.0, 9.0, 1, 101),
    "t1_mixed_trunc": (24, ((4, .5), (9, .25), (13, .25)), 4.5, 5.0, 6.0, 1, 102),
    "t2_nsf13": (16, 13, 5.0, 5.0, 4.0, 2, 103),
}


def make(name, spec):
    natoms, nsf, ra, rb, rc, mt, seed = spec
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=seed)
    a, b, c = case.values()
    o, p, q = oracle.triplets(case.desc)
    oracle.mat_mult(case.desc, a, b, c)
    out = c if mt == 1 else a
    np.savez_compressed(os.path.join(HERE, name + ".npz"), natoms=natoms, nsf=np.asarray(case.nsf),
                        ranges=np.array([ra, rb, rc]), mult_type=mt, seed=seed, trip_out=o.astype(np.int32), trip_p=p.astype(np.int32), trip_q=q.astype(np.int32),
                        result=out)


if __name__ == "__main__":
    for n, s in FIXTURES.items():
        make(n, s)
        print("wrote", n)
