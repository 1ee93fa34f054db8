"""Small seeded systems shaped like the benchmark workloads (sizes the CPU oracle
finishes in seconds). Lengths in Angstrom."""
from __future__ import annotations

import json
import os

import numpy as np

import cqmm


def species_nsf(natoms, mix, seed):
    """nsf per atom: an int, or a mix like ((4, .5), (9, .25), (13, .25))."""
    if isinstance(mix, int):
        return np.full(natoms, mix, np.int32)
    rng = np.random.default_rng(seed + 7919)
    vals = np.array([m[0] for m in mix], np.int32)
    probs = np.array([m[1] for m in mix], np.float64)
    return vals[rng.choice(len(vals), size=natoms, p=probs / probs.sum())]


class Case:
    """A system with ranges and one product.

    mult_type 1: c(rc) = a(ra).b(rb); mult_type 2: a(ra) = c(rc).b(rb)^T."""

    def __init__(self, natoms, nsf, ra, rb, rc, mult_type=1, seed=1, density=0.05, grid=None, decay=2.0,
                 structure=None):
        if structure is None:
            self.xyz, self.edge = cqmm.random_structure(natoms, density, 2.0, seed)
        else:                                   # a fixed (xyz, cubic edge) in the ranges' length unit
            self.xyz, self.edge = np.asarray(structure[0], np.float64), float(structure[1])
            assert len(self.xyz) == natoms
        self.nsf = species_nsf(natoms, nsf, seed)
        rcover = max(ra + rb, rc, ra, rb)
        if grid is None:
            grid = cqmm.default_grid(self.edge, 4.0)
        self.grid = grid
        self.sys = cqmm.System(self.xyz, self.edge, self.nsf, rcover, grid)
        self.ra, self.rb, self.rc = ra, rb, rc
        self.ida = self.sys.add_range(ra)
        self.idb = self.sys.add_range(rb) if rb != ra else self.ida
        self.idc = self.sys.add_range(rc) if rc not in (ra, rb) else (self.ida if rc == ra else self.idb)
        self.mult_type = mult_type
        self.desc = self.sys.mult_desc(mult_type, self.ida, self.idb, self.idc)
        self.seed = seed
        self.decay = decay

    def values(self, tag_a=1, tag_b=2, tag_c=3):
        """Inputs as mat_mult receives them: (a, b, c) with the output zeroed."""
        s = self.sys
        b = s.fill(self.idb, 0, self.seed, tag_b, (self.decay,))
        if self.mult_type == 1:
            a = s.fill(self.ida, 0, self.seed, tag_a, (self.decay,))
            c = np.zeros(s.length(self.idc))
        else:
            c = s.fill(self.idc, 0, self.seed, tag_c, (self.decay,))
            a = np.zeros(s.length(self.ida))
        return a, b, c


def cluster_structure(natoms, radius=3.5, dmin=1.8, box=50.0, seed=1):
    """An isolated cluster (no periodic image within any range used with it):
    natoms uniform in a sphere of `radius` at the centre of a cubic cell."""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < natoms:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) > radius:
            continue
        if all(np.linalg.norm(p - q) >= dmin for q in pts):
            pts.append(p)
    return np.array(pts) + box / 2.0, box


# The structure of the reference's Order-N test 002 (8 Si atoms, 10.36 bohr
# cubic cell, 4 support functions per atom, S range 14.24 bohr, L range 16
# bohr): tests/golden/si8_test002.json, written by tests/golden/make_si8.py.
SI8 = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "si8_test002.json")))

# Products of the LNV step on it: L.S (full), (L.S).L truncated to L's
# pattern, and the type-2 form that builds a result on L's pattern from an
# operand on the L+S range.
SI8_PRODUCTS = [
    (SI8["l_range_bohr"], SI8["s_range_bohr"], SI8["l_range_bohr"] + SI8["s_range_bohr"], 1),
    (SI8["l_range_bohr"] + SI8["s_range_bohr"], SI8["l_range_bohr"], SI8["l_range_bohr"], 1),
    (SI8["l_range_bohr"], SI8["s_range_bohr"], SI8["l_range_bohr"] + SI8["s_range_bohr"], 2),
]


def si8_case(ra, rb, rc, mt, seed=3):
    edge = SI8["cell_bohr"][0][0]
    xyz = np.array(SI8["frac"]) * edge
    return Case(8, SI8["nsf"], ra, rb, rc, mt, seed=seed, grid=(2, 2, 2), structure=(xyz, edge))
