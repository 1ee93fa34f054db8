"""Writes tests/golden/si8_test002.json: the structure of the reference's own
Order-N test (testsuite/test_002_bulk_Si_1proc_OrderN): its cell and
fractional coordinates (coords.dat), the support functions per atom implied by
its basis file (Si.ion: sum over shells of zetas * (2l + 1)), the PAO cutoff
radius (S range = twice the largest radius) and the density-matrix range
(Conquest_input DM.L_range), all in bohr. Data only; run once where the
reference tree is present:  python tests/golden/make_si8.py [reference root]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
T = os.path.join(REF, "testsuite", "test_002_bulk_Si_1proc_OrderN")

with open(os.path.join(T, "coords.dat")) as f:
    lines = [l.split() for l in f if l.strip()]
cell = [[float(x) for x in lines[k][:3]] for k in range(3)]
natoms = int(lines[3][0])
frac = [[float(x) for x in lines[4 + k][:3]] for k in range(natoms)]

ion = open(os.path.join(T, "Si.ion")).read()
info = ion[ion.index("<Conquest_basis_info>"):ion.index("</Conquest_basis_info>")]
nsf, radii = 0, []
for l, z in re.findall(r"l\s*=\s*(\d+),\s*(\d+)\s*zetas", info):
    nsf += int(z) * (2 * int(l) + 1)
for r in re.findall(r"Radii:\s*([0-9. ]+)", info):
    radii += [float(x) for x in r.split()]

inp = open(os.path.join(T, "Conquest_input")).read()
l_range = float(re.search(r"DM\.L_range\s+([0-9.]+)", inp).group(1))

out = dict(source="testsuite/test_002_bulk_Si_1proc_OrderN (coords.dat, Si.ion, Conquest_input)",
           cell_bohr=cell, frac=frac, nsf=nsf, pao_radius_bohr=max(radii), s_range_bohr=2 * max(radii),
           l_range_bohr=l_range)
with open(os.path.join(HERE, "si8_test002.json"), "w") as f:
    json.dump(out, f, indent=1)
print(out)
