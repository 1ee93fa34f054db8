"""CPU test of the transpose pairs of a range (cqmm_sys_transpose_pairs), which
symmetrise_L, matrix_transpose and symmetric-result plans use: block (i, j, s)
and block (j, i, -s), found here by a dictionary over the block list, in storage
order of the first block, with the first block's dimensions; periodic images
and mixed block sizes included."""
import numpy as np
import pytest

import cqmm
from cqmm.dm import Purifier


@pytest.mark.parametrize("nsf", [4, "mix"])
def test_transpose_pairs_match_block_list(nsf):
    xyz, edge = cqmm.random_structure(48, 0.05, 2.0, 9)
    if nsf == "mix":
        nsf = np.random.default_rng(9).choice([4, 9, 13], size=48).astype(np.int32)
    pu = Purifier(xyz, edge, nsf, r_S=4.5, r_H=4.5, r_L=6.0, grid=(2, 2, 2), seed=9, device=False)
    for name in ("S", "L", "LSL"):
        rid = pu.rid[name]
        b = pu.sys.blocks(rid)
        where = {(gi, gj, tuple(sh)): off for gi, gj, sh, off in
                 zip(b["gid_i"], b["gid_j"], b["shift"], b["offset"])}
        ref = []
        for gi, gj, sh, off, n1, n2 in zip(b["gid_i"], b["gid_j"], b["shift"], b["offset"], b["nd1"], b["nd2"]):
            f = where[(gj, gi, tuple(-sh))]
            if f >= off:
                ref.append((off + 1, f + 1, n1, n2))
        got = [tuple(int(x) for x in t) for t in zip(*pu.sys.transpose_pairs(rid))]
        assert got == [tuple(int(x) for x in t) for t in ref], name
        if name == "LSL":  # the range exceeds the cell: pairs (i, i, s) with (i, i, -s), s != 0
            k = np.searchsorted(b["offset"], np.array([t[0] for t in got]) - 1)
            assert any(t[0] != t[1] and b["gid_i"][x] == b["gid_j"][x] for t, x in zip(got, k))
