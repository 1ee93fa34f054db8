"""CPU tests of the matrix files in CONQUEST's store_matrix format
(dump_matrix2, src/store_matrix_module.f90:340-470; read as UpdateMatrix_local
places blocks, src/UpdateInfo_module.f90:1851-2162).

The file layout is checked with a reader written here from the reference's
write statements (record by record), independently of the library's reader;
the geometry of each block (vec_Rij = (R_i - R_j image)/cell) against the
atom positions; round trips binary and formatted, one and two spins; and a
file read into a different range keeps exactly the blocks both ranges hold.
No CONQUEST-written file exists in the reference for this path (parity with
a real dump is unpinned); the format follows the write statements.
"""
import os
import struct

import numpy as np
import pytest

import cqmm


def records(path):
    """Fortran unformatted sequential records: (4-byte length, payload, 4-byte length)."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        payload = data[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        assert n2 == n
        out.append(payload)
        pos += 8 + n
    return out


def system(natoms=30, nsf=((4, .5), (9, .5)), seed=3):
    xyz, edge = cqmm.random_structure(natoms, 0.05, 2.0, seed)
    rng = np.random.default_rng(seed)
    if isinstance(nsf, int):
        n = nsf
    else:
        n = np.array([v for v, _ in nsf])[rng.choice(len(nsf), natoms, p=[p for _, p in nsf])]
    s = cqmm.System(xyz, edge, n, 16.0, (2, 2, 2))
    return s, xyz, edge


def test_file_name_follows_get_file_name_2rank():
    assert cqmm.matrix2_file_name("L", 0, 1) == "Lmatrix2.i00.p000001"
    assert cqmm.matrix2_file_name("out/K", 7, 12) == "out/Kmatrix2.i07.p000012"


def test_binary_layout_matches_the_write_statements(tmp_path):
    s, xyz, edge = system()
    rid = s.add_range(7.0)
    v = s.fill(rid, 0, 5, 1, (2.0,))
    path = tmp_path / cqmm.matrix2_file_name("L")
    s.dump_matrix2(rid, [v], path, binary=True)
    recs = records(path)
    b = s.blocks(rid)
    n = s.natoms
    inode, nprim = struct.unpack("<2i", recs[0])
    assert (inode, nprim) == (1, n)
    nsf_i = np.frombuffer(recs[1], "<i4")
    jmax = np.frombuffer(recs[2], "<i4")
    jbeta = np.frombuffer(recs[3], "<i4")
    (nspin,) = struct.unpack("<i", recs[4])
    assert nspin == 1
    prim = b["prim"]
    order = s.primary_order()
    assert np.array_equal(jmax, np.bincount(prim, minlength=n))
    k = 5
    for p in range(n):
        (idi,) = struct.unpack("<i", recs[k])
        rows = np.nonzero(prim == p)[0]
        assert idi == order[p] + 1 and nsf_i[p] == b["nd1"][rows[0]]
        beta = np.frombuffer(recs[k + 1], "<i4")
        idj = np.frombuffer(recs[k + 2], "<i4")
        assert np.array_equal(beta, b["nd2"][rows]) and np.array_equal(idj, b["gid_j"][rows] + 1)
        assert jbeta[p] == beta.sum()
        k += 3
        for e in rows:
            vec = np.frombuffer(recs[k], "<f8")
            want = (xyz[b["gid_i"][e]] - (xyz[b["gid_j"][e]] + b["shift"][e] * edge)) / edge
            assert np.abs(vec - want).max() < 1e-12
            k += 1
        nval = jbeta[p] * nsf_i[p]
        vals = np.array([struct.unpack("<d", recs[k + q])[0] for q in range(nval)])
        o = b["offset"][rows[0]]
        assert np.array_equal(vals, v[o:o + nval])
        k += nval
    assert k == len(recs)


@pytest.mark.parametrize("binary", [True, False])
def test_round_trip_two_spins(tmp_path, binary):
    s, _, _ = system(natoms=24, nsf=((4, .5), (13, .5)), seed=5)
    rid = s.add_range(6.5)
    up, dn = s.fill(rid, 0, 7, 1, (2.0,)), s.fill(rid, 0, 7, 2, (2.0,))
    path = tmp_path / "Kmatrix2.i00.p000001"
    s.dump_matrix2(rid, [up, dn], path, binary=binary)
    (a, b), loaded, skipped = s.load_matrix2(rid, path, binary=binary)
    assert skipped == 0 and loaded == len(s.blocks(rid)["offset"])
    np.testing.assert_array_equal(a, up)
    np.testing.assert_array_equal(b, dn)
    if not binary:
        line = open(path).read().split("\n")[-2]
        assert len(line) == 30 and "E" in line and line.strip().lstrip("-").startswith("0.")


def test_load_into_another_range_keeps_the_common_blocks(tmp_path):
    s, _, _ = system()
    big, small = s.add_range(8.0), s.add_range(5.0)
    v = s.fill(big, 0, 9, 3, (2.0,))
    path = tmp_path / "Lmatrix2.i00.p000001"
    s.dump_matrix2(big, [v], path)
    (w,), loaded, skipped = s.load_matrix2(small, path)
    bb, bs = s.blocks(big), s.blocks(small)
    assert loaded == len(bs["offset"]) and skipped == len(bb["offset"]) - len(bs["offset"])
    key = lambda b, e: (b["gid_i"][e], b["gid_j"][e], *b["shift"][e])
    where = {key(bb, e): e for e in range(len(bb["offset"]))}
    for e in range(len(bs["offset"])):
        f = where[key(bs, e)]
        n = bs["nd1"][e] * bs["nd2"][e]
        np.testing.assert_array_equal(w[bs["offset"][e]:bs["offset"][e] + n], v[bb["offset"][f]:bb["offset"][f] + n])
    # And back: the small range's file fills only its blocks of the big range.
    path2 = tmp_path / "Smatrix2.i00.p000001"
    s.dump_matrix2(small, [w], path2)
    (z,), loaded2, skipped2 = s.load_matrix2(big, path2)
    assert loaded2 == len(bs["offset"]) and skipped2 == 0
    assert np.count_nonzero(z) == np.count_nonzero(w)


def test_truncated_file_is_an_error(tmp_path):
    s, _, _ = system(natoms=12, nsf=4)
    rid = s.add_range(6.0)
    path = tmp_path / "Lmatrix2.i00.p000001"
    s.dump_matrix2(rid, [s.fill(rid, 0, 1, 1, (2.0,))], path)
    data = open(path, "rb").read()
    open(path, "wb").write(data[:len(data) // 2])
    with pytest.raises(cqmm.CqmmError, match="malformed or truncated"):
        s.load_matrix2(rid, path)
