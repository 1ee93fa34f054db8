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

The InfoGlobal file (dump_InfoMatGlobal, src/store_matrix_module.f90:700-781)
is checked line by line against its write statements and formats, round trip,
and read from a hand-written file in other list-directed forms; the older
dump_matrix / grab_matrix text dump (src/io_module.f90:1960-2075) by its
file name rule, its f25.18 lines and a round trip.
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


def write_records(path, recs):
    with open(path, "wb") as f:
        for r in recs:
            f.write(struct.pack("<i", len(r)) + r + struct.pack("<i", len(r)))


def _edit_ints(rec, k, value):
    a = np.frombuffer(rec, "<i4").copy()
    a[k] = value
    return a.tobytes()


@pytest.mark.parametrize("edit,match", [
    # jbeta_max of the first atom one less than its neighbours' sizes add up to
    (lambda r: r.__setitem__(3, _edit_ints(r[3], 0, np.frombuffer(r[3], "<i4")[0] - 1)), "exceed jbeta_max"),
    (lambda r: r.__setitem__(2, _edit_ints(r[2], 0, -1)), "row sizes"),             # negative jmax
    (lambda r: r.__setitem__(1, _edit_ints(r[1], 0, 100)), "row sizes"),            # nsf above 64
    (lambda r: r.__setitem__(3, _edit_ints(r[3], 0, 2**31 - 1)), "row sizes"),      # jbeta beyond any file
])
def test_corrupt_binary_counts_are_errors(tmp_path, edit, match):
    """Counts read from the file are checked before they size a buffer or a copy."""
    s, _, _ = system(natoms=12, nsf=4)
    rid = s.add_range(6.0)
    path = tmp_path / "Lmatrix2.i00.p000001"
    s.dump_matrix2(rid, [s.fill(rid, 0, 1, 1, (2.0,))], path, binary=True)
    recs = records(path)
    edit(recs)
    write_records(path, recs)
    with pytest.raises(cqmm.CqmmError, match=match):
        s.load_matrix2(rid, path, binary=True)


def test_corrupt_record_marker_and_text_integer_are_errors(tmp_path):
    s, _, _ = system(natoms=12, nsf=4)
    rid = s.add_range(6.0)
    v = s.fill(rid, 0, 1, 1, (2.0,))
    path = tmp_path / "Lmatrix2.i00.p000001"
    s.dump_matrix2(rid, [v], path, binary=True)
    data = bytearray(open(path, "rb").read())
    data[0:4] = struct.pack("<I", 0xFFFFFFF0)          # a record longer than the file
    open(path, "wb").write(bytes(data))
    with pytest.raises(cqmm.CqmmError, match="malformed or truncated"):
        s.load_matrix2(rid, path, binary=True)
    s.dump_matrix2(rid, [v], path, binary=False)
    text = open(path).read().split("\n")
    text[0] = text[0].replace("1", "x", 1)             # inode is not an integer
    open(path, "w").write("\n".join(text))
    with pytest.raises(cqmm.CqmmError, match="malformed or truncated"):
        s.load_matrix2(rid, path, binary=False)


# ---- InfoGlobal (dump_InfoMatGlobal, src/store_matrix_module.f90:700-781) ----

def test_info_global_file_name():
    assert cqmm.info_global_file_name(0) == "InfoGlobal.i00"
    assert cqmm.info_global_file_name(7) == "InfoGlobal.i07"


def test_info_global_layout_matches_the_write_statements(tmp_path):
    """Each line read back the way the reference's write statements lay it out
    (:755-774): list-directed lines, the cell as 3f25.15, and the atom records
    in (5x,i8,3x,3e22.13) columns."""
    s, xyz, edge = system(natoms=25)
    vel = np.random.default_rng(4).normal(size=(25, 3)) * 1e-3
    path = tmp_path / cqmm.info_global_file_name(3)
    s.dump_info_global(path, index=3, md_step=12, velocity=vel)
    lines = open(path).read().split("\n")
    assert lines[0].split()[:2] == ["T", "T"]
    assert [int(t) for t in lines[1].split()[:2]] == [25, 1]
    assert [int(t) for t in lines[2].split()[:3]] == [2, 2, 2]
    cell = [float(lines[3][25 * k:25 * (k + 1)]) for k in range(3)]
    assert lines[3][75:].strip() == "rcellx,y,z"
    np.testing.assert_allclose(cell, [edge] * 3, rtol=0, atol=1e-15 * edge)
    assert [int(t) for t in lines[4].split()] == [1] * 25
    assert int(lines[5].split()[0]) == 12

    def atoms(first):
        ids, vals = [], []
        for rec in lines[first:first + 25]:
            assert rec[:5] == " " * 5 and rec[13:16] == " " * 3 and len(rec) == 16 + 66
            ids.append(int(rec[5:13]))
            vals.append([float(rec[16 + 22 * c:38 + 22 * c]) for c in range(3)])
            for c in range(3):
                f = rec[16 + 22 * c:38 + 22 * c].strip()
                assert f.lstrip("-").startswith("0.") and f[-4] == "E" and len(f.split("E")[0].split(".")[1]) == 13
        return ids, np.array(vals)

    ids, pos = atoms(6)
    assert ids == list(range(1, 26))
    # The system holds positions wrapped into the cell; atom g is xyz[g - 1].
    np.testing.assert_allclose(pos, np.mod(xyz, edge), rtol=1e-12, atol=1e-12 * edge)
    assert lines[31].strip() == "velocity"
    ids, v = atoms(32)
    assert ids == list(range(1, 26))
    np.testing.assert_allclose(v, vel, rtol=1e-12, atol=0)
    assert int(lines[57].split()[0]) == 3


def test_info_global_round_trip(tmp_path):
    s, xyz, edge = system(natoms=20)
    vel = np.random.default_rng(5).normal(size=(20, 3))
    vel[0] = [1e-120, -3.5e105, 0.0]          # three-digit exponents: written without the E
    p1 = tmp_path / "a"
    s.dump_info_global(p1, index=5, md_step=3, velocity=vel)
    g = cqmm.load_info_global(p1, index=5)
    assert (g["natoms"], g["numprocs"], g["npcell"], g["md_step"], g["has_md_step"], g["index"]) == \
        (20, 1, (2, 2, 2), 3, True, 5)
    np.testing.assert_allclose(g["rcell"], [edge] * 3, rtol=1e-15)
    assert (g["glob_to_node"] == 1).all()
    np.testing.assert_allclose(g["coords"], np.mod(xyz, edge), rtol=1e-12, atol=1e-12 * edge)
    np.testing.assert_allclose(g["velocity"], vel, rtol=1e-12, atol=0)
    assert "0.1000000000000-119" in open(p1).read()
    p2 = tmp_path / "b"
    s.dump_info_global(p2, index=0)
    g = cqmm.load_info_global(p2, index=0)
    assert g["velocity"] is None and not g["has_md_step"] and g["md_step"] == 0


def test_info_global_reads_other_list_directed_forms(tmp_path):
    """A file as another compiler's list-directed writes may lay it out:
    .TRUE./.FALSE., commas, an r*c repeat count and a wrapped node list."""
    recs = "\n".join(f"     {g:8d}   " + "".join(f"{x:22.13E}" for x in (0.5 * g, 1.0, -2.25)) for g in (1, 2, 3))
    text = (".FALSE., .TRUE.  = flag_velocity, flag_MDstep\n"
            "3, 4 # of atoms, # of process\n"
            "1 1 2  npcellx,y,z\n"
            "   10.000000000000000   11.5D0   1.2E+01 rcellx,y,z\n"
            "2*1\n"
            "4\n"
            "7  MD step\n" + recs + "\n"
            " 9 = index_local in dump_InfoMatGlobal\n")
    p = tmp_path / "InfoGlobal.i09"
    p.write_text(text)
    g = cqmm.load_info_global(p, index=9)
    assert (g["natoms"], g["numprocs"], g["npcell"], g["md_step"]) == (3, 4, (1, 1, 2), 7)
    assert g["rcell"] == (10.0, 11.5, 12.0)
    assert list(g["glob_to_node"]) == [1, 1, 4]
    np.testing.assert_array_equal(g["coords"], [[0.5, 1, -2.25], [1.0, 1, -2.25], [1.5, 1, -2.25]])
    assert g["velocity"] is None and g["has_md_step"]


def test_info_global_errors(tmp_path):
    s, _, _ = system(natoms=10)
    p = tmp_path / "x"
    s.dump_info_global(p, index=2)
    with pytest.raises(cqmm.CqmmError, match="index"):
        cqmm.load_info_global(p, index=1)
    with pytest.raises(cqmm.CqmmError, match="room"):
        cqmm.load_info_global(p, index=2, natoms_max=5)
    lines = open(p).readlines()
    p.write_text("".join(lines[:9]))
    with pytest.raises(cqmm.CqmmError, match="truncated"):
        cqmm.load_info_global(p, index=2)
    # An atom count no file of this length can hold fails before it sizes anything.
    assert lines[1].split()[0] == "10"
    p.write_text("".join([lines[0], lines[1].replace("10", "2000000000", 1)] + lines[2:]))
    with pytest.raises(cqmm.CqmmError, match="ni_in_cell"):
        cqmm.load_info_global(p, index=2)


# ---- dump_matrix / grab_matrix (src/io_module.f90:1960-2075) ----

@pytest.mark.parametrize("numprocs,inode", [(1, 1), (8, 12), (1000, 7), (1, 12345), (123456, 99)])
def test_matrix_file_name_follows_get_file_name(numprocs, inode):
    # src/io_module.f90:3093-3104: max(3, floor(1+log10(numprocs))) - floor(1+log10(inode)) zeros.
    import math
    pad = max(3, math.floor(1 + math.log10(numprocs))) - math.floor(1 + math.log10(inode))
    assert cqmm.matrix_file_name("L", numprocs, inode) == "Lmatrix." + "0" * max(pad, 0) + str(inode)


def test_dump_matrix_text_round_trip(tmp_path):
    s, _, _ = system()
    rid = s.add_range(6.0)
    v = s.fill(rid, 0, 9, 2, (2.0,))
    v[:3] = [0.0, -1.0, 1e-30]
    p = tmp_path / cqmm.matrix_file_name("K")
    cqmm.dump_matrix_text(p, v)
    lines = open(p).read().split("\n")
    assert lines[-1] == "" and len(lines) == v.size + 1
    assert all(len(l) == 25 and l[-19] == "." for l in lines[:-1])
    back = cqmm.grab_matrix_text(p, v.size)
    # f25.18 keeps 18 decimals: each value back within half a unit of the 18th,
    # plus the rounding of that decimal to the nearest double.
    assert (np.abs(back - v) <= 5e-19 + np.spacing(np.abs(v))).all()
    assert np.max(np.abs(back - v)) < 1e-18
    assert back[2] == 0.0
    with pytest.raises(cqmm.CqmmError, match="missing"):
        cqmm.grab_matrix_text(p, v.size + 1)


def test_dump_matrix_text_too_wide_is_asterisks(tmp_path):
    p = tmp_path / "w"
    cqmm.dump_matrix_text(p, np.array([1.5, -12345678.0]))
    assert open(p).read().split("\n")[1] == "*" * 25
    with pytest.raises(cqmm.CqmmError, match="unreadable"):
        cqmm.grab_matrix_text(p, 2)
