"""The Fortran host (iso_c_binding module cqmm_f + cq_gpu_mat_mult, built with
flang) drives the drop-in multiply the way CONQUEST's matrix_product would; its
results must equal the oracle's for the same seeded system. The density-matrix
driver must match the Python host, and the file driver (host-only calls) must
write the same files as the Python host."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from cases import Case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "conquest-release_amd", "fortran", "build", "cq_mult_driver")
DM_DRIVER = os.path.join(ROOT, "conquest-release_amd", "fortran", "build", "cq_dm_driver")
CQ_DRIVER = os.path.join(ROOT, "conquest-release_amd", "fortran", "build", "cq_conquest_driver")


def write_conquest_arrays(path, desc, a, b, c):
    """The arrays of a one-process type(matrix_mult) in cq_conquest_driver's
    stream layout (what a CONQUEST run holds in its halos, local transpose,
    bmat(:) and bindex), with the operand values."""
    import ctypes as C

    def arr(ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(int(n),)).astype(np.int32) if n else np.zeros(0, np.int32)

    n_prim = desc.n_prim
    out = [np.array([desc.mult_type, n_prim, desc.bmat.n_parts, desc.ncovery, desc.ncoverz], np.int32)]
    for h in (desc.ahalo, desc.chalo):
        out.append(np.array([h.np_in_halo, h.ni_in_halo], np.int32))
        out.append(np.array([h.len_i_halo, h.len_i_hbeg], np.int64))
        for name, n in (("nh_part", h.np_in_halo), ("j_beg", h.np_in_halo), ("lab_hcell", h.np_in_halo),
                        ("lab_hcover", h.np_in_halo), ("j_seq", h.ni_in_halo), ("ndimj", h.ni_in_halo),
                        ("i_h2d", n_prim * h.ni_in_halo), ("i_halo", h.len_i_halo), ("i_hbeg", h.len_i_hbeg),
                        ("ndimi", n_prim)):
            out.append(arr(getattr(h, name), n))
    t, nh = desc.ltrans, desc.ahalo.ni_in_halo
    out.append(np.array([t.len_i_prim], np.int64))
    for name, n in (("n_hnab", nh), ("i_beg", nh), ("i_nd_beg", nh), ("i_prim", t.len_i_prim)):
        out.append(arr(getattr(t, name), n))
    bm = desc.bmat
    for name in ("array_posn", "n_atoms", "part_nabs", "nd_offset"):
        out.append(arr(getattr(bm, name), bm.n_parts))
    out.append(np.array([bm.len_bindex], np.int64))
    out.append(arr(bm.bindex, bm.len_bindex))
    out.append(np.array([desc.n_cell_parts], np.int32))
    out.append(arr(desc.i_cc2seq, desc.n_cell_parts))
    out.append(np.array([a.size, b.size, c.size], np.int64))
    out += [a.astype(np.float64), b.astype(np.float64), c.astype(np.float64)]
    with open(path, "wb") as f:
        for x in out:
            f.write(np.ascontiguousarray(x).tobytes())


def test_fortran_host_is_built():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() (flang builds the Fortran host)"
    assert os.path.exists(DM_DRIVER)
    assert os.path.exists(os.path.join(ROOT, "conquest-release_amd", "fortran", "build", "cq_io_driver"))
    assert os.path.exists(CQ_DRIVER)
    src = open(os.path.join(ROOT, "conquest-release_amd", "fortran", "cqmm_f.f90")).read()
    for name in ("cqmm_mat_mult", "cqmm_ctx_add_mult", "cqmm_mcweeny", "cqmm_lnv", "cqmm_dmm",
                 "cqmm_matrix_transpose", "cqmm_plan_tune", "cqmm_sys_dump_matrix2", "cqmm_load_info_global"):
        assert f"bind(c, name=\"{name}\")" in src


@pytest.mark.gpu
@pytest.mark.parametrize("args", [(40, 4, 6.0, 8.0, 14.0, 1, 3), (36, 13, 8.0, 8.0, 6.0, 2, 4)])
def test_fortran_host_matches_oracle(tmp_path, args):
    natoms, nsf, ra, rb, rc, mt, seed = args
    out = tmp_path / "out.bin"
    print("fortran driver", args, flush=True)
    r = subprocess.run([DRIVER] + [str(x) for x in args] + [str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.uint8)
    ntrip = int(raw[:8].view(np.int64)[0])
    got = raw[8:].view(np.float64)
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=seed)
    a, b, c = case.values()
    oracle.mat_mult(case.desc, a, b, c)
    ref = c if mt == 1 else a
    assert ntrip == len(oracle.triplets(case.desc)[0])
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("args", [(40, 4, 6.0, 8.0, 14.0, 1, 3), (36, 13, 8.0, 8.0, 6.0, 2, 4),
                                  (40, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1, 5)])
def test_fortran_conquest_types_drive_the_drop_in(tmp_path, args):
    """The Fortran host fills its matrix_halo / matrix_trans / matrix(:) /
    matrix_mult types from CONQUEST-layout arrays (not through the library's
    index mirror), flattens them with halo_desc / trans_desc / bmat_desc and
    calls the mat_mult wrapper keyed by mult id: results match the oracle, and
    the second call (B doubled) reuses the cached plan and doubles the result
    exactly."""
    natoms, nsf, ra, rb, rc, mt, seed = args
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=seed)
    a, b, c = case.values()
    src, out = tmp_path / "abc.bin", tmp_path / "out.bin"
    write_conquest_arrays(src, case.desc, a, b, c)
    print("fortran conquest-types driver", args, flush=True)
    r = subprocess.run([CQ_DRIVER, str(src), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "plan built once, reused" in r.stdout
    got = np.fromfile(out, dtype=np.float64)
    oracle.mat_mult(case.desc, a, b, c)
    ref = c if mt == 1 else a
    assert got.size == 2 * ref.size
    first, second = got[:ref.size], got[ref.size:]
    assert np.abs(first - ref).max() <= 1e-12 * np.abs(ref).max()
    np.testing.assert_array_equal(second, 2.0 * first)


@pytest.mark.gpu
def test_fortran_density_matrix_host_matches_python_host():
    """The Fortran host drives Hotelling -> InitMcW -> McWMin -> FindMinDM through
    the handle API; the same sequence from Python (cqmm.dm.Purifier, the same
    system, ranges, mult table and parameters) gives the same iterations and
    energies (same library calls; host arithmetic of H + 3S done without FMA in both)."""
    import cqmm
    from cqmm.dm import Purifier
    natoms, nsf, seed = 24, 4, 5
    out = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"cq_dm_driver_{os.getpid()}.txt")
    print("fortran dm driver", natoms, nsf, seed, flush=True)
    r = subprocess.run([DM_DRIVER, str(natoms), str(nsf), str(seed), out], capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stderr
    lines = open(out).read().split("\n")
    os.remove(out)
    hot_it, mcw_it, mcw_e, mcw_n = lines[0].split()
    done, early, late, e_band, n_el, resid = lines[1].split()
    xyz, edge = cqmm.random_structure(natoms, 0.05, 2.0, seed)
    pu = Purifier(xyz, edge, nsf, r_S=4.5, r_H=4.5, r_L=6.0, grid=(2, 2, 2), seed=seed, lnv=True)
    ne = float(pu.n_orb // 2)
    it, _ = pu.hotelling()
    pu.init_mcweeny(ne)
    mcw = pu.mcweeny(precision=0, max_iter=100)
    res = pu.dmm(n_electrons=ne, tolerance=1e-7, n_L_iterations=100)
    assert int(hot_it) == it and int(mcw_it) == mcw["iterations"]
    assert abs(float(mcw_e) - mcw["energy"]) <= 1e-12 * abs(mcw["energy"])
    assert int(done) == res["done"] == 1
    assert (int(early), int(late)) == (res["early_iterations"], res["late_iterations"])
    assert abs(float(e_band) - res["band_energy"]) <= 1e-12 * abs(res["band_energy"])
    assert abs(float(n_el) - res["n_electrons"]) <= 1e-12 * abs(res["n_electrons"])


IO_DRIVER = os.path.join(ROOT, "conquest-release_amd", "fortran", "build", "cq_io_driver")


def test_fortran_file_io_host(tmp_path):
    """The Fortran host writes and reads the matrix files through cqmm_f
    (dump_matrix2 binary and text, InfoGlobal, dump_matrix); the files equal
    the ones the Python host writes for the same system and values, byte for
    byte, and read back exactly (text dumps to their digits)."""
    import cqmm
    natoms, nsf, seed = 30, 4, 3
    out = tmp_path / "out.txt"
    r = subprocess.run([IO_DRIVER, str(natoms), str(nsf), str(seed), str(tmp_path), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = open(out).read().split("\n")
    xyz, edge = cqmm.random_structure(natoms, 0.05, 2.0, seed)
    s = cqmm.System(xyz, edge, nsf, 6.0, (2, 2, 2))
    rid = s.add_range(6.0)
    vals = s.fill(rid, 1, seed, 11, (0.1, 1.5))
    nblocks = s.range_info(rid).n_blocks
    assert lines[0] == cqmm.matrix2_file_name("L", 0, 1)
    for ln in lines[1:3]:
        ns, nl, nk, diff = ln.split()
        assert (int(ns), int(nl), int(nk), float(diff)) == (1, nblocks, 0, 0.0)
    s.dump_matrix2(rid, [vals], tmp_path / "py_bin", binary=True)
    s.dump_matrix2(rid, [vals], tmp_path / "py_txt", binary=False)
    assert open(tmp_path / lines[0], "rb").read() == open(tmp_path / "py_bin", "rb").read()
    assert open(tmp_path / ("text_" + lines[0])).read() == open(tmp_path / "py_txt").read()
    assert lines[3] == cqmm.info_global_file_name(4)
    assert [int(t) for t in lines[4].split()] == [natoms, 1, 2, 2, 2, 17, 1, 1, 4]
    vdiff, rcell, nodes = lines[5].split()
    assert float(vdiff) < 1e-15 and float(rcell) == edge and int(nodes) == natoms
    coords = np.array([float(t) for t in lines[6:6 + 3 * natoms]]).reshape(natoms, 3)
    g = cqmm.load_info_global(tmp_path / lines[3], index=4)
    np.testing.assert_array_equal(coords, g["coords"])
    assert lines[6 + 3 * natoms] == cqmm.matrix_file_name("K", 1, 1)
    assert float(lines[7 + 3 * natoms]) < 1e-18
    back = cqmm.grab_matrix_text(tmp_path / lines[6 + 3 * natoms], vals.size)
    assert np.abs(back - vals).max() < 1e-18
