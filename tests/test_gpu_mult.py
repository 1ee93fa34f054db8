"""GPU parity of the drop-in mat_mult (HIP triplet builder + FP64 MFMA kernel)
against the CPU oracle, through the C ABI.

Bar: the triplet list (which (i,k,j) products, which output block, which
operand blocks, and the per-block summation order) is bit-exact with the
default kernels' walk; values agree to 1e-12 relative (max |diff| / max |ref|)
and output blocks without products are exactly zero.
"""
import os
import sys

import numpy as np
import pytest

import cqmm
import oracle
from cases import SI8_PRODUCTS, Case, si8_case

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _stable_by_out(o, p, q):
    order = np.argsort(o, kind="stable")
    return o[order], p[order], q[order]


def check_case(case, exact_order=True):
    print(f"case natoms={case.sys.natoms} nsf={np.unique(case.nsf)} ranges={case.ra},{case.rb},{case.rc} "
          f"type={case.mult_type}", flush=True)
    plan = cqmm.Plan(case.desc)
    o_ref, p_ref, q_ref = oracle.triplets(case.desc)
    o_gpu, p_gpu, q_gpu = plan.triplets()
    if case.mult_type == 1:
        o_ref, p_ref, q_ref = _stable_by_out(o_ref, p_ref, q_ref)
    np.testing.assert_array_equal(o_gpu, o_ref)
    np.testing.assert_array_equal(p_gpu, p_ref)
    np.testing.assert_array_equal(q_gpu, q_ref)
    info = plan.info()
    assert info.n_triplets == len(o_ref)
    assert info.flops == oracle.flops(case.desc)
    a, b, c = case.values()
    a_r, b_r, c_r = a.copy(), b.copy(), c.copy()
    oracle.mat_mult(case.desc, a_r, b_r, c_r)
    # Poison the output: the GPU path must overwrite every element.
    if case.mult_type == 1:
        c[:] = np.nan
    else:
        a[:] = np.nan
    plan.mat_mult(a, b, c)
    got, ref = (c, c_r) if case.mult_type == 1 else (a, a_r)
    assert np.all(np.isfinite(got))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  triplets={len(o_ref)} rel err={err:.3e}", flush=True)
    assert err <= TOL
    # Blocks with no products are exactly zero.
    zero = ref == 0.0
    assert np.all(got[zero] == 0.0)
    return plan, got


SMALL = [
    (40, 4, 6.0, 8.0, 14.0, 1),
    (40, 4, 6.0, 8.0, 9.0, 1),
    (36, 13, 6.0, 8.0, 8.0, 2),
    (48, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1),
    (48, ((4, .5), (9, .25), (13, .25)), 7.0, 9.0, 6.0, 2),
    (30, 9, 5.0, 6.0, 11.0, 1),
    (30, 16, 5.0, 6.0, 6.0, 2),
    (40, 5, 6.0, 6.0, 12.0, 1),   # non-instanced uniform size takes the general path
    # Blocks above 16x16 (sub-block kernel): uniform 20, 4/22 mix, 64 (the limit), odd 33.
    (24, 20, 5.0, 6.0, 8.0, 1),
    (24, 20, 5.0, 6.0, 6.0, 2),
    (30, ((4, .5), (22, .5)), 5.0, 6.0, 7.0, 1),
    (30, ((4, .5), (22, .5)), 5.0, 6.0, 5.0, 2),
    (12, 64, 4.0, 4.0, 5.0, 1),
    (16, ((33, .5), (9, .5)), 4.0, 5.0, 4.0, 2),
    # One atom (every neighbour a periodic image of itself) and two atoms.
    (1, 4, 4.0, 5.0, 6.0, 1),
    (1, 13, 4.0, 5.0, 3.0, 2),
    (2, 4, 3.0, 3.0, 3.0, 1),
]


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", SMALL)
def test_small_parity(natoms, nsf, ra, rb, rc, mt):
    check_case(Case(natoms, nsf, ra, rb, rc, mt, seed=21))


@pytest.mark.parametrize("ra,rb,rc,mt", SI8_PRODUCTS)
def test_si8_reference_structure_parity(ra, rb, rc, mt):
    """The structure of the reference's Order-N test 002 (diamond Si, every range
    longer than the cell): many images of each atom per pattern."""
    check_case(si8_case(ra, rb, rc, mt))


def test_cfg1_shape_full_size_parity():
    # 512 atoms, 4x4 blocks, A 8, B 12, C 20 (full product), type 1.
    check_case(Case(512, 4, 8.0, 12.0, 20.0, 1, seed=1))


def test_cfg2_shape_reduced_parity():
    # DZP shape (13x13, A 8, B 16, C truncated to 16, m_kern_min form) on 128 atoms.
    check_case(Case(128, 13, 16.0, 16.0, 8.0, 2, seed=2))


def test_cfg3_shape_mixed_species_parity():
    check_case(Case(160, ((4, .5), (9, .25), (13, .25)), 8.0, 12.0, 16.0, 1, seed=3))


def test_deterministic_bitwise():
    case = Case(48, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1, seed=8)
    plan = cqmm.Plan(case.desc)
    a, b, c = case.values()
    plan.mat_mult(a, b, c)
    c1 = c.copy()
    for _ in range(3):
        c[:] = 0.0
        plan.mat_mult(a, b, c)
        np.testing.assert_array_equal(c, c1)


def test_device_pointer_path_matches_host_path():
    case = Case(40, 13, 6.0, 8.0, 8.0, 2, seed=5)
    plan = cqmm.Plan(case.desc)
    a, b, c = case.values()
    plan.mat_mult(a, b, c)
    da, db, dc = (cqmm.DeviceArray.from_host(x) for x in (np.zeros_like(a), b, c))
    plan.mat_mult_dev(da.ptr, da.n, db.ptr, db.n, dc.ptr, dc.n)
    np.testing.assert_array_equal(da.to_host(), a)


def test_linearity_in_each_operand():
    case = Case(64, 9, 6.0, 8.0, 12.0, 1, seed=6)
    plan = cqmm.Plan(case.desc)
    a, b, c = case.values()
    a2 = case.sys.fill(case.ida, 0, 99, 1, (2.0,))
    c1 = np.zeros_like(c); c2 = np.zeros_like(c); c3 = np.zeros_like(c)
    plan.mat_mult(a, b, c1)
    plan.mat_mult(a2, b, c2)
    plan.mat_mult(2.0 * a - 0.5 * a2, b, c3)
    ref = 2.0 * c1 - 0.5 * c2
    assert np.abs(c3 - ref).max() <= 1e-12 * np.abs(ref).max()


def test_bad_descriptor_is_rejected_before_any_kernel():
    case = Case(30, 4, 5.0, 6.0, 11.0, 1, seed=3)
    d = case.desc
    d.lenb = 10  # B blocks now point past lenb
    with pytest.raises(cqmm.CqmmError, match="lenb|B block"):
        cqmm.Plan(d)
    case2 = Case(30, 4, 5.0, 6.0, 11.0, 1, seed=3)
    d2 = case2.desc
    d2.lenc = 10  # C blocks now point past lenc
    with pytest.raises(cqmm.CqmmError, match="lenc|C block"):
        cqmm.Plan(d2)


TILE_CASES = [
    (64, 13, 8.0, 8.0, 12.0, 2),
    (64, 13, 6.0, 8.0, 12.0, 1),
    (48, 4, 6.0, 8.0, 14.0, 1),
    (40, 9, 6.0, 7.0, 10.0, 2),
    (40, 16, 5.0, 6.0, 9.0, 1),
    (40, 16, 5.0, 6.0, 6.0, 2),
    (200, 13, 12.0, 12.0, 6.0, 2),   # many tiles, multi-chunk entries
    (1, 13, 4.0, 5.0, 3.0, 2),       # one atom: one tile of self-images
    (40, 4, 6.0, 8.0, 6.0, 2),       # 4x4 blocks (dense-quadrant kernel), type 2
    (1, 4, 4.0, 5.0, 6.0, 1),        # 4x4 blocks, one atom
]


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", TILE_CASES)
def test_tile_kernel_matches_oracle_and_triplet_kernel(natoms, nsf, ra, rb, rc, mt):
    """Uniform block sizes get the output-tile schedule. Both FP64 kernels meet
    the 1e-12 bar against the oracle and overwrite every output element. Type 1
    sums every block in the same (ascending k) order in both kernels, so there
    they agree bit for bit — also for NSF = 4m+1, where the tile kernel's last
    contraction index is a VALU fused multiply-add and the triplet kernel's an
    MFMA step with three zero terms."""
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=31)
    print(f"tile case natoms={natoms} nsf={nsf} ranges={ra},{rb},{rc} type={mt}", flush=True)
    plan = cqmm.Plan(case.desc)
    info = plan.info()
    assert info.n_tiles > 0 and info.kernel == cqmm.KERNEL_TILE
    assert 0 < info.n_loads <= 2 * info.n_triplets
    a, b, c = case.values()
    a_r, b_r, c_r = a.copy(), b.copy(), c.copy()
    oracle.mat_mult(case.desc, a_r, b_r, c_r)
    ref = c_r if mt == 1 else a_r
    got = {}
    for kernel in (cqmm.KERNEL_TILE, cqmm.KERNEL_TRIPLET):
        plan.set_kernel(kernel)
        assert plan.info().kernel == kernel
        x, y, z = a.copy(), b.copy(), c.copy()
        out = z if mt == 1 else x
        out[:] = np.nan
        plan.mat_mult(x, y, z)
        assert np.all(np.isfinite(out))
        err = np.abs(out - ref).max() / np.abs(ref).max()
        print(f"  kernel {kernel}: rel err {err:.3e} (tiles {info.n_tiles}, entries {info.n_entries}, "
              f"loads/product {info.n_loads / info.n_triplets:.3f})", flush=True)
        assert err <= TOL
        assert np.all(out[ref == 0.0] == 0.0)
        got[kernel] = out.copy()
    if mt == 1:
        np.testing.assert_array_equal(got[cqmm.KERNEL_TILE], got[cqmm.KERNEL_TRIPLET])


MIXED_TILE_CASES = [
    (48, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1),
    (48, ((4, .5), (9, .25), (13, .25)), 7.0, 9.0, 6.0, 2),
    (40, ((16, .3), (4, .4), (9, .3)), 6.0, 7.0, 9.0, 1),    # 16x16 blocks: the larger stage
    (40, ((16, .3), (4, .4), (9, .3)), 6.0, 6.0, 6.0, 2),
    (40, 5, 6.0, 6.0, 12.0, 1),                              # a uniform size without a fixed-size instance
    (160, ((4, .5), (9, .25), (13, .25)), 10.0, 12.0, 8.0, 2),  # many tiles, multi-chunk entries
    # Blocks of at most 4x4 (single-zeta C/H-like mixes, uniform 1): the dense-quadrant mixed kernel.
    (60, ((4, .6), (1, .4)), 6.0, 7.0, 10.0, 1),
    (60, ((4, .6), (1, .4)), 7.0, 9.0, 6.0, 2),
    (40, ((2, .5), (3, .5)), 6.0, 7.0, 10.0, 1),
    (50, 1, 6.0, 8.0, 12.0, 1),
    (160, ((4, .6), (1, .4)), 10.0, 12.0, 8.0, 2),             # multi-chunk entries
    # Double-zeta C/H-like mixes (blocks up to 8x8): the general mixed kernel.
    (50, ((8, .6), (2, .4)), 6.0, 7.0, 10.0, 1),
    (50, ((8, .6), (2, .4)), 7.0, 9.0, 6.0, 2),
    (40, ((6, .5), (7, .5)), 6.0, 7.0, 9.0, 1),
    (120, ((8, .6), (2, .4)), 9.0, 11.0, 7.0, 2),
]


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", MIXED_TILE_CASES)
def test_mixed_size_tile_kernel_matches_oracle_and_triplet_kernel(natoms, nsf, ra, rb, rc, mt):
    """Mixed block sizes (and uniform sizes without a fixed-size instance) get
    the mixed-size tile schedule: operands packed in LDS by size, per-product
    contraction dimension. Both FP64 kernels meet the 1e-12 bar against the
    oracle, overwrite every output element, and agree with each other to 1e-13."""
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=37)
    print(f"mixed tile case natoms={natoms} nsf={np.unique(case.nsf)} ranges={ra},{rb},{rc} type={mt}", flush=True)
    plan = cqmm.Plan(case.desc)
    info = plan.info()
    assert info.n_tiles > 0 and info.kernel == cqmm.KERNEL_TILE
    assert 0 < info.n_loads <= 2 * info.n_triplets
    a, b, c = case.values()
    a_r, b_r, c_r = a.copy(), b.copy(), c.copy()
    oracle.mat_mult(case.desc, a_r, b_r, c_r)
    ref = c_r if mt == 1 else a_r
    got = {}
    for kernel in (cqmm.KERNEL_TILE, cqmm.KERNEL_TRIPLET):
        plan.set_kernel(kernel)
        x, y, z = a.copy(), b.copy(), c.copy()
        out = z if mt == 1 else x
        out[:] = np.nan
        plan.mat_mult(x, y, z)
        assert np.all(np.isfinite(out))
        err = np.abs(out - ref).max() / np.abs(ref).max()
        print(f"  kernel {kernel}: rel err {err:.3e} (tiles {info.n_tiles}, entries {info.n_entries}, "
              f"loads/product {info.n_loads / info.n_triplets:.3f})", flush=True)
        assert err <= TOL
        assert np.all(out[ref == 0.0] == 0.0)
        got[kernel] = out.copy()
    d = np.abs(got[cqmm.KERNEL_TILE] - got[cqmm.KERNEL_TRIPLET]).max() / np.abs(ref).max()
    assert d <= 1e-13


FP32_CASES = [
    (64, 13, 6.0, 8.0, 12.0, 1),
    (64, 13, 8.0, 8.0, 12.0, 2),
    (48, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1),   # mixed-size tile kernel
    (48, 4, 6.0, 8.0, 14.0, 1),                               # 4x4 blocks: dense-quadrant tile kernel
    (40, 4, 6.0, 8.0, 6.0, 2),
    (30, ((4, .5), (22, .5)), 5.0, 6.0, 7.0, 1),               # blocks above 16x16: sub-block kernel
    (24, 20, 5.0, 6.0, 6.0, 2),
    (60, ((4, .6), (1, .4)), 6.0, 7.0, 10.0, 1),               # dense-quadrant mixed kernel
    (50, ((8, .6), (2, .4)), 6.0, 7.0, 10.0, 1),               # general mixed kernel, blocks up to 8x8
]


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", FP32_CASES)
def test_fp32_products(natoms, nsf, ra, rb, rc, mt):
    """CQMM_FP32: FP32 MFMA products on FP32 copies, FP32 accumulation
    (bar 2e-5 relative to the FP64 oracle: inputs rounded to FP32, sums of a few
    hundred terms in FP32) or, with CQMM_FP32_ACC64, each block's products summed
    in FP64 (bar 2e-6: only the inputs and each 13-term product are FP32) or, with
    CQMM_FP32_RUN64, a block's products of one chunk (at most 15) summed in FP32
    and the chunks in FP64 (bar 5e-6; the per-output-block kernel sums each product
    in FP64 there). The tile
    and triplet kernels sum type-1 blocks of uniform size in the same order and agree bit
    for bit."""
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=41)
    print(f"fp32 case natoms={natoms} nsf={nsf} ranges={ra},{rb},{rc} type={mt}", flush=True)
    plan = cqmm.Plan(case.desc)
    a, b, c = case.values()
    a_r, b_r, c_r = a.copy(), b.copy(), c.copy()
    oracle.mat_mult(case.desc, a_r, b_r, c_r)
    ref = c_r if mt == 1 else a_r
    kernels = (cqmm.KERNEL_TILE, cqmm.KERNEL_TRIPLET) if plan.info().n_tiles else (cqmm.KERNEL_TRIPLET,)
    for flags, bar in ((cqmm.CQMM_FP32, 2e-5), (cqmm.CQMM_FP32 | cqmm.CQMM_FP32_ACC64, 2e-6),
                       (cqmm.CQMM_FP32 | cqmm.CQMM_FP32_RUN64, 5e-6)):
        got = {}
        for kernel in kernels:
            plan.set_kernel(kernel)
            x, y, z = a.copy(), b.copy(), c.copy()
            out = z if mt == 1 else x
            out[:] = np.nan
            plan.mat_mult(x, y, z, flags)
            assert np.all(np.isfinite(out))
            err = np.abs(out - ref).max() / np.abs(ref).max()
            print(f"  flags {flags} kernel {kernel}: rel err {err:.3e}", flush=True)
            assert err <= bar
            assert np.all(out[ref == 0.0] == 0.0)
            got[kernel] = out.copy()
        if mt == 1 and len(got) == 2 and np.ndim(nsf) == 0 and not flags & cqmm.CQMM_FP32_RUN64:
            np.testing.assert_array_equal(got[cqmm.KERNEL_TILE], got[cqmm.KERNEL_TRIPLET])


@pytest.mark.parametrize("nsf,mt", [(13, 2), (((4, .5), (9, .25), (13, .25)), 1)])
def test_plan_tune_keeps_the_faster_kernel_and_a_correct_result(nsf, mt):
    case = Case(64, nsf, 6.0, 8.0, 12.0, mt, seed=43)
    plan = cqmm.Plan(case.desc)
    a, b, c = case.values()
    a_r, b_r, c_r = a.copy(), b.copy(), c.copy()
    oracle.mat_mult(case.desc, a_r, b_r, c_r)
    ref = c_r if mt == 1 else a_r
    da, db, dc = (cqmm.DeviceArray.from_host(x) for x in (a, b, c))
    ms_tile, ms_trip = plan.tune_dev(da.ptr, da.n, db.ptr, db.n, dc.ptr, dc.n, reps=2)
    print(f"tune: tile {ms_tile} ms, per-block {ms_trip} ms -> kernel {plan.info().kernel}", flush=True)
    assert ms_tile is not None and ms_tile > 0 and ms_trip > 0
    want = cqmm.KERNEL_TILE if ms_tile <= ms_trip else cqmm.KERNEL_TRIPLET
    assert plan.info().kernel == want
    out = (dc if mt == 1 else da).to_host()
    assert np.abs(out - ref).max() <= TOL * np.abs(ref).max()
    for x in (da, db, dc):
        x.free()


def test_cfg2_full_size_properties():
    """The DZP drop-in workload at full size (4096 atoms, 13x13 blocks,
    X(16 A) = A(8 A).B(16 A)^T, 273,159,028 block products), checked through
    properties that hold at any size: the tile kernel and the per-output-block
    kernel (different schedules and summation structure) agree to 1e-13
    relative; doubling an operand doubles the result bit for bit; every output
    element is overwritten. The oracle checks the same shape on 128 atoms."""
    xyz, edge = cqmm.random_structure(4096, 0.05, 2.0, 1)
    print("cfg2 full size: building the system", flush=True)
    sysm = cqmm.System(xyz, edge, 13, 32.0, cqmm.default_grid(edge, 4.0))
    ids = {r: sysm.add_range(r) for r in (8.0, 16.0)}
    desc = sysm.mult_desc(2, ids[16.0], ids[16.0], ids[8.0])
    b = sysm.fill(ids[16.0], 0, 1, 2, (2.0,))
    c = sysm.fill(ids[8.0], 0, 1, 3, (2.0,))
    n_out = sysm.length(ids[16.0])
    plan = cqmm.Plan(desc)
    info = plan.info()
    print(f"  products {info.n_triplets}, tiles {info.n_tiles}", flush=True)
    assert info.n_triplets == 273159028 and info.n_tiles > 0
    out = {}
    for kernel in (cqmm.KERNEL_TILE, cqmm.KERNEL_TRIPLET):
        plan.set_kernel(kernel)
        a = np.full(n_out, np.nan)
        plan.mat_mult(a, b, c)
        assert np.isfinite(a).all()
        out[kernel] = a
    ref = out[cqmm.KERNEL_TRIPLET]
    err = np.abs(out[cqmm.KERNEL_TILE] - ref).max() / np.abs(ref).max()
    print(f"  tile vs per-block kernel: rel err {err:.3e}", flush=True)
    assert err <= 1e-13
    plan.set_kernel(cqmm.KERNEL_TILE)
    a2 = np.full(n_out, np.nan)
    plan.mat_mult(a2, b, 2.0 * c)
    np.testing.assert_array_equal(a2, 2.0 * out[cqmm.KERNEL_TILE])


# Species mixes for the class-tile kernel: every row/column packing (blocks up
# to 4 pack four to an MFMA tile, 5-8 two, 9-16 one), both product types,
# rectangular blocks, a 16x16 species and an isolated species (one atom).
CLASS_CASES = [
    (64, ((4, .5), (9, .25), (13, .25)), 6.0, 8.0, 12.0, 1),
    (64, ((4, .5), (9, .25), (13, .25)), 8.0, 8.0, 6.0, 2),
    (48, ((8, .5), (13, .5)), 6.0, 7.0, 10.0, 1),
    (48, ((5, .3), (16, .3), (2, .4)), 6.0, 7.0, 9.0, 1),
    (40, ((6, .5), (9, .5)), 7.0, 8.0, 6.0, 2),
    (40, ((1, .45), (4, .45), (13, .1)), 6.0, 8.0, 12.0, 1),
]


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", CLASS_CASES)
def test_class_tiles_match_oracle_general_tiles_and_persistent_grids(natoms, nsf, ra, rb, rc, mt, monkeypatch):
    """Class tiles (tile rows and columns grouped by block size, packed MFMA
    tiles, entry lists) against the oracle (1e-12 relative) and against the
    general mixed-size schedule (1e-13); persistent grids of 8 to 1024
    workgroups give the one-workgroup-per-tile result bit for bit; FP32 modes
    within their bars of the oracle."""
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=47)
    print(f"class tiles: natoms={natoms} nsf={nsf} ranges={ra},{rb},{rc} type={mt}", flush=True)
    a, b, c = case.values()
    a_r, b_r, c_r = a.copy(), b.copy(), c.copy()
    oracle.mat_mult(case.desc, a_r, b_r, c_r)
    ref = c_r if mt == 1 else a_r
    out_of = lambda x, z: z if mt == 1 else x
    monkeypatch.setenv("CQMM_MIXED_TILES", "class")
    plan = cqmm.Plan(case.desc)
    info = plan.info()
    print(f"  products {info.n_triplets}, tiles {info.n_tiles}, family {info.tile_family}", flush=True)
    assert info.n_tiles > 0 and info.tile_family == cqmm.TILES_CLASS and info.kernel == cqmm.KERNEL_TILE
    x, y, z = a.copy(), b.copy(), c.copy()
    out_of(x, z)[:] = np.nan
    plan.mat_mult(x, y, z)
    got = out_of(x, z).copy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  class tiles vs oracle: rel err {err:.3e}", flush=True)
    assert np.all(np.isfinite(got)) and err <= TOL
    assert np.all(got[ref == 0.0] == 0.0)
    for wgs in (8, 64, 1024):
        plan.set_tile_workgroups(wgs)
        x, y, z = a.copy(), b.copy(), c.copy()
        out_of(x, z)[:] = np.nan
        plan.mat_mult(x, y, z)
        np.testing.assert_array_equal(out_of(x, z), got)
    plan.set_tile_workgroups(0)
    for flags, bar in ((cqmm.CQMM_FP32, 2e-5), (cqmm.CQMM_FP32 | cqmm.CQMM_FP32_ACC64, 2e-6)):
        x, y, z = a.copy(), b.copy(), c.copy()
        out_of(x, z)[:] = np.nan
        plan.mat_mult(x, y, z, flags)
        e32 = np.abs(out_of(x, z) - ref).max() / np.abs(ref).max()
        print(f"  flags {flags}: rel err {e32:.3e}", flush=True)
        assert e32 <= bar
    plan.close()
    monkeypatch.setenv("CQMM_MIXED_TILES", "general")
    gen = cqmm.Plan(case.desc)
    assert gen.info().tile_family in (cqmm.TILES_MIXED, cqmm.TILES_DENSE_MIXED)
    x, y, z = a.copy(), b.copy(), c.copy()
    gen.mat_mult(x, y, z)
    d = np.abs(out_of(x, z) - got).max() / np.abs(ref).max()
    print(f"  class vs general mixed tiles: rel diff {d:.3e}", flush=True)
    assert d <= 1e-13
    gen.close()


def test_cfg3_full_size_class_tiles_properties():
    """The batched uneven-blocks workload at full size (8192 atoms, 50% 4 /
    25% 9 / 25% 13 SF, C(16 A) = A(8 A).B(12 A), one launch over all block
    shapes), checked through size-independent properties: class tiles
    (CQMM_MIXED_TILES=class), the general mixed-size tiles (the default) and
    the per-output-block kernel agree to 1e-13
    relative; a persistent grid gives the same result bit for bit; doubling an
    operand doubles the result bit for bit."""
    import os
    nsf = ((4, .5), (9, .25), (13, .25))
    xyz, edge = cqmm.random_structure(8192, 0.05, 2.0, 1)
    rng = np.random.default_rng(1 + 7919)
    sp = np.array([4, 9, 13], np.int32)[rng.choice(3, size=8192, p=[.5, .25, .25])]
    print("cfg3 full size: building the system", flush=True)
    sysm = cqmm.System(xyz, edge, sp, 20.0, cqmm.default_grid(edge, 4.0))
    ids = {r: sysm.add_range(r) for r in (8.0, 12.0, 16.0)}
    desc = sysm.mult_desc(1, ids[8.0], ids[12.0], ids[16.0])
    a = sysm.fill(ids[8.0], 0, 1, 1, (2.0,))
    b = sysm.fill(ids[12.0], 0, 1, 2, (2.0,))
    n_out = sysm.length(ids[16.0])
    os.environ["CQMM_MIXED_TILES"] = "class"
    try:
        plan = cqmm.Plan(desc)
    finally:
        del os.environ["CQMM_MIXED_TILES"]
    info = plan.info()
    print(f"  products {info.n_triplets}, tiles {info.n_tiles}, family {info.tile_family}", flush=True)
    assert info.tile_family == cqmm.TILES_CLASS
    out = {}
    for kernel in (cqmm.KERNEL_TILE, cqmm.KERNEL_TRIPLET):
        plan.set_kernel(kernel)
        c = np.full(n_out, np.nan)
        plan.mat_mult(a, b, c)
        assert np.isfinite(c).all()
        out[kernel] = c
    ref = out[cqmm.KERNEL_TRIPLET]
    err = np.abs(out[cqmm.KERNEL_TILE] - ref).max() / np.abs(ref).max()
    print(f"  class tiles vs per-block kernel: rel err {err:.3e}", flush=True)
    assert err <= 1e-13
    plan.set_kernel(cqmm.KERNEL_TILE)
    plan.set_tile_workgroups(768)
    c = np.full(n_out, np.nan)
    plan.mat_mult(a, b, c)
    np.testing.assert_array_equal(c, out[cqmm.KERNEL_TILE])
    c2 = np.full(n_out, np.nan)
    plan.mat_mult(2.0 * a, b, c2)
    np.testing.assert_array_equal(c2, 2.0 * out[cqmm.KERNEL_TILE])
    plan.close()
    gen = cqmm.Plan(desc)
    assert gen.info().tile_family == cqmm.TILES_MIXED
    c = np.full(n_out, np.nan)
    gen.mat_mult(a, b, c)
    err = np.abs(c - ref).max() / np.abs(ref).max()
    print(f"  general mixed tiles vs per-block kernel: rel err {err:.3e}", flush=True)
    assert err <= 1e-13
    gen.close()


@pytest.mark.parametrize("natoms,nsf,ra,rb,rc,mt", [(48, 13, 6.0, 8.0, 8.0, 2), (40, 9, 6.0, 8.0, 12.0, 1),
                                                    (48, ((4, .5), (9, .25), (13, .25)), 6.0, 8.0, 12.0, 1)])
def test_class_tiles_uniform_blocks_and_split_bf16(natoms, nsf, ra, rb, rc, mt, monkeypatch):
    """Uniform 9 and 13 blocks run as one class of the class-tile schedule
    (CQMM_UNIFORM_TILES=class): FP64 against the oracle (1e-12) and against the
    fixed-size tile kernel (1e-13). Split-BF16 x3 products (CQMM_BF16X3: each
    operand rounded to BF16 plus its BF16 remainder, hi.hi + hi.lo + lo.hi per
    block product in FP32, FP64 sums) in the class-tile kernel and in the
    fixed-size kernel of 9x9 and 13x13 blocks against the oracle: relative
    error bounded by 5e-5 (16 significant bits per operand; the dropped lo.lo
    term is 2^-16 of a product); the per-product mixed-size kernel refuses it."""
    case = Case(natoms, nsf, ra, rb, rc, mt, seed=53)
    print(f"class uniform / bf16x3: natoms={natoms} nsf={nsf} ranges={ra},{rb},{rc} type={mt}", flush=True)
    a, b, c = case.values()
    a_r, b_r, c_r = a.copy(), b.copy(), c.copy()
    oracle.mat_mult(case.desc, a_r, b_r, c_r)
    ref = c_r if mt == 1 else a_r
    out_of = lambda x, z: z if mt == 1 else x
    fixed = cqmm.Plan(case.desc)
    x, y, z = a.copy(), b.copy(), c.copy()
    fixed.mat_mult(x, y, z)
    got_fixed = out_of(x, z).copy()
    x, y, z = a.copy(), b.copy(), c.copy()
    if fixed.info().tile_family == cqmm.TILES_MIXED:
        # The per-product mixed-size kernel has no split-BF16 mode.
        with pytest.raises(cqmm.CqmmError, match="split-BF16"):
            fixed.mat_mult(x, y, z, cqmm.CQMM_FP32 | cqmm.CQMM_BF16X3)
    else:
        # Uniform 9 and 13: the fixed-size tile kernel's split-BF16 mode.
        out_of(x, z)[:] = np.nan
        fixed.mat_mult(x, y, z, cqmm.CQMM_FP32 | cqmm.CQMM_BF16X3)
        e16 = np.abs(out_of(x, z) - ref).max() / np.abs(ref).max()
        print(f"  fixed-size kernel split-BF16 x3 vs oracle: max rel {e16:.3e}", flush=True)
        assert np.all(np.isfinite(out_of(x, z))) and e16 <= 5e-5
        assert np.all(out_of(x, z)[ref == 0.0] == 0.0)
    fixed.close()
    monkeypatch.setenv("CQMM_UNIFORM_TILES", "class")
    monkeypatch.setenv("CQMM_MIXED_TILES", "class")
    plan = cqmm.Plan(case.desc)
    info = plan.info()
    assert info.tile_family == cqmm.TILES_CLASS
    x, y, z = a.copy(), b.copy(), c.copy()
    out_of(x, z)[:] = np.nan
    plan.mat_mult(x, y, z)
    got = out_of(x, z).copy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    d = np.abs(got - got_fixed).max() / np.abs(ref).max()
    print(f"  class tiles vs oracle {err:.3e}, vs default tile kernel {d:.3e}", flush=True)
    assert err <= TOL and d <= 1e-13
    x, y, z = a.copy(), b.copy(), c.copy()
    out_of(x, z)[:] = np.nan
    plan.mat_mult(x, y, z, cqmm.CQMM_FP32 | cqmm.CQMM_BF16X3)
    got16 = out_of(x, z)
    e16 = np.abs(got16 - ref).max() / np.abs(ref).max()
    f16 = np.linalg.norm(got16 - ref) / np.linalg.norm(ref)
    print(f"  split-BF16 x3 vs oracle: max rel {e16:.3e}, Frobenius rel {f16:.3e}", flush=True)
    assert np.all(np.isfinite(got16)) and e16 <= 5e-5
    assert np.all(got16[ref == 0.0] == 0.0)
    plan.close()
