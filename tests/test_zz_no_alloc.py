"""GPU test of cqmm_mat_mult_fp32_noalloc, the FP32 product entry point that
allocates nothing (its FP32 copies live in the library's static device array).

Bar: bitwise equal to cqmm_mat_mult with the same CQMM_FP32 flags on the same
device operands (same kernels, same summation order), on two calls in a row
with the second one smaller (the array keeps the first call's contents), for
the tile kernels of uniform 13x13 (type 2), 4x4 (type 1) and a 4/9/13 species
mix (type 1); a product too large for the array is refused before any launch.
"""
import numpy as np
import pytest

import cqmm
from cases import Case

pytestmark = pytest.mark.gpu

# (natoms, nsf, ra, rb, rc, type): large first, then smaller, per kernel family.
PAIRS = [
    [(36, 13, 6.0, 8.0, 8.0, 2), (20, 13, 5.0, 6.0, 6.0, 2)],
    [(60, 4, 6.0, 8.0, 14.0, 1), (30, 4, 5.0, 6.0, 10.0, 1)],
    [(48, ((4, .5), (9, .25), (13, .25)), 6.0, 7.0, 10.0, 1), (24, ((4, .5), (9, .25), (13, .25)), 5.0, 6.0, 8.0, 1)],
]
MODES = [("fp32", cqmm.CQMM_FP32), ("fp32_acc64", cqmm.CQMM_FP32 | cqmm.CQMM_FP32_ACC64),
         ("fp32_run64", cqmm.CQMM_FP32 | cqmm.CQMM_FP32_RUN64)]


def run_pair(args, flags, seed):
    case = Case(*args, seed=seed)
    plan = cqmm.Plan(case.desc)
    a, b, c = case.values()
    out_idx = 2 if case.mult_type == 1 else 0
    ref_dev = [cqmm.DeviceArray.from_host(x) for x in (a, b, c)]
    plan.mat_mult_dev(ref_dev[0].ptr, ref_dev[0].n, ref_dev[1].ptr, ref_dev[1].n, ref_dev[2].ptr, ref_dev[2].n,
                      flags)
    ref = ref_dev[out_idx].to_host()
    got_dev = [cqmm.DeviceArray.from_host(x) for x in (a, b, c)]
    plan.mat_mult_fp32_noalloc_dev(got_dev[0].ptr, got_dev[0].n, got_dev[1].ptr, got_dev[1].n, got_dev[2].ptr,
                                   got_dev[2].n, flags & ~cqmm.CQMM_FP32)
    got = got_dev[out_idx].to_host()
    for x in ref_dev + got_dev:
        x.free()
    info = plan.info()
    plan.close()
    return ref, got, info


@pytest.mark.parametrize("mode", [m[0] for m in MODES])
@pytest.mark.parametrize("family", range(len(PAIRS)))
def test_noalloc_equals_allocating_entry_on_two_calls(family, mode):
    cqmm.init(0)
    flags = dict(MODES)[mode]
    for k, args in enumerate(PAIRS[family]):
        print(f"no-alloc {mode}: call {k + 1}, {args[0]} atoms, nsf {args[1]}, type {args[5]}", flush=True)
        ref, got, info = run_pair(args, flags, seed=5 + k)
        assert info.n_triplets > 0
        assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
        np.testing.assert_array_equal(got, ref)


def test_noalloc_refuses_a_product_beyond_its_array():
    """200 atoms x 13 SF at 8/16/16 A: the FP32 copies need more than the
    array holds; the call returns CQMM_ERR_ARG before any launch and the
    allocating entry point still runs the same product."""
    cqmm.init(0)
    limit = cqmm.lib().cqmm_noalloc_scratch_bytes() // 4
    case = Case(200, 13, 16.0, 16.0, 8.0, 2, seed=5)
    a, b, c = case.values()
    assert a.size + b.size + c.size > limit
    plan = cqmm.Plan(case.desc)
    dev = [cqmm.DeviceArray.from_host(x) for x in (a, b, c)]
    with pytest.raises(cqmm.CqmmError, match="no-allocation scratch"):
        plan.mat_mult_fp32_noalloc_dev(dev[0].ptr, dev[0].n, dev[1].ptr, dev[1].n, dev[2].ptr, dev[2].n,
                                       cqmm.CQMM_FP32_ACC64)
    plan.mat_mult_dev(dev[0].ptr, dev[0].n, dev[1].ptr, dev[1].n, dev[2].ptr, dev[2].n,
                      cqmm.CQMM_FP32 | cqmm.CQMM_FP32_ACC64)
    assert np.abs(dev[0].to_host()).max() > 0
    for x in dev:
        x.free()
    plan.close()
