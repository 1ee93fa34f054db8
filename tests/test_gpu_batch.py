"""GPU tests of batched McWeeny purification (cqmm_mcweeny_batch).

Several independent systems step in lockstep: their products run as grouped
tile launches, their traces stay on the device and unchanged steps replay a
captured hipGraph. Each system's result must be the one it gets alone, bit for
bit (same kernels, same summation order), whatever it is batched with, with or
without graphs and grouping; every graph replay is checked against the eager
path through the final L and the per-step energy trace.
"""
import numpy as np
import pytest

import cqmm
from cqmm.dm import BATCH_NO_GRAPH, BATCH_NO_GROUPING, Purifier, mcweeny_batch

pytestmark = pytest.mark.gpu


def systems():
    """Uniform 13 SF (two structures), uniform 4 SF and a 4/9 species mix, at
    ranges that keep the oracle-free check fast but give every system many
    periodic images and tile schedules of all three kernel families."""
    out = []
    for natoms, nsf, seed in ((20, 13, 3), (16, 13, 4), (24, 4, 5), (20, None, 6)):
        xyz, edge = cqmm.random_structure(natoms, 0.05, 2.0, seed)
        if nsf is None:
            nsf = np.where(np.random.default_rng(seed).random(natoms) < 0.5, 4, 9).astype(np.int32)
        pu = Purifier(xyz, edge, nsf, r_S=4.5, r_H=4.5, r_L=6.0, grid=(2, 2, 2), seed=seed)
        pu.hotelling()
        pu.init_mcweeny()
        pu.L0 = pu.ctx.get(pu.matL)
        out.append(pu)
    return out


def reset(pus):
    for pu in pus:
        pu.ctx.set(pu.matL, pu.L0)


@pytest.fixture(scope="module")
def pus():
    return systems()


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_results_equal_alone_bitwise(pus, precision):
    alone = []
    for pu in pus:
        reset([pu])
        r = pu.mcweeny(precision=precision, max_iter=60, fp32_tol=0.0)
        alone.append((r, pu.ctx.get(pu.matL)))
    print("alone:", [(r["iterations"], r["fp32_iterations"], r["stop_reason"]) for r, _ in alone], flush=True)
    for flags in (0, BATCH_NO_GRAPH, BATCH_NO_GROUPING):
        reset(pus)
        out = mcweeny_batch(pus, precision=precision, max_iter=60, fp32_tol=0.0, flags=flags)
        for pu, (st, r), (ra, La) in zip(pus, out, alone):
            assert st == 0
            assert r["iterations"] == ra["iterations"] and r["fp32_iterations"] == ra["fp32_iterations"]
            assert r["stop_reason"] == ra["stop_reason"]
            assert r["energy"] == ra["energy"] and r["electrons"] == ra["electrons"]
            assert r["energy_trace"] == ra["energy_trace"]
            np.testing.assert_array_equal(pu.ctx.get(pu.matL), La)
    # Subsets and another order: each system still gets its own result.
    reset(pus)
    sub = [pus[2], pus[0]]
    out = mcweeny_batch(sub, precision=precision, max_iter=60, fp32_tol=0.0)
    for pu, (st, r) in zip(sub, out):
        k = pus.index(pu)
        assert st == 0 and r["energy_trace"] == alone[k][0]["energy_trace"]
        np.testing.assert_array_equal(pu.ctx.get(pu.matL), alone[k][1])


def test_batch_failure_of_one_system_leaves_the_others(pus):
    """A system whose step fails (L = 0: c2 vanishes) reports its status; the
    others finish with their own results and every stack is left as found."""
    reset(pus)
    ref = []
    for pu in pus[:2]:
        reset([pu])
        ref.append(pu.mcweeny(precision=0, max_iter=60))
    reset(pus)
    pus[1].ctx.scale(0.0, pus[1].matL)
    out = mcweeny_batch(pus[:3], precision=0, max_iter=60)
    assert out[0][0] == 0 and out[0][1]["energy_trace"] == ref[0]["energy_trace"]
    assert out[1][0] == -6  # CQMM_ERR_NOTCONV
    assert out[2][0] == 0
    for pu in pus[:3]:
        h = pu.ctx.allocate(pu.rid["L"])
        assert h == pu.matL + 1
        pu.ctx.free(h)


def test_batch_rejects_a_shared_context(pus):
    pu = pus[0]
    reset([pu])
    out = mcweeny_batch([pu, pu], precision=0, max_iter=5)
    assert out[0][0] == 0
    assert out[1][0] == -1  # CQMM_ERR_ARG: two systems of one batch share a context
