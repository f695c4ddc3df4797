"""One bias table shared by separate handles (the reference's getMultiWalkerPtrs /
setMultiWalkerPtrs, GluedForce.cpp:257-270, CommonGluedKernels.cpp:5234-5262;
its tests/test_multiwalker.py drives two contexts over one MetaD grid).

Two one-walker contexts, the second pointed at the first's OPES table, hill
list and MetaD grid, stepped in turn on one stream, against one two-walker
context whose walkers share the same biases: the OPES and hill tables and the
forces they give agree bit for bit (each deposit is one walker's, in walker
order, in both forms); the grid agrees to rounding (the two-walker deposit adds
both walkers' Gaussians to a node before adding them to the grid)."""
import numpy as np
import pytest
import torch

from gluedmd_amd._lib import GluedError
from gluedmd_amd.glued import Force

pytestmark = pytest.mark.gpu

N = 3


def _force(opes_cap=400):
    f = Force(pbc=True, temperature=300.0)
    a = f.add_distance([0, 1])
    b = f.add_distance([0, 2])
    f.add_opes([a, b], [0.05, 0.04], gamma=10.0, pace=1, max_kernels=opes_cap)
    f.add_metad_hills([a, b], [0.1, 0.1], 0.6, 1, gamma=10.0, max_hills=64)
    f.add_metad([a], [0.1], 0.5, 2, grid_min=[0.5], grid_max=[2.5], bins=[40], gamma=8.0)
    for k in range(3):
        f.share_bias(k)
    return f


def _positions(rng):
    return np.array([(0, 0, 0), (1.0 + 0.15 * rng.normal(), 0, 0), (0, 1.2 + 0.15 * rng.normal(), 0)],
                    dtype=np.float32)


def test_two_handles_share_tables_like_one_batched_handle():
    from gluedmd_amd.context import Context
    f = _force()
    one = Context(f, walkers=[N, N])
    hA = Context(f, walkers=[N])
    hB = Context(f, walkers=[N])
    for g in (one, hA, hB):
        g.set_box(4.0)
    for b in range(3):  # the reference's calling form: force.getMultiWalkerPtrs(context, bias)
        ptrs = f.getMultiWalkerPtrs(hA, b)
        assert len(ptrs) == (5, 3, 1)[b] and all(p != 0 for p in ptrs)
        f.setMultiWalkerPtrs(hB, b, ptrs, primary=hA)
    assert hB.getMultiWalkerPtrs(0) == hA.getMultiWalkerPtrs(0)
    rng = np.random.default_rng(21)
    for step in range(14):
        pos = [_positions(rng), _positions(rng)]
        print(f"step {step}", flush=True)
        one.update_state(step)
        one.set_all_positions(pos)
        one.zero_forces()
        one.execute_batch(use_graph=False)
        hA.update_state(step)
        hB.update_state(step)
        for g, p in ((hA, pos[0]), (hB, pos[1])):
            g.set_all_positions([p])
            g.zero_forces()
            g.execute(0, True, False)
        torch.cuda.synchronize()
        for w, g in ((0, hA), (1, hB)):
            assert np.array_equal(g.getLastCVValues(0), one.getLastCVValues(w)), (step, w)
            fo, fg = one.get_forces_fixed(w).astype(np.int64), g.get_forces_fixed(0).astype(np.int64)
            assert np.max(np.abs(fo - fg)) <= 64, (step, w, fo, fg)  # 64 units of 2^-32 kJ/mol/nm
    for g in (hA, hB):
        mo, mg = one.getOPESMetrics(0, 0), g.getOPESMetrics(0, 0)
        assert mg == mo and mo[2] > 1, (mo, mg)
        np.testing.assert_array_equal(g.getKernelSigmas(0, 0), one.getKernelSigmas(0, 0))
        co, ho = one.get_hills(1, 0)
        cg, hg = g.get_hills(1, 0)
        assert len(ho) == 2 * 13
        np.testing.assert_array_equal(cg, co)
        np.testing.assert_array_equal(hg, ho)
    # the grid: both handles see one table, equal to the two-walker one to rounding
    sa, so = hA.getBiasState(0), one.getBiasState(0)
    assert len(sa) == len(so)
    from gluedmd_amd import bias_state as bs
    ga, go = bs.parse(sa, f).metad[0][1], bs.parse(so, f).metad[0][1]
    assert np.any(go != 0)
    np.testing.assert_allclose(ga, go, rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(bs.parse(hB.getBiasState(0), f).metad[0][1], ga)


def test_redirect_rejects_mismatched_tables():
    from gluedmd_amd.context import Context
    big = Context(_force(opes_cap=400_000), walkers=[N])
    small = Context(_force(opes_cap=50_000), walkers=[N])
    ptrs = small.getMultiWalkerPtrs(0)
    with pytest.raises(GluedError, match="smaller than this bias"):
        big.setMultiWalkerPtrs(0, ptrs)
    with pytest.raises(GluedError, match="expected 5 pointers"):
        big.setMultiWalkerPtrs(0, ptrs[:3])
    with pytest.raises(GluedError, match="not device memory"):
        big.setMultiWalkerPtrs(0, [0] * 5)
    # one table per walker: nothing to hand out
    f = Force(pbc=True, temperature=300.0)
    a = f.add_distance([0, 1])
    f.add_opes([a], [0.05], gamma=10.0, pace=1, max_kernels=100)
    two = Context(f, walkers=[N, N])
    with pytest.raises(GluedError, match="one table per walker"):
        two.getMultiWalkerPtrs(0)
    # the rejected calls left the table in use: the context still evaluates
    big.set_box(4.0)
    big.set_all_positions([_positions(np.random.default_rng(3))])
    big.zero_forces()
    big.execute(0, True, False)
    torch.cuda.synchronize()
