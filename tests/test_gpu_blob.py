"""The engine's bias-state blob read field by field on the host
(gluedmd_amd.bias_state, the 'GPUS' v1 layout of CommonSerialization.cpp
:20-178 with this library's hill-list section), its values checked against
the engine's own accessors, packed back byte for byte, and a BiasStateMerger
merge of two walkers' blobs restored into a fresh context that then writes
the same bytes."""
import numpy as np
import pytest
import torch

from gluedmd_amd import bias_state as bs
from gluedmd_amd.glued import Force

pytestmark = pytest.mark.gpu


def _force():
    f = Force(pbc=True, temperature=300.0)
    a = f.add_distance([0, 1])
    b = f.add_distance([0, 2])
    f.add_opes([a, b], [0.05, 0.04], gamma=10.0, pace=1, max_kernels=200)
    f.add_metad([a, b], [0.1, 0.1], 0.5, 1, grid_min=[0.5, 0.5], grid_max=[2.5, 2.5], bins=[20, 16],
                periodic=[False, True], gamma=8.0)
    f.add_pbmetad([a, b], 0.1, 0.4, 1, grid_min=0.5, grid_max=2.5, bins=[12, 10], periodic=[True, False],
                  gamma=6.0)
    f.add_upper_wall(a, 2.0, 100.0)
    f.add_linear([b], [0.3])
    f.add_opes_expanded([a, b], weights=[1.0, 2.0], pace=1)
    f.add_multithermal(300.0, 400.0, 4, temp0=300.0, pace=1)
    f.add_metad_hills([a, b], [0.1, 0.1], 0.6, 1, gamma=10.0, max_hills=64)
    return f


def test_engine_blob_fields_repack_and_merge():
    from gluedmd_amd.context import Context
    f = _force()
    W, N = 2, 3
    g = Context(f, walkers=[N] * W)
    g.set_box(4.0)
    rng = np.random.default_rng(8)
    fe = [torch.zeros(3 * g.padded[w], dtype=torch.float32, device=g.device) for w in range(W)]
    for step in range(6):
        pos = [np.array([(0, 0, 0), (1.0 + 0.1 * rng.normal(), 0, 0), (0, 1.2 + 0.1 * rng.normal(), 0)],
                        dtype=np.float32) for _ in range(W)]
        for w in range(W):
            g.set_energy_input(-100.0 + rng.normal(), fe[w], w)
        g.update_state(step)
        g.set_all_positions(pos)
        g.zero_forces()
        g.execute_batch(use_graph=False)
    torch.cuda.synchronize()
    blobs = []
    for w in range(W):
        blob = bytes(g.getBiasState(w))
        st = bs.parse(blob, f)
        assert bs.pack(st) == blob
        assert st.header and st.n_sections == len(bs.SECTIONS)
        o = st.opes[0]
        assert o.nk == int(g.getOPESMetrics(0, w)[2]) and o.nk > 0
        # getKernelSigmas hands the widths out as float (CommonGluedKernels.cpp:5200-5232)
        np.testing.assert_array_equal(o.sigmas.reshape(-1).astype(np.float32), g.getKernelSigmas(0, w))
        assert st.metad[0][0] == 5 and st.metad[0][1].shape == (21 * 16,) and np.any(st.metad[0][1] != 0)
        assert [len(x[1]) for x in st.pbmetad[0]] == [12, 11] and all(x[0] == 5 for x in st.pbmetad[0])
        assert (st.n_linear, st.n_wall) == (1, 1)
        assert st.opes_expanded[0][1] == 5
        dF, rct, cnt = st.multithermal[0]
        assert len(dF) == 4 and abs(dF[0]) < 1e-12 and cnt > 0
        hc, hh = g.get_hills(7, w)
        assert np.array_equal(st.hills[0][0], np.asarray(hc, dtype=np.float64).reshape(-1, 2))
        assert np.array_equal(st.hills[0][1], np.asarray(hh, dtype=np.float64))
        blobs.append(blob)
    merged = bs.BiasStateMerger.merge_additive(blobs, f, opes_sum_uprob="recompute")
    m = bs.parse(merged, f)
    s0, s1 = bs.parse(blobs[0], f), bs.parse(blobs[1], f)
    np.testing.assert_array_equal(m.metad[0][1], s0.metad[0][1] + s1.metad[0][1])
    assert m.opes[0].nk == s0.opes[0].nk + s1.opes[0].nk
    assert len(m.hills[0][1]) == len(s0.hills[0][1]) + len(s1.hills[0][1])
    g2 = Context(f, walkers=[N] * W)
    g2.set_box(4.0)
    for w in range(W):
        g2.setBiasState(merged, w)
    for w in range(W):
        assert bytes(g2.getBiasState(w)) == merged
