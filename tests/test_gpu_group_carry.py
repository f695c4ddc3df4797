"""Wide group CVs (over 1,024 atoms or pairs) carried by another launch of the
batched step: by the 8-phase path-row launch (512-thread items) in a plan with
a path CV and no coordination, by the coordination rank launch (1,024-thread
items) in a plan with coordination.  Either way the plain per-walker launches
run the same group code on a launch of their own, so the two forms give the
same bits (CV values and fixed-point forces), and both agree with the FP64
oracle (CommonGluedKernels.cpp:556-639 gyration, :1130-1224 DRMSD).  The plan
also holds a contact map of three 512-pair chunks: in the batched step the
path-row launch adds its chunk partials, in the plain form the last chunk
does (CommonGluedKernels.cpp:1236-1323)."""
import numpy as np
import pytest

from gluedmd_amd.glued import Force
from tests.gpu_util import check

pytestmark = pytest.mark.gpu


def _system(rng, n, M, N, L):
    ref = rng.uniform(0.5, L - 0.5, size=(M, 3))
    frames = [ref.copy()]
    for _ in range(N - 1):
        frames.append(frames[-1] + rng.normal(0, 0.03, size=(M, 3)))
    return ref, np.stack(frames).astype(np.float32)


def _force(rng, n, M, N, L, coordination):
    f = Force(pbc=True, temperature=300.0)
    ref, frames = _system(rng, n, M, N, L)
    s = f.add_path(list(range(M)), [fr.reshape(-1).tolist() for fr in frames], 40.0)
    pairs = rng.choice(n, size=(1100, 2), replace=True)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    f.add_drmsd(pairs.reshape(-1).tolist(), rng.uniform(0.3, 1.5, size=len(pairs)).tolist())
    f.add_gyration(list(range(M, M + 1500)))
    cm = rng.choice(n, size=(1300, 2), replace=True)  # three chunks of up to 512 pairs
    cm = cm[cm[:, 0] != cm[:, 1]]
    f.add_contact_map(cm.reshape(-1).tolist(), r0=rng.uniform(0.4, 1.2, size=len(cm)).tolist())
    if coordination:
        f.add_coordination(list(range(0, 60)), list(range(200, 1400)), 0.3, cutoff=0.9)
    nv = f.getNumCollectiveVariables()
    f.setTestBiasGradients(rng.normal(0, 1.0, size=nv).tolist())
    return f, ref


@pytest.mark.parametrize("coordination", [False, True])
def test_group_cvs_carried_match_plain_and_oracle(coordination, monkeypatch):
    from gluedmd_amd.context import Context
    monkeypatch.setenv("GLUED_AUTOTUNE", "0")  # the default 8 path phases: the carrying launch
    from oracle.oracle import OracleContext
    rng = np.random.default_rng(71 + coordination)
    n, M, N, L, W = 2000, 80, 16, 3.2, 3
    f, ref = _force(rng, n, M, N, L, coordination)
    g = Context(f, walkers=[n] * W)
    g.set_box(L)
    pos = []
    for w in range(W):
        p = rng.uniform(0.0, L, size=(n, 3))
        p[:M] = ref + rng.normal(0, 0.05, size=(M, 3))
        pos.append(p.astype(np.float32))
    g.set_all_positions(pos)
    for graph in (False, True):
        g.zero_forces()
        g.execute_batch(use_graph=graph)
        Fb = [g.get_forces_fixed(w).copy() for w in range(W)]
        cvb = [g.getLastCVValues(w).copy() for w in range(W)]
        g.zero_forces()
        g.execute_all_plain()
        for w in range(W):
            assert np.array_equal(g.getLastCVValues(w), cvb[w]), (graph, w)
            assert np.array_equal(g.get_forces_fixed(w), Fb[w]), (graph, w)
    for w in range(W):
        o = OracleContext(f, n)
        o.set_box(L)
        r = o.execute(pos[w])
        check(g, o, r, g.getLastBias(w), w)
