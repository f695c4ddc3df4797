"""The launch-shape autotuner: at an engine's first (batched) evaluation it
times candidate shapes of the path and coordination kernels and keeps the
fastest; results stay within the oracle's tolerances, the one-walker launches
of the same engine give the batched bits, and a second engine of the same
plan shape takes the cached choice and gives the same bits."""
import numpy as np
import pytest

from gluedmd_amd import workloads as wl
from gluedmd_amd.glued import Force
from tests.gpu_util import check

pytestmark = pytest.mark.gpu


def _path_coord_force(c, nA, nB):
    f = Force(pbc=True, temperature=300.0)
    M = c["M"]
    s = f.add_path(list(range(M)), [fr.reshape(-1).tolist() for fr in c["frames"]], c["lam"])
    f.add_coordination(list(range(M, M + nA)), list(range(M + nA, M + nA + nB)), 0.3, cutoff=0.9)
    f.setTestBiasGradients([0.8, -0.3, 0.01])
    return f, s


def _positions(c, W, nA, nB, L, rng):
    out = []
    for b in range(W):
        extra = rng.uniform(0, L, size=(nA + nB, 3))
        out.append(np.concatenate([c["confs"][b], extra]).astype(np.float32))
    return out


def test_autotune_measures_caches_and_keeps_parity():
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    W, nA, nB = 20, 150, 1500
    c = wl.config3_system(batch=W, M=120, N=40)
    f, _ = _path_coord_force(c, nA, nB)
    n = c["M"] + nA + nB
    L = 3.2  # 3 cells of >= 0.9 nm per axis
    rng = np.random.default_rng(47)
    pos = _positions(c, W, nA, nB, L, rng)
    g = Context(f, walkers=[n] * W)
    g.set_box(L)
    g.set_all_positions(pos)
    g.zero_forces()
    g.execute_batch(use_graph=False)
    rep = g.tuning()
    assert "(measured)" in rep and "coordBlock:" in rep and "pathBlocks:" in rep and "pathPhases:" in rep, rep
    Fb = [g.get_forces_fixed(w).copy() for w in range(W)]
    cvb = [g.getLastCVValues(w).copy() for w in range(W)]
    for w in (0, 7, W - 1):
        o = OracleContext(f, n)
        o.set_box(L)
        r = o.execute(pos[w])
        check(g, o, r, g.getLastBias(w), w)
    # the one-walker launches of the same engine run the tuned shapes
    g.zero_forces()
    g.execute_all_plain()
    for w in range(W):
        assert np.array_equal(g.get_forces_fixed(w), Fb[w]), w
        assert np.array_equal(g.getLastCVValues(w), cvb[w]), w
    # the library's walker loop equals one glued_execute call per walker
    g.zero_forces()
    for w in range(W):
        g.execute(w, True, False)
    for w in range(W):
        assert np.array_equal(g.get_forces_fixed(w), Fb[w]), w
    # a second engine of the same plan shape takes the cached choice
    g2 = Context(f, walkers=[n] * W)
    g2.set_box(L)
    g2.set_all_positions(pos)
    g2.zero_forces()
    g2.execute_batch(use_graph=True)
    rep2 = g2.tuning()
    assert "(cached)" in rep2 and rep2.split(" (")[0] == rep.split(" (")[0], (rep, rep2)
    for w in range(W):
        assert np.array_equal(g2.get_forces_fixed(w), Fb[w]), w


def test_first_plain_evaluation_keeps_defaults(monkeypatch):
    """An engine whose first evaluation is one walker keeps the default
    shapes (no timing), so its later batched evaluations equal its plain ones."""
    from gluedmd_amd.context import Context
    monkeypatch.setenv("GLUED_AUTOTUNE", "0")  # no cached choice from other tests either way
    W, nA, nB = 4, 60, 600
    c = wl.config3_system(batch=W, M=64, N=20)
    f, _ = _path_coord_force(c, nA, nB)
    n = c["M"] + nA + nB
    L = 2.8
    pos = _positions(c, W, nA, nB, L, np.random.default_rng(53))
    g = Context(f, walkers=[n] * W)
    g.set_box(L)
    g.set_all_positions(pos)
    g.zero_forces()
    g.execute_all_plain()
    Fp = [g.get_forces_fixed(w).copy() for w in range(W)]
    g.zero_forces()
    g.execute_batch(use_graph=False)
    assert "(default)" in g.tuning() or "(cached)" in g.tuning()
    for w in range(W):
        assert np.array_equal(g.get_forces_fixed(w), Fp[w]), w
