"""GPU parity of the OPES paths beyond the plain fixed-sigma table: the
expanded ensemble with its running log Z, the Welford-adaptive bandwidth,
and one kernel table shared by several walkers.  Every step is compared with
the FP64 oracle; kernel counts are exact."""
import numpy as np
import pytest

from gluedmd_amd.glued import Force
from tests.gpu_util import assert_close_rel, check

pytestmark = pytest.mark.gpu


def _walker_oracles(f, N, W, box=None, shared=None):
    from oracle.oracle import OracleContext
    out = []
    for _ in range(W):
        o = OracleContext(f, N, shared=shared)
        if box is not None:
            o.set_box(box)
        out.append(o)
    return out


def _expanded_force():
    """Five position ECVs, three of them used by the expanded bias in the
    order (3, 0, 4): non-consecutive value indices, as the reference accepts
    (cvValues[cvIdxList[l]], CommonGluedKernels.cpp:3050-3125)."""
    f = Force(pbc=True, temperature=300.0)
    ecv = [f.add_position(a, 0) for a in range(5)]
    d = f.add_distance([5, 6])
    f.add_opes_expanded([ecv[3], ecv[0], ecv[4]], weights=[1.0, 2.0, 3.0], pace=1, temperature=300.0)
    f.add_harmonic(d, 50.0, 1.0)
    return f


@pytest.mark.parametrize("form", ["batched", "plain", "step_graph"])
def test_opes_expanded_logz_sequence(form):
    """Three walkers, 24 steps with a log Z update every step: energies,
    dU/dCV (the state probabilities) and forces against the oracle at every
    step, in the batched, plain and one-graph-per-step forms."""
    from gluedmd_amd.context import Context
    W, N = 3, 8
    f = _expanded_force()
    g = Context(f, walkers=[N] * W)
    g.set_box(4.0)
    os_ = _walker_oracles(f, N, W, box=4.0)
    rng = np.random.default_rng(17)
    base = [rng.uniform(0.5, 3.5, size=(N, 3)) for _ in range(W)]
    for step in range(24):
        pos = [(b + rng.normal(0, 0.3, size=b.shape)).astype(np.float32) for b in base]
        g.set_all_positions(pos)
        g.zero_forces()
        if form == "plain":
            g.update_state(step)
            g.execute_all_plain()
        elif form == "batched":
            g.update_state(step)
            g.execute_batch(use_graph=False)
        else:
            g.step_batch(step)
        for w, o in enumerate(os_):
            o.update_state(step)
            r = o.execute(pos[w])
            check(g, o, r, g.getLastBias(w), w)
    for o in os_:
        assert int(o.states[0].nUpd[0]) == 23


def _adaptive_force(shared=False, stride=6, pace=2):
    f = Force(pbc=True, temperature=300.0)
    a = f.add_distance([0, 1])
    b = f.add_distance([0, 2])
    ob = f.add_opes([a, b], None, gamma=12.0, pace=pace, sigma_min=1e-3, max_kernels=400,
                    adaptive_sigma_stride=stride)
    if shared:
        f.share_bias(ob)
    return f, ob


def _walk(rng, W, N, steps, spread=0.04):
    base = [np.concatenate([[(0, 0, 0), (1.0, 0, 0), (0, 1.0, 0)],
                            rng.uniform(0.2, 2.8, size=(N - 3, 3))]) for _ in range(W)]
    x = [b.copy() for b in base]
    for _ in range(steps):
        for w in range(W):
            x[w][1:3] += rng.normal(0, spread, size=(2, 3))
        yield [p.astype(np.float32) for p in x]


def test_opes_adaptive_sigma_sequence():
    """sigma=None: the Welford running variance of every evaluated step
    (opesAccumulateWelford, CommonGluedKernels.cpp:3009-3027) sets the
    bandwidth, and deposits start once the warm-up stride has passed
    (:2784).  Two walkers with private tables, 60 steps, pace 2, stride 6:
    energies and forces every step, then kernel counts exact, sigmas and
    metrics within 1e-6."""
    from gluedmd_amd.context import Context
    W, N = 2, 6
    f, ob = _adaptive_force(shared=False)
    g = Context(f, walkers=[N] * W)
    g.set_box(3.0)
    os_ = _walker_oracles(f, N, W, box=3.0)
    rng = np.random.default_rng(23)
    deposits_seen = 0
    for step, pos in enumerate(_walk(rng, W, N, 60)):
        g.update_state(step)
        g.set_all_positions(pos)
        g.zero_forces()
        g.execute_batch(use_graph=False)
        for w, o in enumerate(os_):
            o.update_state(step)
            r = o.execute(pos[w])
            check(g, o, r, g.getLastBias(w), w)
        deposits_seen = max(deposits_seen, os_[0].states[ob].nk)
    assert deposits_seen >= 3, "the adaptive path never deposited"
    for w, o in enumerate(os_):
        st = o.states[ob]
        m = g.getOPESMetrics(ob, w)
        assert int(m[2]) == st.nk, (w, m[2], st.nk)
        assert_close_rel(m, st.metrics(), 1e-6, "OPES metrics")
        _, s, _ = st.table()
        assert np.allclose(g.getKernelSigmas(ob, w), s.reshape(-1), rtol=1e-9)
        # the bandwidth came from the Welford variance, not from sigma0 = 0
        assert np.all(s > 1e-3)


@pytest.mark.parametrize("adaptive", [False, True])
def test_opes_shared_table_four_walkers(adaptive):
    """One OPES state for four walkers (the multi-walker sharing of
    CommonGluedKernels.cpp:2912-2933 / :5234-5262): every step the walkers'
    deposits enter the one table in walker order, merging into a near kernel
    or appending.  40 steps; walkers 0/1 sample one basin (merges), 2/3 walk
    apart (appends).  Kernel count exact, metrics within 1e-6, every walker's
    energy and forces every step; the step graph gives the same bits as
    update_state + the batched evaluation."""
    from gluedmd_amd.context import Context
    from oracle.oracle import _Opes
    W, N = 4, 6
    if adaptive:
        f, ob = _adaptive_force(shared=True, stride=4, pace=1)
    else:
        f = Force(pbc=True, temperature=300.0)
        a = f.add_distance([0, 1])
        b = f.add_distance([0, 2])
        ob = f.add_opes([a, b], [0.05, 0.04], gamma=10.0, pace=1, max_kernels=400)
        f.share_bias(ob)
    g = Context(f, walkers=[N] * W)
    g2 = Context(f, walkers=[N] * W)
    g.set_box(3.0)
    g2.set_box(3.0)
    _, cvs, params, ip = f.biases[ob]
    common = {ob: _Opes(len(cvs), params, ip)}
    os_ = _walker_oracles(f, N, W, box=3.0, shared=common)
    rng = np.random.default_rng(29)
    spreads = [0.005, 0.005, 0.08, 0.08]
    base = [np.concatenate([[(0, 0, 0), (1.0, 0, 0), (0, 1.0, 0)], rng.uniform(0.2, 2.8, size=(N - 3, 3))])
            for _ in range(W)]
    x = [b.copy() for b in base]
    for step in range(40):
        for w in range(W):
            x[w][1:3] += rng.normal(0, spreads[w], size=(2, 3))
        pos = [p.astype(np.float32) for p in x]
        g.update_state(step)
        g.set_all_positions(pos)
        g.zero_forces()
        g.execute_batch(use_graph=False)
        g2.set_all_positions(pos)
        g2.zero_forces()
        g2.step_batch(step)
        for w, o in enumerate(os_):
            o.update_state(step)
        for w, o in enumerate(os_):
            r = o.execute(pos[w])
            check(g, o, r, g.getLastBias(w), w)
            assert np.array_equal(g.get_forces_fixed(w), g2.get_forces_fixed(w)), (step, w)
    st = common[ob]
    deposits = 39 * W if not adaptive else None
    for w in range(W):
        m = g.getOPESMetrics(ob, w)
        assert int(m[2]) == st.nk, (w, m[2], st.nk)
        assert_close_rel(m, st.metrics(), 1e-6, "OPES metrics")
        assert np.array_equal(g.getOPESMetrics(ob, w), g2.getOPESMetrics(ob, w))
    assert st.nk > 4, "no appends"
    if deposits is not None:
        assert st.nk < deposits, "no merges"
    _, s, _ = st.table()
    assert np.allclose(g.getKernelSigmas(ob, 0), s.reshape(-1), rtol=1e-9)


@pytest.mark.parametrize("shared", [False, True])
def test_opes_large_table_deposits(shared):
    """Tables past the deposit kernels' staged prefixes: 9,000 kernels in a
    private table (the 16 deposit workgroups load two kernels per thread,
    8,192 in all, before the search; the rest stream) and 4,500 kernels
    shared by two walkers (the one-workgroup form stages 4,000 kernels of a
    2-CV table in LDS; the rest stream).  Each step puts the CVs 0.1 sigma
    from a kernel past the prefix (a merge there) or far from every kernel
    (an append).  Kernel count exact, metrics within 1e-6, energies and
    forces every step."""
    from gluedmd_amd import workloads as wl
    from gluedmd_amd.context import Context
    from oracle.oracle import _Opes
    W = 2 if shared else 1
    n = 4500 if shared else 9000
    sig = np.array([0.05, 0.04])
    f = Force(pbc=True, temperature=300.0)
    a = f.add_distance([0, 1])
    b = f.add_distance([0, 2])
    ob = f.add_opes([a, b], sig.tolist(), gamma=10.0, pace=1, max_kernels=n + 64)
    if shared:
        f.share_bias(ob)
    rng = np.random.default_rng(41)
    centers, sigmas, logw = wl.opes_prior_kernels(wl.SplitMix64(41), [1.0, 1.0], sig, 10.0, n=n, spread=6.0)
    _, cvs, params, ip = f.biases[ob]
    common = _Opes(len(cvs), params, ip)
    su = wl.sum_uprob_np(centers, sigmas, logw, common.cutoff2)
    common.load(centers, sigmas, logw, su)
    g = Context(f, walkers=[3] * W)
    g.set_box(8.0)
    g.setBiasState(wl.opes_state_blob(centers, sigmas, logw, su))
    if shared:
        os_ = _walker_oracles(f, 3, W, box=8.0, shared={ob: common})
    else:
        os_ = _walker_oracles(f, 3, W, box=8.0)
        os_[0].states[ob] = common
    lo = 4000 if shared else 8192
    targets = [k for k in range(lo, n) if min(centers[k]) > 0.3]
    for step in range(6):
        pos = []
        for w in range(W):
            if (step + w) % 3 == 2:
                d1, d2 = 3.2 + 0.1 * step, 3.2 + 0.07 * w  # far from every prior kernel
            else:
                d1, d2 = centers[targets[int(rng.integers(len(targets)))]] + 0.1 * sig
            pos.append(np.array([(0, 0, 0), (d1, 0, 0), (0, d2, 0)], dtype=np.float32))
        g.update_state(step)
        g.set_all_positions(pos)
        g.zero_forces()
        g.execute_batch(use_graph=False)
        for o in os_:
            o.update_state(step)
        for w, o in enumerate(os_):
            r = o.execute(pos[w])
            check(g, o, r, g.getLastBias(w), w)
    for w in range(W):
        m = g.getOPESMetrics(ob, w)
        assert int(m[2]) == common.nk, (w, m[2], common.nk)
        assert_close_rel(m, common.metrics(), 1e-6, "OPES metrics")
    assert common.nk > n, "no appends"
    assert common.nk < n + 6 * W, "no merges"
    _, s, _ = common.table()
    assert np.allclose(g.getKernelSigmas(ob, 0), s.reshape(-1), rtol=1e-9)
