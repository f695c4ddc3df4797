"""Pins of the oracle's PBMetaD (CPU): the reference's own test cases
(tests/test_bias_pbmetad.py in MarvinTaterra/GluedMD: zero before a
deposit, one deposit raising V to about h0), the reduction of one CV to 1-D
grid metadynamics (the oracle's independent BIAS_METAD path, over a
sequence of well-tempered deposits), the log-sum-exp combination written
without the minimum shift, and central finite differences."""
import math

import numpy as np

from gluedmd_amd.glued import Force


def _ctx(oracle, f, n):
    c = oracle.OracleContext(f, n)
    c.set_box(3.0)
    return c


def _pos(rng, n=8):
    return rng.uniform(0.2, 2.8, size=(n, 3)).astype(np.float32)


def test_pbmetad_zero_before_deposit(oracle):
    # test_bias_pbmetad.py:82-118: empty grids give V = 0 (one and two CVs)
    for ncv in (1, 2):
        f = Force(pbc=True, temperature=300.0)
        cvs = [f.add_distance([0, 1]), f.add_distance([2, 3])][:ncv]
        f.add_pbmetad(cvs, 0.1, 1.0, 1000, grid_min=0.0, grid_max=3.0, bins=60, gamma=10.0)
        r = _ctx(oracle, f, 8).execute(_pos(np.random.default_rng(1)))
        assert r["energy"] == 0.0
        assert np.all(r["dUdCV"] == 0.0)


def test_pbmetad_one_deposit_raises_energy(oracle):
    # test_bias_pbmetad.py:122-152: after one deposit at the current CV the
    # bias is about h0 (0.5 h0 < V < 1.5 h0)
    f = Force(pbc=True, temperature=300.0)
    d = f.add_distance([0, 1])
    f.add_pbmetad([d], 0.1, 2.0, 1, grid_min=0.0, grid_max=3.0, bins=120, gamma=10.0)
    c = _ctx(oracle, f, 8)
    pos = _pos(np.random.default_rng(2))
    c.execute(pos)
    c.update_state(1)
    E = c.execute(pos)["energy"]
    assert 1.0 < E < 3.0


def test_pbmetad_one_cv_is_grid_metad(oracle):
    # one CV: w = 1 and V = V_1, so PBMetaD is 1-D well-tempered grid MetaD
    # (heights h0 e^(-V hif), the same grid geometry and Gaussian cut)
    rng = np.random.default_rng(3)
    fp = Force(pbc=True, temperature=300.0)
    fm = Force(pbc=True, temperature=300.0)
    for f in (fp, fm):
        f.add_dihedral([0, 1, 2, 3])
    fp.add_pbmetad([0], 0.35, 1.2, 1, grid_min=-math.pi, grid_max=math.pi, bins=72, periodic=True, gamma=6.0)
    fm.add_metad([0], 0.35, 1.2, 1, grid_min=-math.pi, grid_max=math.pi, bins=72, periodic=True, gamma=6.0)
    cp, cm = _ctx(oracle, fp, 8), _ctx(oracle, fm, 8)
    base = _pos(rng)
    for step in range(8):
        cp.update_state(step)
        cm.update_state(step)
        p = (base + rng.normal(0, 0.2, size=base.shape)).astype(np.float32)
        rp, rm = cp.execute(p), cm.execute(p)
        assert abs(rp["energy"] - rm["energy"]) <= 1e-12 * max(1.0, abs(rm["energy"]))
        assert np.allclose(rp["dUdCV"], rm["dUdCV"], rtol=1e-12, atol=1e-12)
    assert rm["energy"] > 0.1


def test_pbmetad_log_sum_exp_and_fd(oracle):
    # two CVs: V = -kT log(mean_i e^(-V_i/kT)) from the sub-grid values, and
    # the forces are minus the central differences of V
    rng = np.random.default_rng(4)
    f = Force(pbc=True, temperature=300.0)
    a = f.add_distance([0, 1])
    t = f.add_dihedral([2, 3, 4, 5])
    f.add_pbmetad([a, t], [0.08, 0.3], 1.5, 1, grid_min=[0.0, -math.pi], grid_max=[3.0, math.pi], bins=[150, 90],
                  periodic=[False, True], gamma=8.0)
    c = _ctx(oracle, f, 8)
    base = _pos(rng)
    for step in range(6):
        c.update_state(step)
        r = c.execute((base + rng.normal(0, 0.05, size=base.shape)).astype(np.float32))
    st = c.states[0]
    kT = f.biases[0][2][2]
    assert abs(r["energy"] - (-kT * math.log(np.mean(np.exp(-st.Vi / kT))))) < 1e-9 * max(1.0, abs(r["energy"]))
    assert abs(st.wi.sum() - 1.0) < 1e-14
    # dV/ds_i = w_i dV_i/ds_i against central differences in CV space (V is
    # linear in each s_i inside a grid cell; positions are rounded to FP32
    # by the oracle, CVs are not)
    sv = r["cv"][[a, t]].copy()
    V0, g = st.eval(sv)
    h = 1e-7
    for i in range(2):
        sp, sm = sv.copy(), sv.copy()
        sp[i] += h
        sm[i] -= h
        fd = (st.eval(sp)[0] - st.eval(sm)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i]))
    assert np.allclose(r["dUdCV"][[a, t]], g, rtol=1e-12, atol=0.0)
