"""Edge cases of the GPU path against the FP64 oracle: the atom counts at
which work moves between kernels, and a hill list that starts empty and is
filled to its capacity."""
import math

import numpy as np
import pytest

from gluedmd_amd.glued import Force
from tests.gpu_util import make_pair, run_both

pytestmark = pytest.mark.gpu


def _batched_equals_plain(g):
    """One walker: the batched launch gives the plain launch's fixed-point forces."""
    F = g.get_forces_fixed(0)
    g.zero_forces()
    g.execute_batch(use_graph=False)
    assert np.array_equal(g.get_forces_fixed(0), F)


@pytest.mark.parametrize("n", [3, 64, 1024, 1025])
def test_kabsch_size_boundaries(n):
    """Kabsch RMSD on the register path (up to 1,024 atoms, 4 per thread) and
    on the general group path above it."""
    rng = np.random.default_rng(100 + n)
    ref = rng.uniform(0, 2, size=(n, 3))
    f = Force()
    f.add_rmsd(list(range(n)), ref.tolist(), fit=True)
    f.setTestBiasGradients([0.7])
    g, o = make_pair(f, n + 3)
    Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    pos = np.vstack([ref @ Q.T + rng.normal(0, 0.1, size=(n, 3)) + 1.0, rng.uniform(0, 2, size=(3, 3))])
    run_both(g, o, pos)
    _batched_equals_plain(g)


@pytest.mark.parametrize("n", [2, 512, 513, 1024, 1025, 5000])
def test_group_size_boundaries(n):
    """Gyration over n atoms and a contact map over n pairs, on both sides of
    the 1,024-entry split between the 256- and the 1,024-thread group kernels
    and of the 512-pair chunk of a contact map (longer maps are summed from
    chunk partials by the last chunk to finish)."""
    rng = np.random.default_rng(200 + n)
    N = max(2 * n + 10, 60)
    L = (N / 100) ** (1 / 3)
    f = Force(pbc=True)
    f.add_gyration(list(range(n)))
    pairs = rng.integers(0, N, size=(n + 8, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]][:n]
    f.add_contact_map(pairs.reshape(-1).tolist(), r0=0.5)
    f.setTestBiasGradients(rng.normal(size=2).tolist())
    g, o = make_pair(f, N, box=L)
    run_both(g, o, rng.uniform(0, L, size=(N, 3)))
    _batched_equals_plain(g)


def test_hill_list_empty_then_full():
    """Hill MetaD that starts with no hills (V = 0, no force) and deposits
    every step past its capacity: the count stops at the capacity, and
    energies, forces and the stored hills follow the oracle."""
    cap = 5
    f = Force(pbc=True, temperature=300.0)
    d = f.add_distance([0, 1])
    t = f.add_dihedral([2, 3, 4, 5])
    hb = f.add_metad_hills([d, t], sigma=[0.1, 0.3], height=1.0, pace=1, gamma=8.0, max_hills=cap,
                           period=[0.0, 2 * math.pi])
    g, o = make_pair(f, 20, box=2.0)
    rng = np.random.default_rng(9)
    base = rng.uniform(0, 2, size=(20, 3))
    eg, r = run_both(g, o, base)
    assert eg == 0.0 and r["energy"] == 0.0
    assert np.abs(g.get_forces(0)).max() == 0.0
    for step in range(1, cap + 4):
        g.update_state(step)
        o.update_state(step)
        run_both(g, o, base + rng.normal(0, 0.02, size=base.shape))
    hc, hh = g.get_hills(hb)
    assert len(hh) == cap == len(o.states[hb].heights)
    # heights h0 exp(-V hif) carry V's tolerance (bias energies: 1e-6 relative)
    assert np.allclose(hh, o.states[hb].heights, rtol=1e-6, atol=0.0)
    _batched_equals_plain(g)
