"""glued_hill_adjoint_noalloc (the hill adjoint with its partials in the
library's fixed device array) against glued_hill_adjoint through the C-ABI:
bit-identical height and sigma gradients on two calls in a row, the second
on a smaller problem, and a clear error for a problem beyond the array."""
import ctypes

import numpy as np
import pytest
import torch

from gluedmd_amd import _lib
from gluedmd_amd.glued import Force

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _hill_context(W, cap, n_hills, seed, D=2, shared=False):
    from gluedmd_amd.context import Context
    f = Force(pbc=True, temperature=300.0)
    cvs = [f.add_distance([0, 1]), f.add_distance([0, 2]), f.add_distance([1, 2])][:D]
    hb = f.add_metad_hills(cvs, [0.05] * D, 1.0, 1, gamma=8.0, max_hills=cap)
    if shared:
        f.share_bias(hb)
    N = 4
    g = Context(f, walkers=[N] * W)
    g.set_box(3.0)
    rng = np.random.default_rng(seed)
    for w in range(1 if shared else W):
        c = rng.uniform(0.7, 1.3, size=(n_hills, D))
        h = rng.uniform(0.1, 1.2, size=n_hills)
        g.set_hills(hb, c, h, w)
    pos = [np.array([(0, 0, 0), (1.0, 0, 0), (0, 1.0, 0), (2, 2, 2)]) + rng.normal(0, 0.08, size=(N, 3))
           for _ in range(W)]
    g.set_all_positions([p.astype(np.float32) for p in pos])
    g.zero_forces()
    g.execute_batch(use_graph=False)
    T = 1 if shared else W
    gV = torch.as_tensor(rng.normal(size=W), dtype=torch.float64, device=g.device)
    return g, hb, gV, T * cap, D


def _adjoint(g, hb, gV, nh, D, name):
    gh = torch.zeros(nh, dtype=torch.float64, device=g.device)  # entries past a table's count stay 0
    gs = torch.full((D,), np.nan, dtype=torch.float64, device=g.device)
    _lib.call(name, g.h, g.stream, hb, _p(gV), _p(gh), _p(gs))
    torch.cuda.synchronize()
    return gh.cpu().numpy(), gs.cpu().numpy()


def test_noalloc_adjoint_matches_allocating_form():
    big = _hill_context(W=8, cap=6000, n_hills=5000, seed=1)
    small = _hill_context(W=3, cap=1500, n_hills=1200, seed=2, D=1, shared=True)
    # two calls in a row into the one array, the second smaller
    out_big = _adjoint(*big, "glued_hill_adjoint_noalloc")
    out_small = _adjoint(*small, "glued_hill_adjoint_noalloc")
    ref_big = _adjoint(*big, "glued_hill_adjoint")
    ref_small = _adjoint(*small, "glued_hill_adjoint")
    for got, want in ((out_big, ref_big), (out_small, ref_small)):
        assert np.isfinite(want[0]).all() and np.isfinite(want[1]).all() and np.isfinite(got[1]).all()
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        assert np.abs(want[0]).max() > 0 and np.abs(want[1]).max() > 0
    # and again after the smaller call: the array keeps no state between calls
    again = _adjoint(*big, "glued_hill_adjoint_noalloc")
    assert np.array_equal(again[0], ref_big[0]) and np.array_equal(again[1], ref_big[1])


def test_noalloc_adjoint_rejects_oversized_problem():
    """ceil(200,000 / 256) * 40 tables * 3 CVs = 93,840 doubles of partials,
    beyond the 65,536 of the fixed array: an error, and nothing launched."""
    from gluedmd_amd.context import Context
    f = Force(pbc=True, temperature=300.0)
    cvs = [f.add_distance([0, 1]), f.add_distance([0, 2]), f.add_distance([1, 2])]
    hb = f.add_metad_hills(cvs, [0.05] * 3, 1.0, 1, gamma=8.0, max_hills=200000)
    g = Context(f, walkers=[4] * 40)
    gV = torch.zeros(40, dtype=torch.float64, device=g.device)
    gh = torch.zeros(40 * 200000, dtype=torch.float64, device=g.device)
    gs = torch.zeros(3, dtype=torch.float64, device=g.device)
    rc = _lib.load().glued_hill_adjoint_noalloc(g.h, g.stream, hb, _p(gV), _p(gh), _p(gs))
    assert rc != 0
    assert b"workspace" in _lib.load().glued_last_error()
