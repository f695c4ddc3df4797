"""GPU parity of every CV kind against the FP64 oracle (through the C-ABI),
with injected dU/dCV as the reference's CV tests do (setTestBiasGradients,
CommonGluedKernels.cpp:4670-4671), plus the reference tests' closed forms."""
import math

import numpy as np
import pytest

from gluedmd_amd.glued import Force
from tests.gpu_util import assert_close_rel, check, make_pair, run_both

pytestmark = pytest.mark.gpu


def test_distance_reference_cases():
    f = Force()
    f.add_distance([0, 1])
    f.setTestBiasGradients([1.0])
    g, o = make_pair(f, 3)
    run_both(g, o, [(0, 0, 0), (0.5, 0, 0), (5, 5, 5)])
    assert abs(g.getLastCVValues()[0] - 0.5) < 1e-7
    F = g.get_forces()
    assert abs(F[0][0] - 1.0) < 1e-9 and abs(F[1][0] + 1.0) < 1e-9
    fp = Force(pbc=True)
    fp.add_distance([0, 1])
    g, o = make_pair(fp, 2, box=1.0)
    run_both(g, o, [(0.005, 0, 0), (0.995, 0, 0)])
    assert abs(g.getLastCVValues()[0] - 0.01) < 1e-6


@pytest.mark.parametrize("pbc", [False, True])
def test_small_cvs_random(pbc):
    rng = np.random.default_rng(1)
    f = Force(pbc=pbc)
    f.add_distance([0, 5])
    f.add_angle([1, 2, 3])
    f.add_dihedral([4, 6, 8, 9])
    f.add_dihedral([6, 8, 9, 2])
    f.add_position(7, 1)
    f.setTestBiasGradients([0.7, -1.3, 2.1, 0.4, -0.9])
    g, o = make_pair(f, 10, box=2.5 if pbc else None)
    for _ in range(3):
        run_both(g, o, rng.uniform(0, 2.5, size=(10, 3)))


def test_coordination_reference_cases():
    f = Force()
    f.add_coordination([0], [1], r0=1.5, n=1, m=2)
    f.setTestBiasGradients([1.0])
    g, o = make_pair(f, 3)
    run_both(g, o, [(0, 0, 0), (0.5, 0, 0), (5, 5, 5)])
    assert abs(g.getLastCVValues()[0] - 0.75) < 1e-6
    assert abs(g.get_forces()[0][0] + 0.375) < 1e-6
    f = Force()
    f.add_coordination([0], [1], r0=1.0, n=6, m=12)
    g, o = make_pair(f, 3)
    run_both(g, o, [(0, 0, 0), (2, 0, 0), (5, 5, 5)])
    assert abs(g.getLastCVValues()[0] - 63.0 / 4095.0) < 1e-7


@pytest.mark.parametrize("cutoff,pbc,n", [(None, False, 300), (1.0, True, 3000), (0.9, True, 300)])
def test_coordination_random(cutoff, pbc, n):
    rng = np.random.default_rng(2)
    L = (n / 100.0) ** (1 / 3)
    nA = n // 10
    f = Force(pbc=pbc)
    f.add_coordination(list(range(nA)), list(range(nA, n)), r0=0.4, n=6, m=12, cutoff=cutoff)
    f.setTestBiasGradients([1.0])
    g, o = make_pair(f, n, box=L)
    run_both(g, o, rng.uniform(0, L, size=(n, 3)))
    if cutoff:
        assert g.coordination_pairs(0) == o.coord_pairs[0]


def test_coordination_per_atom_neighbour_counts():
    """Neighbour counts per group-A atom, exact: 48 coordination CVs that
    share one B group of 3,000 atoms (cell list, cutoff 0.9 nm) and each hold
    one A atom, so every CV's in-cutoff pair count is one A atom's neighbour
    count; every count and every CN against the oracle, batched and plain."""
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    rng = np.random.default_rng(23)
    nB, nA, L = 3000, 48, 3.1
    n = nA + nB
    f = Force(pbc=True)
    B = list(range(nA, n))
    for a in range(nA):
        f.add_coordination([a], B, r0=0.3, n=6, m=12, cutoff=0.9)
    f.setTestBiasGradients([1.0] * nA)
    pos = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    o = OracleContext(f, n)
    o.set_box(L)
    r = o.execute(pos)
    assert len(set(o.coord_pairs)) > 5  # the counts differ from atom to atom
    g = Context(f, walkers=[n, n])
    g.set_box(L)
    g.set_all_positions([pos, pos])
    for form in ("batched", "plain"):
        g.zero_forces()
        if form == "batched":
            g.execute_batch(use_graph=False)
        else:
            g.execute_all_plain()
        for w in range(2):
            got = [g.coordination_pairs(k, w) for k in range(nA)]
            assert got == list(o.coord_pairs), (form, w)
            cv = g.getLastCVValues(w)
            assert np.all(np.abs(cv - r["cv"]) <= 1e-6 * np.abs(r["cv"]) + 1e-12), form


def test_coordination_cells_bit_reproducible():
    """The cell-list sort keeps list order within a cell, so CN, the pair
    count and every Jacobian row come out with the same bits on every
    evaluation, in the batched and the one-walker launches alike (3 walkers,
    ~40 atoms per cell, 12 evaluations each)."""
    from gluedmd_amd.context import Context
    rng = np.random.default_rng(17)
    n, L, W = 8000, 4.4, 3
    nA = 1000
    f = Force(pbc=True)
    f.add_coordination(list(range(nA)), list(range(nA, n)), r0=0.4, n=6, m=12, cutoff=1.2)
    f.setTestBiasGradients([1.0])
    g = Context(f, walkers=[n] * W)
    g.set_box(L)
    g.set_all_positions([rng.uniform(0, L, size=(n, 3)).astype(np.float32) for _ in range(W)])
    ref = None
    for rep in range(12):
        g.zero_forces()
        if rep % 2:
            g.execute_all_plain()
        else:
            g.execute_batch(use_graph=False)
        got = [(g.getLastCVValues(w).copy(), g.get_forces_fixed(w).copy(), g.coordination_pairs(0, w))
               for w in range(W)]
        if ref is None:
            ref = got
            continue
        for w in range(W):
            assert np.array_equal(got[w][0], ref[w][0]), (rep, w)
            assert np.array_equal(got[w][1], ref[w][1]), (rep, w)
            assert got[w][2] == ref[w][2]


def test_group_cvs_random():
    rng = np.random.default_rng(3)
    n = 700
    ref = rng.uniform(0, 2, size=(200, 3))
    f = Force(pbc=True)
    f.add_gyration(list(range(0, 500)))
    f.add_com_distance(list(range(0, 50)), list(range(50, 120)))
    f.add_rmsd(list(range(300, 500)), ref.tolist())
    f.add_rmsd(list(range(300, 500)), ref.tolist(), fit=True)
    pairs = rng.integers(0, n, size=(150, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    f.add_drmsd(pairs.reshape(-1).tolist(), rng.uniform(0.2, 1.0, size=len(pairs)).tolist())
    f.add_contact_map(pairs.reshape(-1).tolist(), r0=0.5, n=6, m=12, w=1.0)
    grads = rng.normal(size=f.getNumCollectiveVariables())
    f.setTestBiasGradients(grads.tolist())
    masses = rng.uniform(1, 16, size=n)
    g, o = make_pair(f, n, box=(n / 100) ** (1 / 3), masses=masses)
    pos = rng.uniform(0, (n / 100) ** (1 / 3), size=(n, 3))
    pos[300:500] = ref @ np.linalg.qr(rng.normal(size=(3, 3)))[0].T + rng.normal(0, 0.1, size=(200, 3)) + 1.0
    run_both(g, o, pos)


def test_kabsch_rotated_copy_is_zero():
    rng = np.random.default_rng(4)
    ref = rng.uniform(0, 2, size=(500, 3))
    f = Force()
    f.add_rmsd(list(range(500)), ref.tolist(), fit=True)
    g, o = make_pair(f, 500)
    Q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    g.set_positions((ref @ Q.T + 0.7).astype(np.float32))
    g.execute()
    assert g.getLastCVValues()[0] < 1e-5


def test_path_random():
    rng = np.random.default_rng(5)
    M, N = 40, 12
    base = rng.uniform(0, 2, size=(M, 3))
    frames = [base + 0.05 * k * rng.normal(size=(M, 3)) for k in range(N)]
    f = Force()
    f.add_path(list(range(M)), [fr.reshape(-1).tolist() for fr in frames], 30.0)
    f.setTestBiasGradients([1.3, -0.6])
    g, o = make_pair(f, M + 5)
    pos = np.vstack([base + rng.normal(0, 0.05, size=(M, 3)), rng.uniform(0, 2, size=(5, 3))])
    run_both(g, o, pos)


@pytest.mark.parametrize("M,N,batched", [(1100, 37, False), (1100, 37, True), (70, 300, True)])
def test_path_tiles(M, N, batched):
    """Path shapes that exercise the work split: atom counts above the LDS
    staging chunk (1,024) and not a multiple of the 64-atom row tile, frame
    counts not a multiple of the 16-frame MSD block, two paths in one force,
    plain and batched forms."""
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    rng = np.random.default_rng(11)
    base = rng.uniform(0, 4, size=(M, 3))
    frames = [base + 0.02 * k * rng.normal(size=(M, 3)) for k in range(N)]
    f = Force()
    f.add_path(list(range(M)), [fr.reshape(-1).tolist() for fr in frames], 40.0)
    sub = list(range(0, M, 3))[:50]
    f.add_path(sub, [fr[sub].reshape(-1).tolist() for fr in frames[:9]], 25.0)
    f.setTestBiasGradients([1.3, -0.6, 0.7, 0.2])
    W = 2 if batched else 1
    g = Context(f, walkers=[M + 3] * W)
    pos = [np.vstack([base + rng.normal(0, 0.04, size=(M, 3)), rng.uniform(0, 2, size=(3, 3))]).astype(np.float32)
           for _ in range(W)]
    for w in range(W):
        g.set_positions(pos[w], w)
    g.zero_forces()
    if batched:
        g.execute_batch(use_graph=True)
    else:
        g.execute(0)
    for w in range(W):
        o = OracleContext(f, M + 3)
        r = o.execute(pos[w])
        check(g, o, r, g.getLastBias(w), w)


def test_path_all_weights_underflow():
    """Every frame's weight underflows (lambda * msd > 745 for all frames):
    s = z = 0 as in the reference, rows 0 instead of the reference's division
    by S = 0; no NaN reaches the force buffer."""
    rng = np.random.default_rng(12)
    M, N = 30, 5
    base = rng.uniform(0, 2, size=(M, 3))
    frames = [base + 0.01 * rng.normal(size=(M, 3)) for _ in range(N)]
    f = Force()
    f.add_path(list(range(M)), [fr.reshape(-1).tolist() for fr in frames], 1e5)
    f.setTestBiasGradients([1.0, 1.0])
    g, o = make_pair(f, M)
    pos = (base + 1.0).astype(np.float32)  # msd = 3 nm^2 to every frame
    eg, r = run_both(g, o, pos, compare=False)
    assert np.array_equal(g.getLastCVValues(), [0.0, 0.0]) and np.array_equal(r["cv"], [0.0, 0.0])
    assert np.array_equal(g.get_forces_fixed(), np.zeros_like(g.get_forces_fixed()))
    assert np.all(r["forces"] == 0.0)


def test_atom_reorder():
    """GPU atom order != user order (rebuildGpuAtomIndices,
    CommonCVKernels.cpp:113-317): forces land on the GPU index."""
    rng = np.random.default_rng(6)
    f = Force()
    f.add_distance([0, 3])
    f.add_dihedral([1, 2, 3, 4])
    f.setTestBiasGradients([1.0, 0.5])
    g, o = make_pair(f, 6)
    pos = rng.uniform(0, 2, size=(6, 3)).astype(np.float32)
    perm = rng.permutation(6)  # gpu index -> user atom
    g.set_atom_order(perm)
    gpu_pos = pos[perm]
    g.set_positions(gpu_pos)
    g.zero_forces()
    g.execute()
    r = o.execute(pos)
    assert_close_rel(g.getLastCVValues(), r["cv"], 1e-6, "CVs")
    Fg = g.get_forces()  # in GPU order
    Fu = np.zeros_like(Fg)
    Fu[perm] = Fg
    assert np.allclose(Fu, r["forces"], atol=1e-6)
