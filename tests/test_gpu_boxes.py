"""GPU parity under periodic boxes other than one shared cube: a triclinic
box in OpenMM's reduced form (deltaPBC applies the c, b, a images in that
order, CommonGluedKernels.cpp:74-112), and walkers whose boxes differ while
a cutoff coordination CV shares one cell grid between them."""
import numpy as np
import pytest

from gluedmd_amd.glued import Force
from tests.gpu_util import check, make_pair

pytestmark = pytest.mark.gpu

# a = (2.4, 0, 0), b = (0.9, 2.2, 0), c = (-1.1, 0.8, 2.6): reduced form
# (|b_x| <= a_x / 2, |c_x| <= a_x / 2, |c_y| <= b_y / 2) with every
# off-diagonal entry nonzero
TRICLINIC = [2.4, 0.0, 0.0, 0.9, 2.2, 0.0, -1.1, 0.8, 2.6]


def _fractional_to_cart(frac, box):
    B = np.asarray(box, dtype=np.float64).reshape(3, 3)
    return frac @ B


def _triclinic_force(nA, nB):
    f = Force(pbc=True)
    n = 12 + nA + nB
    f.add_distance([0, 1])
    f.add_angle([1, 2, 3])
    f.add_dihedral([2, 3, 4, 5])
    f.add_com_distance([0, 1, 2], [6, 7, 8, 9])
    f.add_gyration(list(range(10, 12 + nA)))
    A = list(range(12, 12 + nA))
    B = list(range(12 + nA, n))
    f.add_coordination(A, B, 0.35)                 # every A x B pair (the reference's form)
    f.add_coordination(A, B, 0.35, cutoff=0.9)     # cutoff; the cell list needs an orthorhombic box
    f.add_contact_map([0, 5, 1, 6, 2, 7, 3, 8], r0=0.6)
    f.add_drmsd([4, 9, 5, 10, 6, 11], [0.8, 1.1, 0.9])
    return f, n


@pytest.mark.parametrize("batched", [False, True])
def test_triclinic_cvs_and_coordination(batched):
    """Every CV kind with a periodic image on a triclinic box: values and
    forces (unit dU/dCV on every value) against the oracle, exact pair
    counts of both coordination CVs, on seeded configurations (one plain
    walker, three batched) whose atoms straddle every face."""
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    nA, nB = 40, 400
    f, n = _triclinic_force(nA, nB)
    f.setTestBiasGradients([1.0] * f.getNumCollectiveVariables())
    rng = np.random.default_rng(41)
    W = 3 if batched else 1
    g = Context(f, walkers=[n] * W)
    g.set_box(TRICLINIC)
    pos = [_fractional_to_cart(rng.uniform(-0.2, 1.2, size=(n, 3)), TRICLINIC).astype(np.float32) for _ in range(W)]
    g.set_all_positions(pos)
    g.zero_forces()
    if batched:
        g.execute_batch(use_graph=False)
    for w in range(W):
        o = OracleContext(f, n)
        o.set_box(TRICLINIC)
        if not batched:
            g.zero_forces()
            g.execute(w)
        r = o.execute(pos[w])
        check(g, o, r, g.getLastBias(w), w)
        assert [g.coordination_pairs(k, w) for k in range(2)] == o.coord_pairs, w
        assert 0 < o.coord_pairs[1] < o.coord_pairs[0] == nA * nB


def _coord_force(nA, nB, cutoff):
    f = Force(pbc=True)
    f.add_coordination(list(range(nA)), list(range(nA, nA + nB)), 0.3, cutoff=cutoff)
    f.add_distance([0, nA])
    f.setTestBiasGradients([0.7, -0.4])
    return f


def test_coordination_cells_follow_every_walker_box():
    """Two walkers with different cubic boxes share the coordination cell
    grid: it must use the smallest per-axis cell count over both boxes, and
    no cells at all once one box is too small for three cells per axis.
    The boxes are set walker 1 first, then walker 0 (larger), then walker 1
    is shrunk below three cells and grown again, then walker 0 shrunk
    (NPT-like).  After every change each walker's CN, pair count and
    forces match the oracle, in the batched and the plain form."""
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    nA, nB, rc = 120, 1500, 0.55
    n = nA + nB
    f = _coord_force(nA, nB, rc)
    g = Context(f, walkers=[n, n])
    rng = np.random.default_rng(43)
    frac = [rng.uniform(0, 1, size=(n, 3)) for _ in range(2)]
    boxes = {}
    # cells per axis after each change: (unset walker 0) -, 3, none, 4, 3
    schedule = [(1, 2.0), (0, 3.3), (1, 1.5), (1, 2.6), (0, 1.9)]
    for w_set, L in schedule:
        boxes[w_set] = L
        g.set_box(L, w_set)
        if len(boxes) < 2:
            continue
        pos = [(frac[w] * boxes[w]).astype(np.float32) for w in range(2)]
        g.set_all_positions(pos)
        for form in ("batched", "plain"):
            g.zero_forces()
            if form == "batched":
                g.execute_batch(use_graph=False)
            else:
                g.execute_all_plain()
            for w in range(2):
                o = OracleContext(f, n)
                o.set_box(boxes[w])
                r = o.execute(pos[w])
                check(g, o, r, g.getLastBias(w), w)
                assert g.coordination_pairs(0, w) == o.coord_pairs[0], (w_set, L, form, w)
