"""C-ABI and host logic on the CPU (no GPU calls are made): the library loads
and exports every entry point include/glued.h declares, the specification
store and the initialization-time validation behave like the reference's
(GluedForce.cpp:82-98, CommonGluedKernels.cpp:3879-3927), and the Python
Force assembles the reference's parameter layouts (python/glued.py)."""
import ctypes
import math
import os
import re

import pytest

from gluedmd_amd import _lib
from gluedmd_amd.glued import Force, multithermal_temperature_ladder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "glued.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(glued_\w+)\s*\(", src, re.M)))


def test_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    # the ctypes signature table covers the same set
    assert set(_lib.SIGNATURES) | {"glued_last_error"} == set(names)


def arr(kind, vals):
    a = (kind * max(len(vals), 1))(*vals)
    return a


def handle_with(cvs, biases=()):
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.glued_create(ctypes.byref(h)) == 0
    out = ctypes.c_int()
    for t, atoms, params in cvs:
        assert L.glued_add_cv(h, t, arr(ctypes.c_int, atoms), len(atoms), arr(ctypes.c_double, params), len(params),
                              ctypes.byref(out)) == 0
    for t, c, p, ip in biases:
        rc = L.glued_add_bias(h, t, arr(ctypes.c_int, c), len(c), arr(ctypes.c_double, p), len(p),
                              arr(ctypes.c_int, ip), len(ip), ctypes.byref(out))
        if rc:
            return h, L.glued_last_error().decode()
    return h, None


def init_error(cvs, biases=(), n=10):
    L = _lib.load()
    h, err = handle_with(cvs, biases)
    if err:
        L.glued_destroy(h)
        return err
    na = arr(ctypes.c_int, [n])
    rc = L.glued_initialize_batch(h, 1, na, na, None, 0)
    msg = L.glued_last_error().decode()
    L.glued_destroy(h)
    assert rc != 0
    return msg


def test_value_indices_path_takes_two():
    L = _lib.load()
    h = ctypes.c_void_p()
    L.glued_create(ctypes.byref(h))
    out = ctypes.c_int()
    a = arr(ctypes.c_int, [0, 1])
    L.glued_add_cv(h, 1, a, 2, arr(ctypes.c_double, [0]), 0, ctypes.byref(out))
    assert out.value == 0
    p = [1.0, 2.0] + [0.0] * 12
    L.glued_add_cv(h, 10, a, 2, arr(ctypes.c_double, p), len(p), ctypes.byref(out))
    assert out.value == 1
    L.glued_add_cv(h, 1, a, 2, arr(ctypes.c_double, [0]), 0, ctypes.byref(out))
    assert out.value == 3
    n = ctypes.c_int()
    L.glued_num_cv_values(h, ctypes.byref(n))
    assert n.value == 4
    L.glued_destroy(h)


def test_bias_cv_index_range():
    h, err = handle_with([(1, [0, 1], [])], [(3, [1], [1.0], [1])])
    _lib.load().glued_destroy(h)
    assert "out of range" in err


@pytest.mark.parametrize("cvs,expect", [
    ([(1, [0, 12], [])], "out of range"),
    ([(3, [0, 1, 2], [])], "expected >= 4 atoms"),
    ([(6, [5, 0, 1], [0.4, 6, 12])], "group-A size"),
    ([(6, [1, 0, 1], [0.4, 6])], "expected >= 3 parameters"),
    ([(7, [0, 1], [0.0] * 5)], "expected >= 6 parameters"),
    ([(10, [0], [1.0, 1.0, 0, 0, 0])], "path needs >= 2 frames"),
    ([(16, [0, 1, 2, 3, 4], [5, 0])], "not supported"),
])
def test_initialize_validation(cvs, expect):
    assert expect in init_error(cvs)


def test_bias_layout_validation():
    err = init_error([(1, [0, 1], [])], [(3, [0], [1.0, 0.1], [1])])
    assert "METAD params" in err
    err = init_error([(1, [0, 1], [])], [(15, [0], [2.5, 0.4], [1])])
    assert "CV_ENERGY" in err
    err = init_error([(1, [0, 1], [])], [(5, [0], [2.5, 1.0, 0.1, 0.01], [2, 1, 10])])
    assert "gamma > 1" in err


def test_no_device_fails_cleanly():
    """Initialization on a host without a HIP device reports an error
    instead of crashing (this container has no GPU)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    msg = init_error([(1, [0, 1], [])])
    assert "hip" in msg.lower() or "device" in msg.lower()


# ---------------------------------------------------------------------------
# Python parameter assembly (python/glued.py in the reference)
# ---------------------------------------------------------------------------
def test_param_layouts_match_reference():
    f = Force(temperature=300.0)
    cn = f.add_coordination([1, 2], [3, 4, 5], r0=0.4)
    assert f.cvs[cn] == (6, [2, 1, 2, 3, 4, 5], [0.4, 6.0, 12.0])          # glued.py:173-176
    cm = f.add_contact_map([0, 1, 2, 3], r0=0.5, w=[1.0, 2.0])
    assert f.cvs[-1][2] == [0.5, 6.0, 12.0, 1.0, 0.5, 6.0, 12.0, 2.0]     # glued.py:216-222
    f.add_metad([cn, cm], sigma=[0.1, 0.2], height=1.2, pace=10, grid_min=[0, 0], grid_max=[5, 4],
                bins=[50, 40], periodic=[False, True], gamma=6.0)
    kT = 8.314462618e-3 * 300.0
    t, c, p, ip = f.biases[-1]
    assert p == [1.2, 0.1, 0.2, 6.0, kT, 0.0, 0.0, 5.0, 4.0] and ip == [10, 50, 40, 0, 1]   # glued.py:639-642
    f.add_opes([cn], sigma=0.05, gamma=12.0, pace=5)
    t, c, p, ip = f.biases[-1]
    assert p == [kT, 12.0, 0.05, 0.05 * 0.05] and ip == [0, 5, 100000]    # glued.py:602-612
    f.add_opes(cn, sigma=None, pace=3)
    assert f.biases[-1][3] == [0, 3, 100000, 30] and f.biases[-1][2][2] == 0.0
    with pytest.raises(ValueError):
        f.add_opes(cn, sigma=None, mode="fixed_uniform")
    with pytest.raises(ValueError):
        f.add_opes(cn, sigma=0.1, mode="explore", gamma=math.inf)


def test_multithermal_ladder_and_params():
    lad = multithermal_temperature_ladder(300.0, 500.0, 32)
    assert abs(lad[0] - 300) < 1e-12 and abs(lad[-1] - 500) < 1e-9
    assert all(abs(lad[k + 1] / lad[k] - lad[1] / lad[0]) < 1e-12 for k in range(31))
    f = Force(temperature=300.0)
    e, b = f.add_multithermal(300.0, 500.0, 4)
    t, c, p, ip = f.biases[b]
    assert t == 15 and c == [e] and len(p) == 5 and abs(p[0] - 8.314462618e-3 * 300) < 1e-15
    assert f.cvs[e][0] == 23
