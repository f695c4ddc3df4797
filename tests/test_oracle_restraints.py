"""Pins of the oracle's static restraints (CPU): harmonic, linear and the
one-sided walls, against the known answers and closed forms of the
reference's own tests (tests/test_bias_harmonic.py, test_bias_linear.py,
test_bias_walls.py in MarvinTaterra/GluedMD, cited per test), central
finite differences and the FP32 parameter storage of the reference."""
import math

import numpy as np

from gluedmd_amd.glued import Force

TOL_E, TOL_F = 1e-5, 1e-4


def _run(oracle, f, pos):
    c = oracle.OracleContext(f, len(pos))
    return c.execute(np.asarray(pos, dtype=np.float32))


def _harmonic_distance(k, s0, dist):
    f = Force()
    d = f.add_distance([0, 1])
    f.add_harmonic([d], k, s0)
    return f, [(0.0, 0.0, 0.0), (dist, 0.0, 0.0), (5.0, 5.0, 5.0)]


def test_harmonic_known_answers(oracle):
    # test_bias_harmonic.py:72-82 (E = 12.5), :86-94 (E = 0 at the minimum),
    # :98-114 (F1.x = -50, Newton's third law), :117-130 (s < s0: F1.x = +100)
    for k, s0, dist, E, F1 in ((100.0, 0.5, 1.0, 12.5, -50.0), (200.0, 1.0, 1.0, 0.0, 0.0),
                               (100.0, 2.0, 1.0, 50.0, 100.0)):
        f, pos = _harmonic_distance(k, s0, dist)
        r = _run(oracle, f, pos)
        assert abs(r["energy"] - E) < TOL_E
        assert abs(r["forces"][1][0] - F1) < TOL_F
        assert abs(r["forces"][0][0] + r["forces"][1][0]) < TOL_F


def test_harmonic_two_cvs(oracle):
    # test_bias_harmonic.py:134-150: one bias over two distances, E = V1 + V2 = 12.5 + 6.25
    f = Force()
    a = f.add_distance([0, 1])
    b = f.add_distance([0, 2])
    f.add_harmonic([a, b], [100.0, 50.0], [0.5, 1.5])
    r = _run(oracle, f, [(0, 0, 0), (1.0, 0, 0), (0, 2.0, 0), (5, 5, 5)])
    assert abs(r["energy"] - 18.75) < 10 * TOL_E
    assert abs(r["dUdCV"][a] - 50.0) < TOL_F and abs(r["dUdCV"][b] - 25.0) < TOL_F


def test_harmonic_fd(oracle):
    # test_bias_harmonic.py:154-182: F_x = -(V(+dx) - V(-dx)) / 2dx, k = 200, s0 = 1, s = 1.5
    f, _ = _harmonic_distance(200.0, 1.0, 1.5)
    dx = 1e-2
    Ep = _run(oracle, f, [(0, 0, 0), (1.5 + dx, 0, 0), (5, 5, 5)])["energy"]
    Em = _run(oracle, f, [(0, 0, 0), (1.5 - dx, 0, 0), (5, 5, 5)])["energy"]
    r = _run(oracle, f, [(0, 0, 0), (1.5, 0, 0), (5, 5, 5)])
    assert abs(r["forces"][1][0] - (-(Ep - Em) / (2 * dx))) < 10 * TOL_F
    assert abs(r["forces"][1][0] + 100.0) < TOL_F


def test_harmonic_periodic_folds_the_seam(oracle):
    # the periodic flag (CommonGluedKernels.cpp:2530-2533): a dihedral at
    # +170 degrees restrained to -170 degrees is 20 degrees away, not 340
    f = Force()
    t = f.add_dihedral([0, 1, 2, 3])
    phi = math.radians(170.0)
    f.add_harmonic([t], 10.0, math.radians(-170.0), periodic=True)
    pos = [(1.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (math.cos(phi), math.sin(phi), 1.0)]
    r = _run(oracle, f, pos)
    s = r["cv"][t]
    ds = s - np.float32(math.radians(-170.0))
    ds -= 2 * math.pi * round(ds / (2 * math.pi))
    assert abs(abs(ds) - math.radians(20.0)) < 1e-4
    assert abs(r["energy"] - 0.5 * 10.0 * ds * ds) < 1e-9
    g = Force()
    t2 = g.add_dihedral([0, 1, 2, 3])
    g.add_harmonic([t2], 10.0, math.radians(-170.0), periodic=False)
    assert _run(oracle, g, pos)["energy"] > 100 * r["energy"]


def test_harmonic_parameters_stored_as_fp32(oracle):
    # the reference uploads k and s0 as float (CommonBiasKernels.cpp:118-126)
    f, pos = _harmonic_distance(0.1, 0.3, 1.0)
    r = _run(oracle, f, pos)
    k, s0 = float(np.float32(0.1)), float(np.float32(0.3))
    assert r["energy"] == 0.5 * k * (r["cv"][0] - s0) * (r["cv"][0] - s0)


def _position_x(x):
    return [(x, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)]


def _wall_energy(s, at, kappa, eps, n, lower):
    # test_bias_walls.py:19-28
    delta = (at - s) if lower else (s - at)
    return 0.0 if delta <= 0.0 else kappa * delta ** n * math.exp(eps * delta)


def _wall_grad(s, at, kappa, eps, n, lower):
    # test_bias_walls.py:30-39
    delta = (at - s) if lower else (s - at)
    if delta <= 0.0:
        return 0.0
    return (-1.0 if lower else 1.0) * kappa * (n * delta ** (n - 1.0) + eps * delta ** n) * math.exp(eps * delta)


def test_walls_known_answers(oracle):
    # test_bias_walls.py:71-161: upper / lower, outside (E = 0) and inside,
    # the force sign, and eps = 2 (E = 50 * 0.09 * e^0.6); plus a
    # non-integer exponent (the pow() branch)
    cases = [(0.2, 0.5, 100.0, 0.0, 2.0, False), (0.8, 0.5, 100.0, 0.0, 2.0, False),
             (0.8, 0.5, 100.0, 0.0, 2.0, True), (0.2, 0.5, 100.0, 0.0, 2.0, True),
             (0.8, 0.5, 50.0, 2.0, 2.0, False), (0.9, 0.5, 30.0, 0.5, 3.5, False),
             (0.1, 0.5, 30.0, 0.5, 1.0, True)]
    for s, at, kappa, eps, n, lower in cases:
        f = Force()
        c = f.add_position(0, 0)
        (f.add_lower_wall if lower else f.add_upper_wall)(c, at, kappa, n=n, eps=eps)
        r = _run(oracle, f, _position_x(s))
        sv = float(np.float32(s))
        p = [float(np.float32(x)) for x in (at, kappa, eps, n)]
        assert abs(r["energy"] - _wall_energy(sv, *p, lower)) < TOL_E * max(1.0, r["energy"])
        fx = r["forces"][0][0]
        assert abs(fx + _wall_grad(sv, *p, lower)) < TOL_F * max(1.0, abs(fx))
        if r["energy"] > 0:
            assert (fx > 0) == lower  # an upper wall pushes down, a lower wall up
    assert abs(_wall_energy(0.8, 0.5, 50.0, 2.0, 2, False) - 50.0 * 0.09 * math.exp(0.6)) < 1e-12


def test_linear_known_answers(oracle):
    # test_bias_linear.py:49-155: E = k s, F = -k along the CV's direction,
    # two CVs, E = 0 at the origin
    f = Force()
    x = f.add_position(0, 0)
    z = f.add_position(0, 2)
    f.add_linear([x, z], [3.0, -2.0])
    r = _run(oracle, f, [(0.7, 0.1, 0.4), (0, 0, 0)])
    assert abs(r["energy"] - (3.0 * 0.7 + (-2.0) * 0.4)) < TOL_E
    assert abs(r["forces"][0][0] + 3.0) < TOL_F and abs(r["forces"][0][2] - 2.0) < TOL_F
    assert abs(r["forces"][0][1]) < TOL_F
    assert _run(oracle, f, [(0, 0, 0), (0, 0, 0)])["energy"] == 0.0
