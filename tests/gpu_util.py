"""Shared helpers of the GPU parity tests: run one force on the HIP engine and
on the FP64 oracle with the same FP32 inputs and compare.

Tolerances (the acceptance bar of this path):
  CV values and bias energies: |gpu - oracle| <= 1e-6 |oracle| (+1e-12 absolute)
  forces: ||F_gpu - F_64||_2 / ||F_64||_2 <= 1e-5 over all atoms
  neighbour counts, hill and kernel counts: exact
"""
import numpy as np

CV_RTOL = 1e-6
E_RTOL = 1e-6
F_RMS = 1e-5


def rel_rms(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a - b))


def assert_close_rel(got, want, rtol, what, atol=1e-12):
    got = np.atleast_1d(np.asarray(got, dtype=np.float64))
    want = np.atleast_1d(np.asarray(want, dtype=np.float64))
    err = np.abs(got - want)
    bad = err > rtol * np.abs(want) + atol
    assert not bad.any(), f"{what}: got {got[bad]} want {want[bad]} (rtol {rtol})"


def make_pair(force, natoms, box=None, masses=None):
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    g = Context(force, natoms, masses=masses)
    o = OracleContext(force, natoms, masses=masses)
    if box is not None:
        g.set_box(box)
        o.set_box(box)
    return g, o


def run_both(g, o, pos, walker=0, compare=True, f_rms=F_RMS):
    pos = np.asarray(pos, dtype=np.float32)
    g.set_positions(pos, walker)
    g.zero_forces()
    eg = g.execute(walker)
    r = o.execute(pos)
    if compare:
        check(g, o, r, eg, walker, f_rms)
    return eg, r


def check(g, o, r, eg, walker=0, f_rms=F_RMS):
    cv = g.getLastCVValues(walker)
    assert_close_rel(cv, r["cv"], CV_RTOL, "CV values")
    if o.force.getNumBiases() and o.force._test_grads is None:
        # end to end (oracle CVs from the same positions) ...
        assert_close_rel(g.getBiasEnergies(walker), r["bias_energies"], E_RTOL, "bias energies", atol=1e-10)
        assert_close_rel(eg, r["energy"], E_RTOL, "total bias energy", atol=1e-10)
        # ... and the bias arithmetic alone, at the GPU's own CV values
        be, dU = o.eval_biases(cv, welford=False)
        assert_close_rel(g.getBiasEnergies(walker), be, E_RTOL, "bias energies at GPU CVs", atol=1e-10)
        scale = max(np.abs(dU).max(), 1e-300)
        assert np.abs(g.getDUDCV(walker) - dU).max() <= 1e-6 * scale, "dU/dCV at GPU CVs"
    F = g.get_forces(walker)
    if np.linalg.norm(r["forces"]) > 0:
        e = rel_rms(F, r["forces"])
        assert e <= f_rms, f"force rel RMS {e:.3e} > {f_rms}"
    else:
        assert np.abs(F).max() < 1e-6
