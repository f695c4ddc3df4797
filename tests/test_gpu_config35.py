"""GPU parity on the path-CV pipeline workload (256 configurations batched as
walkers, path over 64 frames, MetaD hills shared by all configurations) and
the large mixed workload (contact map, DRMSD, 256-frame path, multithermal on
a dense energy CV), against the FP64 oracle."""
import numpy as np
import pytest

from gluedmd_amd import workloads as wl
from gluedmd_amd.glued import Force
from tests.gpu_util import assert_close_rel, check, rel_rms

pytestmark = pytest.mark.gpu


def config3_force(c, hills_c, hills_h, sig):
    f = Force(temperature=300.0)
    s = f.add_path(list(range(c["M"])), [fr.reshape(-1).tolist() for fr in c["frames"]], c["lam"])
    hb = f.add_metad_hills([s, s + 1], sigma=sig.tolist(), height=1.2, pace=1, gamma=None,
                           max_hills=len(hills_h) + 16)
    f.share_bias(hb)
    return f, hb


def test_config3_batched_path_metad():
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    c = wl.config3_system(batch=32)
    f0 = Force()
    f0.add_path(list(range(c["M"])), [fr.reshape(-1).tolist() for fr in c["frames"]], c["lam"])
    sz = np.array([OracleContext(f0, c["M"]).eval_cvs(p)[0] for p in c["confs"]])
    hc, hh, sig = wl.config3_hills(c["rng"], (sz[:, 0].min(), sz[:, 0].max()), (sz[:, 1].min(), sz[:, 1].max()),
                                   n_hills=4000)
    f, hb = config3_force(c, hc, hh, sig)
    B = len(c["confs"])
    g = Context(f, walkers=[c["M"]] * B)
    g.set_hills(hb, hc, hh)
    for b in range(B):
        g.set_positions(c["confs"][b], b)
    g.zero_forces()
    g.execute_batch(use_graph=True)
    for b in range(B):
        o = OracleContext(f, c["M"])
        o.states[hb].centers, o.states[hb].heights = hc, hh
        r = o.execute(c["confs"][b])
        check(g, o, r, g.getLastBias(b), b)


def test_config5_reduced():
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    import torch
    c = wl.config5(n_walkers=2, N=20000, n_cm=1024, n_drmsd=400, path_frames=64, path_atoms=300, n_temps=32)
    f = c["force"]
    W = 2
    g = Context(f, walkers=[c["natoms"]] * W)
    g.set_box(c["box"])
    os_ = []
    for w in range(W):
        g.set_positions(c["positions"][w], w)
        p = g.padded[w]
        Fext = torch.zeros(3 * p, dtype=torch.float32, device=g.device)
        Fext.view(3, p)[:, :c["natoms"]] = torch.as_tensor(c["F"][w].T.copy(), device=g.device)
        g.set_energy_input(float(c["U"][w]), Fext, w)
        o = OracleContext(f, c["natoms"])
        o.set_box(c["box"])
        o.set_energy_input(float(c["U"][w]), c["F"][w].astype(np.float64))
        os_.append(o)
    for step in range(3):
        g.update_state(step)
        for o in os_:
            o.update_state(step)
        g.zero_forces()
        g.execute_batch(use_graph=True)
        for w in range(W):
            r = os_[w].execute(c["positions"][w])
            check(g, os_[w], r, g.getLastBias(w), w)


def test_shared_hill_chunks_match_plain():
    """A hill list shared by every configuration is evaluated by chunk
    workgroups that take four walkers each (the chunk loaded once).  With 30
    configurations (the last workgroup holds 2) the batched step must give
    the plain per-walker launches' bias energies and fixed-point forces bit
    for bit."""
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    c = wl.config3_system(batch=30, M=60, N=16)
    f0 = Force()
    f0.add_path(list(range(c["M"])), [fr.reshape(-1).tolist() for fr in c["frames"]], c["lam"])
    sz = np.array([OracleContext(f0, c["M"]).eval_cvs(p)[0] for p in c["confs"]])
    hc, hh, sig = wl.config3_hills(c["rng"], (sz[:, 0].min(), sz[:, 0].max()), (sz[:, 1].min(), sz[:, 1].max()),
                                   n_hills=3000)
    f, hb = config3_force(c, hc, hh, sig)
    B = len(c["confs"])
    g = Context(f, walkers=[c["M"]] * B)
    g.set_hills(hb, hc, hh)
    for b in range(B):
        g.set_positions(c["confs"][b], b)
    g.zero_forces()
    for b in range(B):
        g.execute(b)
    plain = [(g.getBiasEnergies(b).copy(), g.get_forces_fixed(b).copy()) for b in range(B)]
    g.zero_forces()
    g.execute_batch(use_graph=False)
    for b in range(B):
        assert np.array_equal(g.getBiasEnergies(b), plain[b][0]), b
        assert np.array_equal(g.get_forces_fixed(b), plain[b][1]), b
