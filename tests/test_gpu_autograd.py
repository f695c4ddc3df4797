"""The differentiable path-CV -> MetaD pipeline (torch.autograd.Function with
hand-written adjoints) against FP64 central finite differences of the oracle
(h = 1e-5) and the oracle's closed forms."""
import math

import numpy as np
import pytest
import torch

from gluedmd_amd import workloads as wl

pytestmark = pytest.mark.gpu


def oracle_V(frames, lam, x, hc, hh, sig, exact=True):
    """V(s(x), z(x)) of one configuration in FP64 (oracle path CV + hill sum)."""
    from oracle import oracle as o
    from gluedmd_amd.glued import Force
    f = Force()
    M = x.shape[0]
    f.add_path(list(range(M)), [fr.reshape(-1).tolist() for fr in frames], lam)
    c = o.OracleContext(f, M)
    sz = c.eval_cvs_f64(x, exact_params=exact)[0]
    V, _ = o.hills_eval(sz, hc, hh, sig, np.zeros(2))
    return V, sz


@pytest.fixture(scope="module")
def setup():
    from gluedmd_amd.autograd import PathMetadPipeline
    c = wl.config3_system(batch=8, M=60, N=16, n_hills=1500)
    from oracle import oracle as o
    from gluedmd_amd.glued import Force
    f0 = Force()
    f0.add_path(list(range(c["M"])), [fr.reshape(-1).tolist() for fr in c["frames"]], c["lam"])
    sz = np.array([o.OracleContext(f0, c["M"]).eval_cvs(p)[0] for p in c["confs"]])
    hc, hh, sig = wl.config3_hills(c["rng"], (sz[:, 0].min(), sz[:, 0].max()), (sz[:, 1].min(), sz[:, 1].max()),
                                   n_hills=1500)
    pipe = PathMetadPipeline(c["frames"], c["lam"], hc, hh, sig, batch=len(c["confs"]))
    return c, pipe, hc, hh, sig


def test_forward_matches_oracle(setup):
    c, pipe, hc, hh, sig = setup
    x = torch.as_tensor(c["confs"], dtype=torch.float64, device=pipe.device)
    h = torch.as_tensor(hh, device=pipe.device)
    V = pipe(x, h, torch.tensor(c["lam"], dtype=torch.float64)).cpu().numpy()
    for b in range(len(c["confs"])):
        Vo, _ = oracle_V(c["frames"], c["lam"], c["confs"][b].astype(np.float64), hc, hh, sig)
        assert abs(V[b] - Vo) <= 1e-6 * abs(Vo) + 1e-10


def test_backward_against_fd(setup):
    from oracle import oracle as o
    c, pipe, hc, hh, sig = setup
    B = len(c["confs"])
    rng = np.random.default_rng(0)
    gvec = rng.normal(size=B)
    x = torch.as_tensor(c["confs"], dtype=torch.float64, device=pipe.device).requires_grad_(True)
    h = torch.as_tensor(hh, device=pipe.device).requires_grad_(True)
    lam = torch.tensor(c["lam"], dtype=torch.float64, requires_grad=True)
    sg = torch.as_tensor(sig, dtype=torch.float64, device=pipe.device).requires_grad_(True)
    V = pipe(x, h, lam, sg)
    (V * torch.as_tensor(gvec, device=pipe.device)).sum().backward()
    gx = x.grad.cpu().numpy()
    step = 1e-5
    # positions: FD of V_b in FP64, skipping coordinates whose +-h moves a hill across E = 8
    checked = 0
    for b in range(2):
        xb = c["confs"][b].astype(np.float64)
        for (j, k) in [(0, 0), (7, 1), (33, 2), (59, 0)]:
            xp, xm = xb.copy(), xb.copy()
            xp[j, k] += step
            xm[j, k] -= step
            Vp, szp = oracle_V(c["frames"], c["lam"], xp, hc, hh, sig)
            Vm, szm = oracle_V(c["frames"], c["lam"], xm, hc, hh, sig)
            if min(o.hill_margin(szp, hc, sig, np.zeros(2)), o.hill_margin(szm, hc, sig, np.zeros(2))) < 1e-6:
                continue
            fd = gvec[b] * (Vp - Vm) / (2 * step)
            assert abs(gx[b, j, k] - fd) <= 1e-5 * max(abs(fd), 1e-3), (b, j, k, gx[b, j, k], fd)
            checked += 1
    assert checked >= 4
    # heights: V is linear in h -> dL/dh_k = sum_b g_b exp(-E_bk)[E_bk <= 8]
    gh = h.grad.cpu().numpy()
    want = np.zeros(len(hh))
    for b in range(B):
        _, sz = oracle_V(c["frames"], c["lam"], c["confs"][b].astype(np.float32).astype(np.float64), hc, hh, sig)
        dc = sz[None, :] - hc
        E = np.sum(dc * dc / (2 * sig ** 2), axis=1)
        want += gvec[b] * np.where(E <= 8.0, np.exp(-E), 0.0)
    assert np.linalg.norm(gh - want) <= 1e-6 * np.linalg.norm(want)
    # lambda: FD in FP64 of sum_b g_b V_b
    def L_of(lv):
        return sum(gvec[b] * oracle_V(c["frames"], lv, c["confs"][b].astype(np.float32).astype(np.float64), hc, hh,
                                      sig)[0] for b in range(B))
    fd_l = (L_of(c["lam"] + step) - L_of(c["lam"] - step)) / (2 * step)
    assert abs(lam.grad.item() - fd_l) <= 1e-5 * max(abs(fd_l), 1e-6), (lam.grad.item(), fd_l)
    # widths: FD of the hill sum in sigma.  The truncated Gaussian is not
    # differentiable where a hill sits on the E = 8 edge, so the step is sized
    # from the smallest distance of any hill's E to the edge (a relative step
    # ds/sigma moves E_d by at most 2 E_d ds/sigma <= 16 ds/sigma).
    gs = sg.grad.cpu().numpy()
    szs = [oracle_V(c["frames"], c["lam"], c["confs"][b].astype(np.float32).astype(np.float64), hc, hh, sig)[1]
           for b in range(B)]
    margin = min(o.hill_margin(sz, hc, sig, np.zeros(2)) for sz in szs)
    rel = min(1e-6, margin / 64.0)
    assert rel > 1e-9, margin
    for d in range(2):
        ds = rel * sig[d]
        sp, sm = sig.copy(), sig.copy()
        sp[d] += ds
        sm[d] -= ds
        tot_p = tot_m = 0.0
        for b in range(B):
            tot_p += gvec[b] * o.hills_eval(szs[b], hc, hh, sp, np.zeros(2))[0]
            tot_m += gvec[b] * o.hills_eval(szs[b], hc, hh, sm, np.zeros(2))[0]
        fd_s = (tot_p - tot_m) / (2 * ds)
        assert abs(gs[d] - fd_s) <= 1e-5 * max(abs(fd_s), 1e-3), (d, gs[d], fd_s)


def test_backward_recomputes_after_other_forward(setup):
    c, pipe, hc, hh, sig = setup
    x = torch.as_tensor(c["confs"], dtype=torch.float64, device=pipe.device).requires_grad_(True)
    h = torch.as_tensor(hh, device=pipe.device)
    V1 = pipe(x, h, torch.tensor(c["lam"], dtype=torch.float64))
    V1.sum().backward()
    g1 = x.grad.clone()
    x.grad = None
    V2 = pipe(x, h, torch.tensor(c["lam"], dtype=torch.float64))
    _ = pipe(x * 1.001, h, torch.tensor(c["lam"], dtype=torch.float64))  # moves the engine on
    V2.sum().backward()
    assert torch.equal(x.grad, g1)


def test_backward_frames(setup):
    """dL/d(reference frames) from the frame adjoint against FP64 central
    finite differences of the oracle (path CV + hill sum), and the
    translation identity: V depends on x - r_f only, so per component the
    sum over frames and atoms of dL/dr equals minus the sum of dL/dx."""
    from oracle import oracle as o
    c, pipe, hc, hh, sig = setup
    B = len(c["confs"])
    rng = np.random.default_rng(5)
    gvec = rng.normal(size=B)
    frames0 = np.asarray(c["frames"], dtype=np.float64)
    x = torch.as_tensor(c["confs"], dtype=torch.float64, device=pipe.device).requires_grad_(True)
    h = torch.as_tensor(hh, device=pipe.device)
    fr = torch.as_tensor(frames0, device=pipe.device).requires_grad_(True)
    V = pipe(x, h, torch.tensor(c["lam"], dtype=torch.float64), frames=fr)
    (V * torch.as_tensor(gvec, device=pipe.device)).sum().backward()
    gF = fr.grad.cpu().numpy()
    gX = x.grad.cpu().numpy()
    assert gF.shape == frames0.shape
    lhs = gF.sum(axis=(0, 1))
    rhs = -gX.sum(axis=(0, 1))
    scale = np.abs(gX).sum(axis=(0, 1))
    assert np.all(np.abs(lhs - rhs) <= 1e-6 * scale + 1e-8), (lhs, rhs)
    # FD in FP64 on the oracle, around the frames as the engine stores them (FP32 values)
    fr32 = frames0.astype(np.float32).astype(np.float64)
    xs = [c["confs"][b].astype(np.float32).astype(np.float64) for b in range(B)]
    step = 1e-5
    checked = 0
    for (f, j, k) in [(0, 0, 0), (3, 11, 1), (9, 40, 2), (15, 59, 0), (7, 30, 1)]:
        fp, fm = fr32.copy(), fr32.copy()
        fp[f, j, k] += step
        fm[f, j, k] -= step
        Lp = Lm = 0.0
        ok = True
        for b in range(B):
            Vp, szp = oracle_V(fp, c["lam"], xs[b], hc, hh, sig)
            Vm, szm = oracle_V(fm, c["lam"], xs[b], hc, hh, sig)
            ok &= min(o.hill_margin(szp, hc, sig, np.zeros(2)), o.hill_margin(szm, hc, sig, np.zeros(2))) >= 1e-6
            Lp += gvec[b] * Vp
            Lm += gvec[b] * Vm
        if not ok:  # a hill crosses the E = 8 edge within the step
            continue
        fd = (Lp - Lm) / (2 * step)
        assert abs(gF[f, j, k] - fd) <= 1e-5 * max(abs(fd), 1e-3), (f, j, k, gF[f, j, k], fd)
        checked += 1
    assert checked >= 3
    # the frames are restored for the other tests
    pipe(x.detach(), h, torch.tensor(c["lam"], dtype=torch.float64), frames=torch.as_tensor(frames0, device=pipe.device))
