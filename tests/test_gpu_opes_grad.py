"""Gradients of an OPES bias with respect to its kernel table (OpesPipeline),
at the 50,000-atom workload (coordination 2,000 x 20,000 within 1.2 nm and Rg
over 5,000 atoms, OPES-METAD with a 5,000-kernel table shared by 3
configurations): the forward against the oracle, the adjoint for centres,
widths, log-weights and sum_uprob against FP64 central differences of the
oracle's OPES evaluation (orc_opes_eval) at the GPU's CV values, and the
coordinate gradient against the oracle's forces."""
import numpy as np
import pytest
import torch

from gluedmd_amd import workloads as wl
from tests.gpu_util import rel_rms

pytestmark = pytest.mark.gpu


def test_opes_table_gradients_50k_atoms_5000_kernels():
    from gluedmd_amd.autograd import OpesPipeline
    from oracle.oracle import OracleContext, _Opes
    c = wl.config2_system()
    f, N, B = c["force"], c["natoms"], 3
    o = OracleContext(f, N)
    o.set_box(c["box"])
    rng = np.random.default_rng(11)
    xs = [c["positions"]] + [(c["positions"] + rng.normal(0, 0.002, size=(N, 3))).astype(np.float32)
                             for _ in range(B - 1)]
    r0 = o.eval_cvs(xs[0])[0]
    centers, sigmas, logw = wl.opes_prior_kernels(c["rng"], r0, [0.05, 0.02], c["gamma"], n=5000)
    su = wl.sum_uprob_np(centers, sigmas, logw, o.states[0].cutoff2)
    _, cvs, params, ip = f.biases[0]
    D = len(cvs)

    pipe = OpesPipeline(f, 0, N, B, box=c["box"])
    dev = pipe.device
    x = torch.as_tensor(np.stack(xs), device=dev).requires_grad_(True)
    C = torch.as_tensor(centers, device=dev).requires_grad_(True)
    S = torch.as_tensor(sigmas, device=dev).requires_grad_(True)
    L = torch.as_tensor(logw, device=dev).requires_grad_(True)
    U = torch.tensor(su, dtype=torch.float64, requires_grad=True)
    V = pipe(x, C, S, L, U)
    Vn = V.detach().cpu().numpy()
    s_gpu = [pipe.ctx.getLastCVValues(b)[:D].copy() for b in range(B)]

    def table_loss(g, c_, s_, l_, u_):
        ob = _Opes(D, params, ip)
        ob.load(c_, s_, l_, u_)
        return sum(g[b] * ob.eval(s_gpu[b])[0] for b in range(B))

    # forward: end to end against the oracle's CVs, and at the GPU's own CVs
    o.states[0].load(centers, sigmas, logw, su)
    for b in range(B):
        r = o.execute(xs[b])
        assert abs(Vn[b] - r["energy"]) <= 1e-6 * abs(r["energy"]) + 1e-10, (b, Vn[b], r["energy"])
        e = np.eye(B)[b]
        Vg = table_loss(e, centers, sigmas, logw, su)
        assert abs(Vn[b] - Vg) <= 1e-10 * abs(Vg) + 1e-12, (b, Vn[b], Vg)
    gvec = rng.normal(size=B)
    (V * torch.as_tensor(gvec, device=dev)).sum().backward()
    gx, gc, gs, gl = (t.grad.cpu().numpy() for t in (x, C, S, L))
    gu = U.grad.item()

    # coordinates: g_b times minus the oracle's forces on configuration b
    for b in range(B):
        r = o.execute(xs[b])
        err = rel_rms(gx[b] / gvec[b], -r["forces"])
        assert err <= 1e-5, (b, err)

    # kernel parameters: central differences of the oracle's table evaluation
    active = np.flatnonzero(gl != 0.0)
    assert 0 < len(active) < 5000, len(active)
    picks = list(active[np.argsort(-np.abs(gl[active]))[:4]]) + list(rng.choice(active, 3, replace=False))
    inactive = np.flatnonzero(gl == 0.0)
    picks += list(rng.choice(inactive, 2, replace=False))
    h = 1e-6
    scale = {"c": np.abs(gc).max(), "s": np.abs(gs).max(), "l": np.abs(gl).max()}
    for k in picks:
        for name, arr, grad in (("c", centers, gc), ("s", sigmas, gs), ("l", logw, gl)):
            for d in (range(D) if name != "l" else [None]):
                p, m = arr.copy(), arr.copy()
                if d is None:
                    p[k] += h
                    m[k] -= h
                    got = grad[k]
                else:
                    p[k, d] += h
                    m[k, d] -= h
                    got = grad[k, d]
                args_p = {"c": (p, sigmas, logw), "s": (centers, p, logw), "l": (centers, sigmas, p)}[name]
                args_m = {"c": (m, sigmas, logw), "s": (centers, m, logw), "l": (centers, sigmas, m)}[name]
                fd = (table_loss(gvec, *args_p, su) - table_loss(gvec, *args_m, su)) / (2 * h)
                assert abs(got - fd) <= 1e-6 * scale[name] + 1e-6 * abs(fd), (name, k, d, got, fd)
                if k in inactive:
                    assert got == 0.0
    hu = 1e-6 * su
    fd = (table_loss(gvec, centers, sigmas, logw, su + hu) - table_loss(gvec, centers, sigmas, logw, su - hu)) / (2 * hu)
    assert abs(gu - fd) <= 1e-6 * abs(fd) + 1e-12, (gu, fd)
