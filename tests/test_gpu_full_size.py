"""GPU parity of the benchmark workloads at their stated sizes, against the
FP64 oracle:
  - the large mixed workload: 2 walkers at the full per-walker shape
    (100,000 atoms, 4,096-pair contact map, 1,500-pair DRMSD, 256-frame x
    1,000-atom path, multithermal over 32 temperatures), 3 deposit steps;
  - the multi-walker workload: 64 walkers of 2,000-40,000 atoms with 20,000
    private hills each and the shared 200 x 200 grid, evaluation then deposits;
  - the differentiable path-CV pipeline: the backward of 256 configurations
    of 300 atoms, 64 frames, 20,000 hills, checked by FP64 central finite
    differences (h = 1e-5) on samples of coordinates, hill heights, lambda and
    reference frames."""
import numpy as np
import pytest
import torch

from gluedmd_amd import workloads as wl
from tests.gpu_util import check

pytestmark = pytest.mark.gpu


def test_mixed_large_full_shape_two_walkers():
    from gluedmd_amd.context import Context
    from oracle.oracle import OracleContext
    c = wl.config5(n_walkers=2)
    f = c["force"]
    W = 2
    assert c["natoms"] == 100000 and len(c["ladder"]) == 32
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
    t, atoms, params = f.cvs[2]
    assert t == f.CV_PATH and int(params[1]) == 256 and len(atoms) == 1000
    assert len(f.cvs[0][1]) == 2 * 4096 and len(f.cvs[1][1]) == 2 * 1500
    for step in range(3):
        g.update_state(step)
        for o in os_:
            o.update_state(step)
        g.zero_forces()
        g.execute_batch(use_graph=True)
        for w in range(W):
            r = os_[w].execute(c["positions"][w])
            check(g, os_[w], r, g.getLastBias(w), w)


def test_multiwalker_full_64x20k():
    from tests.test_gpu_bias import _config4_pair
    c, g, os_ = _config4_pair(64, 20000)
    W = len(os_)
    assert W == 64 and min(c["sizes"]) >= 2000 and max(c["sizes"]) <= 40000
    rng = np.random.default_rng(8)
    for step in range(3):
        g.update_state(step)
        for o in os_:
            o.update_state(step)  # the shared grid takes the walkers in walker order
        pos = [(p + rng.normal(0, 0.01, size=p.shape)).astype(np.float32) for p in c["positions"]]
        g.set_all_positions(pos)
        g.zero_forces()
        g.execute_batch(use_graph=True)
        for w in range(W):
            r = os_[w].execute(pos[w])
            check(g, os_[w], r, g.getLastBias(w), w)
    for w in range(W):
        assert len(g.get_hills(c["hill_bias"], w)[1]) == len(os_[w].states[c["hill_bias"]].heights) == 20002


def _path_eval(frames, lam, M):
    """An oracle evaluator of the path CV (s, z) with the given frames, in FP64."""
    from oracle import oracle as o
    from gluedmd_amd.glued import Force
    f = Force()
    f.add_path(list(range(M)), [fr.reshape(-1).tolist() for fr in frames], lam)
    ctx = o.OracleContext(f, M)
    return lambda x: ctx.eval_cvs_f64(x, exact_params=True)[0]


def test_path_pipeline_full_backward():
    from gluedmd_amd.autograd import PathMetadPipeline
    from oracle import oracle as o
    c = wl.config3_system(batch=256)
    B, M = len(c["confs"]), c["M"]
    assert (B, M, c["N"]) == (256, 300, 64)
    xs = [c["confs"][b].astype(np.float64) for b in range(B)]
    sz0 = np.array([_path_eval(c["frames"], c["lam"], M)(x) for x in xs])
    hc, hh, sig = wl.config3_hills(c["rng"], (sz0[:, 0].min(), sz0[:, 0].max()), (sz0[:, 1].min(), sz0[:, 1].max()),
                                   n_hills=20000)
    pipe = PathMetadPipeline(c["frames"], c["lam"], hc, hh, sig, batch=B)
    rng = np.random.default_rng(9)
    gvec = rng.normal(size=B)
    frames0 = np.asarray(c["frames"], dtype=np.float64)
    x = torch.as_tensor(c["confs"], dtype=torch.float64, device=pipe.device).requires_grad_(True)
    h = torch.as_tensor(hh, device=pipe.device).requires_grad_(True)
    lam = torch.tensor(c["lam"], dtype=torch.float64, requires_grad=True)
    fr = torch.as_tensor(frames0, device=pipe.device).requires_grad_(True)
    V = pipe(x, h, lam, frames=fr)
    Vn = V.detach().cpu().numpy()
    zero = np.zeros(2)
    for b in range(0, B, 37):  # forward spot check
        Vo = o.hills_eval(sz0[b], hc, hh, sig, zero)[0]
        assert abs(Vn[b] - Vo) <= 1e-6 * abs(Vo) + 1e-10, (b, Vn[b], Vo)
    (V * torch.as_tensor(gvec, device=pipe.device)).sum().backward()
    gx, gh, gF = x.grad.cpu().numpy(), h.grad.cpu().numpy(), fr.grad.cpu().numpy()
    step = 1e-5
    fr32 = frames0.astype(np.float32).astype(np.float64)

    def included(sz):
        dc = sz[None, :] - hc
        return np.sum(dc * dc / (2.0 * sig * sig), axis=1) <= 8.0

    def V_frozen(sz, mask, heights=hh):
        """The hill sum over the hills with E <= 8 at the unperturbed point:
        the truncated sum jumps by h e^-8 where a hill crosses the edge, so a
        central difference is taken of this smooth branch, the one whose
        derivative the adjoint returns."""
        dc = sz[None, :] - hc
        E = np.sum(dc * dc / (2.0 * sig * sig), axis=1)
        return float(np.sum(np.where(mask, heights * np.exp(-E), 0.0)))

    masks = [included(sz0[b]) for b in range(B)]
    for b in range(0, B, 51):  # the frozen branch is the oracle's sum at the point itself
        assert abs(V_frozen(sz0[b], masks[b]) - o.hills_eval(sz0[b], hc, hh, sig, zero)[0]) <= 1e-12 * abs(Vn[b])

    # coordinates: FD of g_b V_b(x_b)
    ev = _path_eval(fr32, c["lam"], M)
    for (b, j, k) in [(0, 0, 0), (17, 123, 1), (64, 299, 2), (130, 7, 0), (201, 250, 1), (255, 42, 2)]:
        xp, xm = xs[b].copy(), xs[b].copy()
        xp[j, k] += step
        xm[j, k] -= step
        fd = gvec[b] * (V_frozen(ev(xp), masks[b]) - V_frozen(ev(xm), masks[b])) / (2 * step)
        assert abs(gx[b, j, k] - fd) <= 1e-5 * max(abs(fd), 1e-3), (b, j, k, gx[b, j, k], fd)
    # heights: FD of L = sum_b g_b V_b over the full batch
    for k in (0, 4321, 12345, 19999):
        hp, hm = hh.copy(), hh.copy()
        hp[k] += step
        hm[k] -= step
        Lp = sum(gvec[b] * o.hills_eval(sz0[b], hc, hp, sig, zero)[0] for b in range(B))
        Lm = sum(gvec[b] * o.hills_eval(sz0[b], hc, hm, sig, zero)[0] for b in range(B))
        fd = (Lp - Lm) / (2 * step)
        assert abs(gh[k] - fd) <= 1e-6 * max(np.abs(gh).max(), 1e-12), (k, gh[k], fd)
    # lambda
    Ls = []
    for lv in (c["lam"] + step, c["lam"] - step):
        e = _path_eval(fr32, lv, M)
        Ls.append(sum(gvec[b] * V_frozen(e(xs[b]), masks[b]) for b in range(B)))
    fd_l = (Ls[0] - Ls[1]) / (2 * step)
    assert abs(lam.grad.item() - fd_l) <= 1e-5 * max(abs(fd_l), 1e-6), (lam.grad.item(), fd_l)
    # reference frames
    for (f_, j, k) in [(0, 0, 0), (21, 150, 1), (63, 299, 2)]:
        fp, fm = fr32.copy(), fr32.copy()
        fp[f_, j, k] += step
        fm[f_, j, k] -= step
        ep, em = _path_eval(fp, c["lam"], M), _path_eval(fm, c["lam"], M)
        Lp = sum(gvec[b] * V_frozen(ep(xs[b]), masks[b]) for b in range(B))
        Lm = sum(gvec[b] * V_frozen(em(xs[b]), masks[b]) for b in range(B))
        fd = (Lp - Lm) / (2 * step)
        assert abs(gF[f_, j, k] - fd) <= 1e-5 * max(abs(fd), 1e-3), (f_, j, k, gF[f_, j, k], fd)
