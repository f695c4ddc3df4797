"""Differentiable CV -> bias pipeline on the HIP engine.

``PathMetadPipeline`` evaluates, for a batch of B configurations of M atoms,
the path CV (s, z) over N reference frames (MSD metric, CommonGluedKernels.cpp:
667-794) and a hill-list metadynamics bias V(s, z) with a hill set shared by
the batch.  ``PathMetadFunction`` is its ``torch.autograd.Function``:

  forward   one batched evaluation on the GPU (the configurations are the
            walkers of one engine, one captured hipGraph);
  backward  hand-written adjoints, no autograd tape:
            dL/dx       = g_b dV_b/dx_b, from the chain-rule force scatter
                          (the force buffer holds -dV/dx in 2^32 fixed point);
            dL/dh_k     = sum_b g_b exp(-E_bk)[E_bk <= 8]  (k_hill_adjoint);
            dL/dsigma_d = sum_b g_b sum_k h_k exp(-E_bk) dc_bkd^2 / sigma_d^3;
            dL/dlambda  = sum_b g_b (dV/ds ds/dlambda + dV/dz dz/dlambda), with
                          ds/dlambda = -(1/S) sum_f (f+1-s) msd_f w_f and
                          dz/dlambda = (<msd>_w - z)/lambda from the path kernel;
            dL/dr_fj    = sum_b c_bf (x_bj - r_fj) with
                          c_bf = g_b (2/M)(w_bf/S_b)[lambda Vs_b (f+1-s_b) - Vz_b]
                          (k_path_frames_coef / k_path_frames_grad).
If the engine was used for another evaluation between forward and backward,
backward recomputes the forward first (nothing per pair is stored).

``OpesPipeline`` evaluates one OPES bias (opesEvalBias,
CommonGluedKernels.cpp:2642-2723) for a batch of configurations sharing one
kernel table, and ``OpesFunction`` differentiates it with respect to the
coordinates and the table itself.  With A = kT/invGF, P_b = sum_k h_k
(G_bk - eps_c)[n2_bk < cutoff2], h_k = exp(logw_k), G_bk = exp(-n2_bk/2),
n2_bk = sum_d (s_bd - c_kd)^2/sigma_kd^2 and V_b = A log(P_b n/U + eps):
  dV_b/dlogw_k  = F_b h_k (G_bk - eps_c),  F_b = A n/(U (P_b n/U + eps));
  dV_b/dc_kd    = F_b h_k G_bk (s_bd - c_kd)/sigma_kd^2;
  dV_b/dsigma_kd = F_b h_k G_bk (s_bd - c_kd)^2/sigma_kd^3;
  dV_b/dU       = -A (P_b n/U)/(U (P_b n/U + eps)),
U = sum_uprob, held as the stored input it is in the evaluation
(k_opes_adjoint).
"""
import ctypes

import numpy as np
import torch

from ._lib import call
from .context import Context
from .glued import Force

FIXED_SCALE = 4294967296.0


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class PathMetadPipeline:
    def __init__(self, frames, lam, hill_centers, hill_heights, sigma, batch, *, temperature=300.0,
                 device="cuda:0"):
        frames = np.asarray(frames, dtype=np.float64)
        self.N, self.M = frames.shape[0], frames.shape[1]
        self.B = int(batch)
        self.H = len(hill_heights)
        f = Force(temperature=temperature)
        self.s_index = f.add_path(list(range(self.M)), [fr.reshape(-1).tolist() for fr in frames], float(lam))
        self.bias = f.add_metad_hills([self.s_index, self.s_index + 1], sigma=[float(x) for x in sigma], height=1.0,
                                      pace=1, gamma=None, max_hills=self.H)
        f.share_bias(self.bias)
        self.force = f
        self.ctx = Context(f, walkers=[self.M] * self.B, device=device)
        self.ctx.set_hills(self.bias, hill_centers, hill_heights)
        self.device = self.ctx.device
        self._lam = float(lam)
        self._sigma = [float(x) for x in sigma]
        self._frames = frames.astype(np.float32)  # as the engine stores them
        self.generation = 0
        self.grad_sigma = None

    def __call__(self, x, heights, lam, sigma=None, frames=None):
        """V_b for configurations x [B, M, 3]; frames [N, M, 3] (optional,
        differentiable) replaces the path's reference frames."""
        if sigma is None:
            sigma = torch.tensor(self._sigma, dtype=torch.float64, device=self.device)
        if frames is None:
            frames = torch.as_tensor(self._frames.astype(np.float64), device=self.device)
        return PathMetadFunction.apply(x, heights, lam, sigma, frames, self)

    # ---- engine plumbing ----
    def _evaluate(self, x, heights, lam, sigma, frames):
        e = self.ctx
        P = e.padded[0]
        posq = e.posq.view(self.B, P, 4)
        posq[:, :self.M, :3].copy_(x.detach().to(torch.float32))
        h = heights.detach().to(torch.float64).contiguous()
        call("glued_set_hill_heights", e.h, 0, self.bias, _p(h), self.H, e.stream)
        lamv = float(lam)
        if lamv != self._lam:
            call("glued_set_path_lambda", e.h, 0, lamv)
            self._lam = lamv
        sig = [float(v) for v in sigma.detach().cpu()]
        if sig != self._sigma:
            arr = (ctypes.c_double * len(sig))(*sig)
            call("glued_set_hill_sigma", e.h, self.bias, arr, len(sig))
            self._sigma = sig
        fr = frames.detach().to("cpu", torch.float64).numpy().reshape(self.N, self.M, 3)
        if not np.array_equal(fr.astype(np.float32), self._frames):
            flat = np.ascontiguousarray(fr, dtype=np.float64).reshape(-1)
            call("glued_set_path_frames", e.h, 0, flat.ctypes.data_as(ctypes.c_void_p), flat.size)
            self._frames = fr.astype(np.float32)
        e.zero_forces()
        e.execute_batch(use_graph=True)
        self.generation += 1
        self._keep = h  # the heights must outlive the asynchronous copy
        return self.generation

    def _buffer(self, which, n):
        out = torch.empty(n, dtype=torch.float64, device=self.device)
        call("glued_copy_buffer", self.ctx.h, int(which), _p(out), out.numel() * 8, self.ctx.stream)
        return out

    def energies(self):
        nb = max(self.force.getNumBiases(), 1)
        return self._buffer(2, self.B * nb).view(self.B, nb)[:, self.bias]


class PathMetadFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, heights, lam, sigma, frames, pipe):
        gen = pipe._evaluate(x, heights, lam, sigma, frames)
        V = pipe.energies().clone()
        ctx.pipe = pipe
        ctx.gen = gen
        ctx.lam_shape = lam.shape if torch.is_tensor(lam) else None
        ctx.lam_device = lam.device if torch.is_tensor(lam) else None
        ctx.save_for_backward(x.detach(), heights.detach(), torch.as_tensor(float(lam), dtype=torch.float64),
                              sigma.detach(), frames.detach())
        return V

    @staticmethod
    def backward(ctx, gV):
        pipe = ctx.pipe
        x, heights, lam, sigma, frames = ctx.saved_tensors
        if pipe.generation != ctx.gen:  # the engine moved on: recompute this forward
            pipe._evaluate(x, heights, lam, sigma, frames)
        e = pipe.ctx
        B, M, P = pipe.B, pipe.M, e.padded[0]
        g = gV.detach().to(torch.float64).contiguous()
        grad_x = grad_h = grad_lam = grad_sig = grad_fr = None
        if ctx.needs_input_grad[0]:
            F = e.forces.view(B, 3, P)[:, :, :M].to(torch.float64) / FIXED_SCALE  # = -dV/dx
            grad_x = (-F.transpose(1, 2) * g[:, None, None]).to(x.dtype)
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[3]:
            gh = torch.zeros(pipe.H, dtype=torch.float64, device=pipe.device)
            gs = torch.zeros(2, dtype=torch.float64, device=pipe.device)
            call("glued_hill_adjoint", e.h, e.stream, pipe.bias, _p(g), _p(gh), _p(gs))
            grad_h = gh.to(heights.dtype) if ctx.needs_input_grad[1] else None
            grad_sig = gs.to(sigma.dtype) if ctx.needs_input_grad[3] else None
            pipe.grad_sigma = gs
        if ctx.needs_input_grad[2]:
            ng = sum(len(b[1]) for b in pipe.force.biases)
            bg = pipe._buffer(1, B * ng).view(B, ng)
            gslot = sum(len(b[1]) for b in pipe.force.biases[:pipe.bias])
            aux = pipe._buffer(4, B * 2).view(B, 2)
            dl = bg[:, gslot] * aux[:, 0] + bg[:, gslot + 1] * aux[:, 1]
            grad_lam = (g * dl).sum()
            if ctx.lam_shape is not None:
                grad_lam = grad_lam.reshape(ctx.lam_shape).to(ctx.lam_device)
        if ctx.needs_input_grad[4]:
            gf = torch.empty(pipe.N * pipe.M * 3, dtype=torch.float64, device=pipe.device)
            call("glued_path_frames_adjoint", e.h, e.stream, 0, _p(g), _p(gf))
            grad_fr = gf.view(pipe.N, pipe.M, 3).to(frames.dtype).to(frames.device)
        return grad_x, grad_h, grad_lam, grad_sig, grad_fr, None


class OpesPipeline:
    """V_b of the OPES bias ``bias_index`` of ``force`` for B configurations of
    ``natoms`` atoms (the walkers of one engine, one kernel table for all).
    The force holds this one bias; its CVs are any the engine evaluates."""

    def __init__(self, force, bias_index, natoms, batch, *, box=None, device="cuda:0"):
        if force.getNumBiases() != 1:
            raise ValueError("OpesPipeline: the force must hold exactly one bias, the OPES one")
        if force.biases[bias_index][0] != force.BIAS_OPES:
            raise ValueError("OpesPipeline: bias_index must name an OPES bias")
        force.share_bias(bias_index)
        self.force = force
        self.bias = int(bias_index)
        self.B, self.natoms = int(batch), int(natoms)
        self.D = len(force.biases[bias_index][1])
        self.cap = int(force.biases[bias_index][3][2])
        self.ctx = Context(force, walkers=[self.natoms] * self.B, device=device)
        if box is not None:
            self.ctx.set_box(box)
        self.device = self.ctx.device
        self.generation = 0
        self.n = 0

    def __call__(self, x, centers, sigmas, logw, sum_uprob):
        """x [B, natoms, 3]; centers, sigmas [n, D]; logw [n]; sum_uprob
        scalar.  Every argument may require grad."""
        return OpesFunction.apply(x, centers, sigmas, logw, sum_uprob, self)

    def _evaluate(self, x, centers, sigmas, logw, sum_uprob):
        e = self.ctx
        n = int(logw.shape[0])
        if n > self.cap:
            raise ValueError(f"OpesPipeline: {n} kernels exceed the capacity {self.cap}")
        c = np.ascontiguousarray(centers.detach().to("cpu", torch.float64).numpy().reshape(n, self.D))
        sg = np.ascontiguousarray(sigmas.detach().to("cpu", torch.float64).numpy().reshape(n, self.D))
        lw = np.ascontiguousarray(logw.detach().to("cpu", torch.float64).numpy())
        e.set_opes_kernels(self.bias, c, sg, lw, float(sum_uprob))
        self.n = n
        P = e.padded[0]
        e.posq.view(self.B, P, 4)[:, :self.natoms, :3].copy_(x.detach().to(torch.float32))
        e.zero_forces()
        e.execute_batch(use_graph=True)
        self.generation += 1
        return self.generation

    def energies(self):
        out = torch.empty(self.B, dtype=torch.float64, device=self.device)
        call("glued_copy_buffer", self.ctx.h, 2, _p(out), out.numel() * 8, self.ctx.stream)
        return out


class OpesFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, centers, sigmas, logw, sum_uprob, pipe):
        gen = pipe._evaluate(x, centers, sigmas, logw, sum_uprob)
        V = pipe.energies().clone()
        ctx.pipe, ctx.gen = pipe, gen
        ctx.u_shape = sum_uprob.shape if torch.is_tensor(sum_uprob) else None
        ctx.u_device = sum_uprob.device if torch.is_tensor(sum_uprob) else None
        ctx.save_for_backward(x.detach(), centers.detach(), sigmas.detach(), logw.detach(),
                              torch.as_tensor(float(sum_uprob), dtype=torch.float64))
        return V

    @staticmethod
    def backward(ctx, gV):
        pipe = ctx.pipe
        x, centers, sigmas, logw, su = ctx.saved_tensors
        if pipe.generation != ctx.gen:  # the engine moved on: recompute this forward
            pipe._evaluate(x, centers, sigmas, logw, su)
        e = pipe.ctx
        B, N, P, D, n, cap = pipe.B, pipe.natoms, e.padded[0], pipe.D, pipe.n, pipe.cap
        g = gV.detach().to(torch.float64).contiguous()
        grad_x = grad_c = grad_s = grad_l = grad_u = None
        if ctx.needs_input_grad[0]:
            F = e.forces.view(B, 3, P)[:, :, :N].to(torch.float64) / FIXED_SCALE  # = -dV/dx
            grad_x = (-F.transpose(1, 2) * g[:, None, None]).to(x.dtype)
        if any(ctx.needs_input_grad[1:5]):
            gc = torch.empty(cap * D, dtype=torch.float64, device=pipe.device)
            gs = torch.empty(cap * D, dtype=torch.float64, device=pipe.device)
            gl = torch.empty(cap, dtype=torch.float64, device=pipe.device)
            gu = torch.empty(1, dtype=torch.float64, device=pipe.device)
            call("glued_opes_adjoint", e.h, e.stream, pipe.bias, _p(g), _p(gc), _p(gs), _p(gl), _p(gu))
            if ctx.needs_input_grad[1]:
                grad_c = gc.view(cap, D)[:n].reshape(centers.shape).to(centers.dtype).to(centers.device)
            if ctx.needs_input_grad[2]:
                grad_s = gs.view(cap, D)[:n].reshape(sigmas.shape).to(sigmas.dtype).to(sigmas.device)
            if ctx.needs_input_grad[3]:
                grad_l = gl[:n].to(logw.dtype).to(logw.device)
            if ctx.needs_input_grad[4]:
                grad_u = gu[0]
                if ctx.u_shape is not None:
                    grad_u = grad_u.reshape(ctx.u_shape).to(ctx.u_device)
        return grad_x, grad_c, grad_s, grad_l, grad_u, None
