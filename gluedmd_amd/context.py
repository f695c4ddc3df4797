"""``Context``: evaluates a ``glued.Force`` on the GPU through the C-ABI.

Stands in for the part of ``openmm.Context`` this path touches: device
positions (``posq``, float4 in GPU atom order) go in, 2^32 fixed-point forces
(int64, field-major, OpenMM's long force buffer) come out, and the bias state
lives on the device (GluedForceImpl / CommonCalcGluedForceKernel in the
reference).  PyTorch is used only to own device memory and name the stream.

One Context holds one walker (the reference's structure) or W walkers of
uneven size (batched form): walker w's atoms start at sum_{v<w} padded_v.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call

FIXED_SCALE = 4294967296.0


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_lib.c_int_p)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_lib.c_double_p)


def pad64(n):
    return ((int(n) + 63) // 64) * 64


def box_vectors(box):
    """Accept an edge length (cubic), 3 edge lengths, a 3x3 matrix or 9 values."""
    b = np.asarray(box, dtype=np.float64)
    if b.ndim == 0:
        return np.diag([float(b)] * 3).reshape(9)
    if b.size == 3:
        return np.diag(b).reshape(9)
    return b.reshape(9)



def _current_raw_stream(index):
    get = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if get is not None:
        return get(index)
    return torch.cuda.current_stream(index).cuda_stream


class Context:
    def __init__(self, force, num_particles=None, masses=None, *, walkers=None, padded=None, device="cuda:0"):
        if walkers is None:
            if num_particles is None:
                raise ValueError("num_particles or walkers is required")
            walkers = [int(num_particles)]
        self.natoms = [int(n) for n in walkers]
        self.W = len(self.natoms)
        self.padded = [int(p) for p in padded] if padded is not None else [pad64(n) for n in self.natoms]
        self.offsets = np.concatenate([[0], np.cumsum(self.padded)]).astype(np.int64)
        self.device = torch.device(device)
        self._dev_index = self.device.index if self.device.index is not None else torch.cuda.current_device()
        self.force = force
        self.step_count = 0
        h = ctypes.c_void_p()
        call("glued_create", ctypes.byref(h))
        self.h = h
        try:
            self._replay(force, masses)
        except Exception:
            _lib.load().glued_destroy(self.h)
            self.h = None
            raise
        total = int(self.offsets[-1])
        self.posq = torch.zeros((total, 4), dtype=torch.float32, device=self.device)
        self.forces = torch.zeros(3 * total, dtype=torch.int64, device=self.device)

    def _replay(self, force, masses):
        call("glued_set_pbc", self.h, 1 if force.usesPeriodicBoundaryConditions() else 0)
        if force.getTemperature() > 0:
            call("glued_set_temperature", self.h, force.getTemperature())
        for t, atoms, params in force.cvs:
            a, ap = _i32(atoms if atoms else [0])
            p, pp = _f64(params if params else [0.0])
            out = ctypes.c_int()
            call("glued_add_cv", self.h, t, ap, len(atoms), pp, len(params), ctypes.byref(out))
        for b, (t, cvs, params, ip) in enumerate(force.biases):
            c, cp = _i32(cvs if cvs else [0])
            p, pp = _f64(params if params else [0.0])
            q, qp = _i32(ip if ip else [0])
            out = ctypes.c_int()
            call("glued_add_bias", self.h, t, cp, len(cvs), pp, len(params), qp, len(ip), ctypes.byref(out))
        for b in sorted(force.shared_biases):
            call("glued_share_bias", self.h, int(b), 1)
        if force._test_grads is not None:
            g, gp = _f64(force._test_grads)
            call("glued_set_test_bias_gradients", self.h, gp, len(force._test_grads))
        n, np_ = _i32(self.natoms)
        pd, pdp = _i32(self.padded)
        if masses is None:
            mp, nm = None, 0
        else:
            m, mp = _f64(masses)
            self._masses = m
            nm = len(m)
        call("glued_initialize_batch", self.h, self.W, np_, pdp, mp, nm)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                _lib.load().glued_destroy(h)
            except Exception:
                pass
            self.h = None

    @property
    def stream(self):
        # torch's current stream on this device (the raw handle: the public
        # accessor builds a Stream object, ~2 us per call)
        return ctypes.c_void_p(_current_raw_stream(self._dev_index))

    def walker_slice(self, w):
        return slice(int(self.offsets[w]), int(self.offsets[w + 1]))

    # ---- inputs ----
    def set_box(self, box, walker=None):
        b, bp = _f64(box_vectors(box))
        for w in range(self.W) if walker is None else [walker]:
            call("glued_set_box", self.h, int(w), bp)

    def set_positions(self, pos, walker=0):
        """pos: (N, 3) array/tensor in user order (identity GPU order)."""
        t = torch.as_tensor(pos, dtype=torch.float32).to(self.device)
        n = t.shape[0]
        o = int(self.offsets[walker])
        self.posq[o:o + n, :3].copy_(t)

    def set_all_positions(self, pos_list):
        for w, p in enumerate(pos_list):
            self.set_positions(p, w)

    def set_atom_order(self, gpu_to_user, walker=0):
        a, ap = _i32(gpu_to_user)
        call("glued_set_atom_order", self.h, int(walker), ap, len(a))

    def set_energy_input(self, U, force_field, walker=0):
        """U (kJ/mol) and the external force (torch float32 on the device,
        field-major 3*padded) of the CV_ENERGY for this walker."""
        assert force_field.dtype == torch.float32 and force_field.is_cuda
        assert force_field.numel() >= 3 * self.padded[walker]
        self._energy_keep = getattr(self, "_energy_keep", {})
        self._energy_keep[walker] = force_field
        call("glued_set_energy_input", self.h, int(walker), float(U), ctypes.c_void_p(force_field.data_ptr()),
             self.stream)

    # ---- the step ----
    def zero_forces(self):
        self.forces.zero_()

    def execute(self, walker=0, include_forces=True, include_energy=True):
        e = ctypes.c_double()
        o = int(self.offsets[walker])
        pos_ptr = self.posq.data_ptr() + 16 * o
        f_ptr = self.forces.data_ptr() + 8 * 3 * o
        call("glued_execute", self.h, self.stream, int(walker), ctypes.c_void_p(pos_ptr), ctypes.c_void_p(f_ptr),
             int(include_forces), int(include_energy), ctypes.byref(e))
        return e.value

    def execute_all_plain(self, include_forces=True):
        """Plain form for every walker: the reference's one-walker-per-launch
        structure, walker after walker, looped inside the library."""
        call("glued_execute_plain_all", self.h, self.stream, ctypes.c_void_p(self.posq.data_ptr()),
             ctypes.c_void_p(self.forces.data_ptr()), int(include_forces))

    def execute_batch(self, include_forces=True, use_graph=True):
        call("glued_execute_batch", self.h, self.stream, ctypes.c_void_p(self.posq.data_ptr()),
             ctypes.c_void_p(self.forces.data_ptr()), int(include_forces), int(use_graph))

    def execute_stage(self, stage):
        """0 = CVs, 1 = biases, 2 = dU/dCV + scatter, 3 = streamed hill /
        OPES-kernel sums only, 4 = the force scatter only (the combined dense
        energy-CV + atom-slot pass where the batched step uses it); every walker."""
        call("glued_execute_stage", self.h, self.stream, ctypes.c_void_p(self.posq.data_ptr()),
             ctypes.c_void_p(self.forces.data_ptr()), int(stage))

    def update_state(self, step):
        call("glued_update_state", self.h, self.stream, int(step))

    def step_batch(self, step, include_forces=True):
        """update_state(step) then execute_batch(), as one graph launch."""
        call("glued_step_batch", self.h, self.stream, int(step), ctypes.c_void_p(self.posq.data_ptr()),
             ctypes.c_void_p(self.forces.data_ptr()), int(include_forces))

    def md_step(self, walker=0):
        """One integrator step as OpenMM orders it: updateContextState(step)
        (deposition of the previous evaluation), then the force evaluation."""
        self.update_state(self.step_count)
        self.zero_forces()
        e = self.execute(walker)
        self.step_count += 1
        return e

    # ---- read-backs ----
    def get_forces(self, walker=0):
        """Forces of one walker as float64 (N, 3), kJ/mol/nm."""
        torch.cuda.synchronize(self.device)
        o, p, n = int(self.offsets[walker]), self.padded[walker], self.natoms[walker]
        blk = self.forces[3 * o:3 * o + 3 * p].view(3, p)[:, :n].cpu().numpy()
        return (blk.T.astype(np.float64)) / FIXED_SCALE

    def get_forces_fixed(self, walker=0):
        torch.cuda.synchronize(self.device)
        o, p, n = int(self.offsets[walker]), self.padded[walker], self.natoms[walker]
        return self.forces[3 * o:3 * o + 3 * p].view(3, p)[:, :n].cpu().numpy().T.copy()

    def getLastCVValues(self, walker=0):
        n = self.force.getNumCollectiveVariables()
        out = np.zeros(max(n, 1))
        call("glued_get_cv_values", self.h, int(walker), out.ctypes.data_as(_lib.c_double_p), len(out))
        return out[:n]

    def getBiasEnergies(self, walker=0):
        n = self.force.getNumBiases()
        out = np.zeros(max(n, 1))
        call("glued_get_bias_energies", self.h, int(walker), out.ctypes.data_as(_lib.c_double_p), len(out))
        return out[:n]

    def getLastBias(self, walker=0):
        return float(np.sum(self.getBiasEnergies(walker)))

    def getDUDCV(self, walker=0):
        n = self.force.getNumCollectiveVariables()
        out = np.zeros(max(n, 1))
        call("glued_get_dudcv", self.h, int(walker), out.ctypes.data_as(_lib.c_double_p), len(out))
        return out[:n]

    def coordination_pairs(self, coord_index=0, walker=0):
        n = ctypes.c_longlong()
        call("glued_get_coordination_pairs", self.h, int(walker), int(coord_index), ctypes.byref(n))
        return n.value

    def getBiasState(self, walker=0):
        size = ctypes.c_size_t()
        call("glued_get_bias_state", self.h, int(walker), None, 0, ctypes.byref(size))
        buf = ctypes.create_string_buffer(size.value)
        call("glued_get_bias_state", self.h, int(walker), buf, size.value, ctypes.byref(size))
        return buf.raw[:size.value]

    def setBiasState(self, blob, walker=0):
        call("glued_set_bias_state", self.h, int(walker), blob, len(blob))

    def getMultiWalkerPtrs(self, bias_index):
        """Device addresses of a single-table bias (MetaD grid, hill list or
        OPES table) for setMultiWalkerPtrs on another context of the device."""
        n = ctypes.c_int()
        buf = (ctypes.c_int64 * 8)()
        call("glued_get_multiwalker_ptrs", self.h, int(bias_index), buf, 8, ctypes.byref(n))
        return [int(buf[k]) for k in range(n.value)]

    def setMultiWalkerPtrs(self, bias_index, ptrs, primary=None):
        """Share another context's table of this bias (its getMultiWalkerPtrs);
        `primary`, when given, is kept alive as long as this context."""
        arr = (ctypes.c_int64 * max(len(ptrs), 1))(*[int(p) for p in ptrs])
        call("glued_redirect_to_primary_bias", self.h, int(bias_index), arr, len(ptrs))
        if primary is not None:
            self._primaries = getattr(self, "_primaries", []) + [primary]

    def getOPESMetrics(self, bias_index=0, walker=0):
        out = np.zeros(4)
        call("glued_get_opes_metrics", self.h, int(walker), int(bias_index), out.ctypes.data_as(_lib.c_double_p))
        return list(out)

    def getKernelSigmas(self, bias_index=0, walker=0):
        n = ctypes.c_int()
        call("glued_get_kernel_sigmas", self.h, int(walker), int(bias_index), None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), dtype=np.float32)
        call("glued_get_kernel_sigmas", self.h, int(walker), int(bias_index), out.ctypes.data_as(_lib.c_float_p),
             len(out), ctypes.byref(n))
        return out[:n.value]

    def set_hills(self, bias_index, centers, heights, walker=0):
        c, cp = _f64(np.asarray(centers, dtype=np.float64).reshape(-1))
        h, hp = _f64(heights)
        call("glued_set_hills", self.h, int(walker), int(bias_index), cp, hp, len(h))

    def get_hills(self, bias_index, walker=0):
        n = ctypes.c_int()
        call("glued_get_hills", self.h, int(walker), int(bias_index), None, None, 0, ctypes.byref(n))
        D = len(self.force.biases[bias_index][1])
        c = np.zeros((max(n.value, 1), D))
        h = np.zeros(max(n.value, 1))
        call("glued_get_hills", self.h, int(walker), int(bias_index), c.ctypes.data_as(_lib.c_double_p),
             h.ctypes.data_as(_lib.c_double_p), max(n.value, 1), ctypes.byref(n))
        return c[:n.value], h[:n.value]

    def set_grid(self, bias_index, grid, walker=0):
        g, gp = _f64(np.asarray(grid).reshape(-1))
        call("glued_set_grid", self.h, int(walker), int(bias_index), gp, len(g))

    def set_opes_kernels(self, bias_index, centers, sigmas, logw, sum_uprob, walker=0):
        """An OPES kernel table (kernel-major centres and sigmas [n][D],
        log-weights [n]) and its sum_uprob."""
        c, cp = _f64(np.asarray(centers, dtype=np.float64).reshape(-1))
        s, sp = _f64(np.asarray(sigmas, dtype=np.float64).reshape(-1))
        lw, lp = _f64(logw)
        call("glued_set_opes_kernels", self.h, int(walker), int(bias_index), cp, sp, lp, len(lw), float(sum_uprob))

    def tuning(self):
        """Launch shapes the autotuner chose for this plan, and how."""
        n = ctypes.c_size_t()
        call("glued_get_tuning", self.h, None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value + 1)
        call("glued_get_tuning", self.h, buf, n.value + 1, ctypes.byref(n))
        return buf.value.decode()

    def plan_info(self):
        r, a = ctypes.c_int(), ctypes.c_int()
        call("glued_plan_info", self.h, ctypes.byref(r), ctypes.byref(a))
        return r.value, a.value
