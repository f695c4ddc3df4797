"""FP64 CPU oracle of the GLUED enhanced-sampling step (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker or the timed CPU baseline.  The product
path (gluedmd_amd) never imports it.

The arithmetic lives in glued_oracle.c (one function per reference kernel,
each citing CommonGluedKernels.cpp); this module replays a force
specification through it with the reference's step semantics:
  * execute: CVs, then biases, then dU/dCV = sum of bias gradients (or the
    injected test gradients), then the chain rule
    (CommonGluedKernels.cpp:4646-5048);
  * update_state: skipped before the first evaluation, once per step, never
    at step 0, pace-gated; OPES, OPES-expanded, multithermal, then MetaD, all
    from the previous evaluation's CVs and energies (:5050-5157).
Forces come out in FP64 and, separately, through the reference's per-row
fixed-point truncation with FP32 Jacobian rows (:2430-2456).
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libglued_oracle.so")

CV_DISTANCE, CV_ANGLE, CV_DIHEDRAL, CV_COM_DISTANCE, CV_GYRATION, CV_COORDINATION, CV_RMSD = 1, 2, 3, 4, 5, 6, 7
CV_PATH, CV_POSITION, CV_DRMSD, CV_CONTACTMAP, CV_ENERGY = 10, 11, 12, 13, 23
BIAS_METAD, BIAS_OPES, BIAS_OPES_EXPANDED, BIAS_OPES_MULTITHERMAL, BIAS_METAD_HILLS = 3, 5, 11, 15, 101
BIAS_HARMONIC, BIAS_UPPER_WALL, BIAS_LOWER_WALL, BIAS_LINEAR, BIAS_PBMETAD = 1, 8, 9, 10, 4
FIXED = 4294967296.0

_lib = None
dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)
fp = ctypes.POINTER(ctypes.c_float)
llp = ctypes.POINTER(ctypes.c_longlong)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        d, i = ctypes.c_double, ctypes.c_int
        sig = {
            "orc_delta_pbc": (None, [dp, dp]),
            "orc_cv_distance": (d, [dp, i, i, dp, i, dp]),
            "orc_cv_angle": (d, [dp, ip, dp, i, dp]),
            "orc_cv_dihedral": (d, [dp, ip, dp, i, dp]),
            "orc_cv_com_distance": (d, [dp, ip, i, i, dp, dp, i, dp]),
            "orc_cv_gyration": (d, [dp, ip, i, dp, dp, i, dp]),
            "orc_switch": (d, [d, d, d, d, dp]),
            "orc_cv_coordination": (d, [dp, ip, i, ip, i, d, d, d, d, dp, i, dp, dp, llp]),
            "orc_cv_rmsd": (d, [dp, ip, i, dp, i, dp, dp]),
            "orc_cv_path": (None, [dp, ip, i, dp, i, d, dp, dp, dp, dp]),
            "orc_cv_contactmap": (d, [dp, ip, i, dp, dp, i, dp]),
            "orc_cv_drmsd": (d, [dp, ip, i, dp, dp, i, dp]),
            "orc_scatter_fixed": (None, [i, ip, fp, fp, fp, ip, dp, llp, i]),
            "orc_metad_grid_eval": (d, [dp, i, dp, dp, dp, ip, ip, ip, dp]),
            "orc_metad_grid_deposit": (None, [dp, i, i, dp, dp, ip, ip, ip, dp, dp, d]),
            "orc_hills_eval": (d, [dp, i, i, dp, dp, dp, dp, dp]),
            "orc_hill_min_margin": (d, [dp, i, i, dp, dp, dp]),
            "orc_opes_eval": (d, [dp, i, i, dp, dp, dp, d, d, d, d, d, d, dp]),
            "orc_opes_constants": (None, [i, d, d, dp]),
            "orc_opes_welford": (None, [dp, i, dp, dp, ip]),
            "orc_opes_deposit": (i, [dp, i, d, d, i, i, d, d, dp, d, i, dp, dp, dp, dp, ip, dp, dp]),
            "orc_opes_sum_uprob": (d, [i, i, dp, dp, dp, d]),
            "orc_opes_expanded_eval": (d, [dp, i, dp, d, d, dp]),
            "orc_opes_expanded_update_logz": (None, [dp, i, dp, d, dp, ip]),
            "orc_multithermal_eval": (d, [d, dp, dp, i, d, dp]),
            "orc_multithermal_update": (None, [d, d, dp, dp, dp, dp, i, d]),
            "orc_harmonic_eval": (d, [dp, i, dp, i, dp]),
            "orc_linear_eval": (d, [dp, i, dp, dp]),
            "orc_wall_eval": (d, [dp, i, dp, i, dp]),
            "orc_pbmetad_eval": (d, [dp, i, dp, dp, d, dp, dp, dp]),
            "orc_pbmetad_deposit": (None, [dp, dp, d, d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def P(a, kind=np.float64):
    """Contiguous array and its ctypes pointer (the array must be kept alive)."""
    a = np.ascontiguousarray(a, dtype=kind)
    ptr = {np.float64: dp, np.int32: ip, np.float32: fp, np.int64: llp}[kind]
    return a, a.ctypes.data_as(ptr)


def f32(x):
    """Value as the reference stores it (float)."""
    return float(np.float32(x))


def f32v(xs):
    """Values as the reference stores them (float), back in FP64."""
    return np.asarray(xs, dtype=np.float64).astype(np.float32).astype(np.float64)


# ---------------------------------------------------------------------------
# thin wrappers (arrays in, numbers out)
# ---------------------------------------------------------------------------
def box9(box):
    b = np.asarray(box, dtype=np.float64)
    if b.ndim == 0:
        return np.diag([float(b)] * 3).reshape(9)
    if b.size == 3:
        return np.diag(b).reshape(9)
    return b.reshape(9).copy()


def switch(x, r0, n, m):
    g = ctypes.c_double()
    f = lib().orc_switch(x, r0, n, m, ctypes.byref(g))
    return f, g.value


def opes_constants(variant, gamma, kT):
    out = np.zeros(4)
    lib().orc_opes_constants(variant, gamma, kT, out.ctypes.data_as(dp))
    return tuple(out)


def hills_eval(cv, centers, heights, sigma, period):
    cv, cvp = P(cv)
    c, cp = P(np.asarray(centers).reshape(-1))
    h, hp = P(heights)
    s, sp = P(sigma)
    per, pp = P(period)
    D = len(cv)
    dV = np.zeros(D)
    V = lib().orc_hills_eval(cvp, D, len(h), cp, hp, sp, pp, dV.ctypes.data_as(dp))
    return V, dV


def hill_margin(cv, centers, sigma, period):
    cv, cvp = P(cv)
    c, cp = P(np.asarray(centers).reshape(-1))
    s, sp = P(sigma)
    per, pp = P(period)
    return lib().orc_hill_min_margin(cvp, len(cv), len(c) // len(cv), cp, sp, pp)


def opes_sum_uprob(centers, sigmas, logw, cutoff2):
    c, cp = P(np.asarray(centers).reshape(-1))
    s, sp = P(np.asarray(sigmas).reshape(-1))
    lw, lp = P(logw)
    D = len(c) // max(len(lw), 1) if len(lw) else 1
    return lib().orc_opes_sum_uprob(len(lw), D, cp, sp, lp, cutoff2)


# ---------------------------------------------------------------------------
# Bias states
# ---------------------------------------------------------------------------
class _Grid:
    def __init__(self, D, params, iparams):
        self.D = D
        self.h0 = params[0]
        self.sigma = np.array(params[1:1 + D])
        gamma, kT = params[1 + D], params[2 + D]
        mn = np.array(params[3 + D:3 + 2 * D])
        mx = np.array(params[3 + 2 * D:3 + 3 * D])
        self.pace = iparams[0]
        bins = np.array(iparams[1:1 + D])
        self.periodic = np.array(iparams[1 + D:1 + 2 * D], dtype=np.int32)
        self.origin = mn
        self.invSpacing = bins / (mx - mn)
        self.npts = (bins + np.where(self.periodic != 0, 0, 1)).astype(np.int32)
        self.strides = np.ones(D, dtype=np.int32)
        for d in range(1, D):
            self.strides[d] = self.strides[d - 1] * self.npts[d - 1]
        self.total = int(self.strides[-1] * self.npts[-1])
        self.grid = np.zeros(self.total)
        self.hif = 0.0 if gamma <= 1.0 else 1.0 / ((gamma - 1.0) * kT)
        self.numDeposited = 0

    def eval(self, s):
        s, sp = P(s)
        dV = np.zeros(self.D)
        V = lib().orc_metad_grid_eval(sp, self.D, self.grid.ctypes.data_as(dp), self.origin.ctypes.data_as(dp),
                                      self.invSpacing.ctypes.data_as(dp),
                                      self.npts.ctypes.data_as(ip), self.strides.ctypes.data_as(ip),
                                      self.periodic.ctypes.data_as(ip), dV.ctypes.data_as(dp))
        return V, dV

    def deposit(self, center, V):
        height = self.h0 * math.exp(-V * self.hif) if self.hif > 0.0 else self.h0
        c, cp = P(center)
        lib().orc_metad_grid_deposit(self.grid.ctypes.data_as(dp), self.D, self.total, self.origin.ctypes.data_as(dp),
                                     self.invSpacing.ctypes.data_as(dp), self.npts.ctypes.data_as(ip),
                                     self.strides.ctypes.data_as(ip), self.periodic.ctypes.data_as(ip), cp,
                                     self.sigma.ctypes.data_as(dp), height)
        self.numDeposited += 1


def _fp32(a):
    """FP64 array of the FP32-rounded values (round to nearest)."""
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


class _Hills:
    """Hill-list metadynamics state.  Centres and heights are kept as FP32
    values (set or deposited values are rounded to nearest), the precision
    the engine stores them in; the sums over them are FP64."""

    def __init__(self, D, params, iparams):
        self.D = D
        self.h0 = params[0]
        self.sigma = np.array(params[1:1 + D])
        gamma, kT = params[1 + D], params[2 + D]
        self.period = np.array(params[3 + D:3 + 2 * D])
        self.hif = 0.0 if gamma <= 1.0 else 1.0 / ((gamma - 1.0) * kT)
        self.pace, self.cap = iparams[0], iparams[1]
        self.centers = np.zeros((0, D))
        self.heights = np.zeros(0)

    @property
    def centers(self):
        return self._centers

    @centers.setter
    def centers(self, c):
        self._centers = _fp32(c).reshape(-1, self.D)

    @property
    def heights(self):
        return self._heights

    @heights.setter
    def heights(self, h):
        self._heights = _fp32(h).reshape(-1)

    def eval(self, s):
        if len(self.heights) == 0:
            return 0.0, np.zeros(self.D)
        return hills_eval(s, self.centers, self.heights, self.sigma, self.period)

    def deposit(self, center, V):
        if len(self.heights) >= self.cap:
            return
        h = self.h0 * math.exp(-V * self.hif) if self.hif > 0.0 else self.h0
        self.centers = np.vstack([self.centers, np.asarray(center, dtype=np.float64)[None, :]])
        self.heights = np.append(self.heights, h)


class _Opes:
    def __init__(self, D, params, iparams):
        self.D = D
        self.kT, self.gamma = params[0], params[1]
        self.sigma0 = np.array(params[2:2 + D])
        self.sigmaMin = params[2 + D]
        self.variant = iparams[0] if len(iparams) > 0 else 0
        self.pace = iparams[1] if len(iparams) > 1 else 500
        self.cap = iparams[2] if len(iparams) > 2 else 100000
        adaptive = self.sigma0[0] <= 0.0
        self.adaptiveStride = (iparams[3] if len(iparams) > 3 else 10 * self.pace) if adaptive else 0
        self.invGF, self.cutoff2, self.eps, self.barrier = opes_constants(self.variant, self.gamma, self.kT)
        if self.variant == 2 and not adaptive:
            self.sigma0 = self.sigma0 * math.sqrt(self.gamma)
        self.centers = np.zeros(self.cap * D)
        self.sigmas = np.ones(self.cap * D)
        self.logw = np.zeros(self.cap)
        # {sum_uprob, sumWeights, sum_w, sum_w2} and {numKernels, numAllocated, stepCount, nSamples}
        self.st_d = np.array([0.0, math.exp(-self.gamma), math.exp(-self.gamma), math.exp(-2.0 * self.gamma)])
        self.st_i = np.zeros(4, dtype=np.int32)
        self.mean = np.zeros(D)
        self.M2 = np.zeros(D)
        self.depositMirror = 0

    @property
    def nk(self):
        return int(self.st_i[0])

    def table(self):
        n, D = self.nk, self.D
        return self.centers[:n * D].reshape(n, D).copy(), self.sigmas[:n * D].reshape(n, D).copy(), self.logw[:n].copy()

    def eval(self, s):
        s, sp = P(s)
        dV = np.zeros(self.D)
        V = lib().orc_opes_eval(sp, self.D, self.nk, self.centers.ctypes.data_as(dp), self.sigmas.ctypes.data_as(dp),
                                self.logw.ctypes.data_as(dp), self.st_d[0], self.invGF, self.kT, self.cutoff2,
                                self.eps, self.barrier, dV.ctypes.data_as(dp))
        return V, dV

    def welford(self, s):
        s, sp = P(s)
        ns = np.array([self.st_i[3]], dtype=np.int32)
        lib().orc_opes_welford(sp, self.D, self.mean.ctypes.data_as(dp), self.M2.ctypes.data_as(dp),
                               ns.ctypes.data_as(ip))
        self.st_i[3] = ns[0]

    def deposit(self, center, V):
        c, cp = P(center)
        return lib().orc_opes_deposit(cp, self.D, V, 1.0 / self.kT, self.variant, self.adaptiveStride, self.gamma,
                                      self.sigmaMin, self.sigma0.ctypes.data_as(dp), self.cutoff2, self.cap,
                                      self.centers.ctypes.data_as(dp), self.sigmas.ctypes.data_as(dp),
                                      self.logw.ctypes.data_as(dp), self.st_d.ctypes.data_as(dp),
                                      self.st_i.ctypes.data_as(ip), self.mean.ctypes.data_as(dp),
                                      self.M2.ctypes.data_as(dp))

    def load(self, centers, sigmas, logw, sum_uprob, nSamples=0, mean=None, M2=None):
        """The restore rule of CommonSerialization.cpp:214-264 (KDNorm recomputed);
        the slot-claim counter follows the restored count."""
        n = len(logw)
        D = self.D
        self.centers[:n * D] = np.asarray(centers).reshape(-1)
        self.sigmas[:n * D] = np.asarray(sigmas).reshape(-1)
        self.logw[:n] = logw
        self.st_i[0] = n
        self.st_i[1] = n
        self.st_i[3] = nSamples
        self.st_d[0] = sum_uprob
        self.st_d[1] = math.exp(-self.gamma) + float(np.sum(np.exp(logw)))
        if mean is not None:
            self.mean[:] = mean
        if M2 is not None:
            self.M2[:] = M2
        self.depositMirror = n

    def metrics(self):
        nker = self.nk
        counter = float(self.st_i[2])
        KDEnorm = counter if self.variant == 2 else self.st_d[1]
        zed = self.st_d[0] / (KDEnorm * nker) if (nker > 0 and KDEnorm > 0) else 1.0
        rct = self.kT * math.log(self.st_d[1] / counter) if (counter > 0 and self.st_d[1] > 0) else 0.0
        sw, sw2 = self.st_d[2], self.st_d[3]
        return [zed, rct, float(nker), (1 + sw) ** 2 / (1 + sw2)]


class _Expanded:
    def __init__(self, D, params, iparams):
        self.kT = params[0]
        w = np.array(params[1:1 + D])
        ws = w.sum()
        if ws <= 0:
            ws = D
        self.logW = np.log(w / ws)
        self.pace = iparams[0] if iparams else 500
        self.logZ = np.zeros(1)
        self.nUpd = np.zeros(1, dtype=np.int32)

    def eval(self, ecv):
        e, epp = P(ecv)
        dV = np.zeros(len(e))
        V = lib().orc_opes_expanded_eval(epp, len(e), self.logW.ctypes.data_as(dp), 1.0 / self.kT, self.logZ[0],
                                         dV.ctypes.data_as(dp))
        return V, dV

    def update(self, ecv):
        e, epp = P(ecv)
        lib().orc_opes_expanded_update_logz(epp, len(e), self.logW.ctypes.data_as(dp), 1.0 / self.kT,
                                            self.logZ.ctypes.data_as(dp), self.nUpd.ctypes.data_as(ip))


class _Multithermal:
    def __init__(self, params, iparams):
        self.kT0 = params[0]
        beta0 = 1.0 / self.kT0
        self.dbeta = np.array(params[1:]) - beta0
        self.pace = iparams[0] if iparams else 500
        self.dF = np.zeros(len(self.dbeta))
        self.rct = np.zeros(1)
        self.counter = np.ones(1)

    def eval(self, U):
        g = ctypes.c_double()
        V = lib().orc_multithermal_eval(U, self.dbeta.ctypes.data_as(dp), self.dF.ctypes.data_as(dp), len(self.dF),
                                        self.kT0, ctypes.byref(g))
        return V, np.array([g.value])

    def update(self, U, V):
        lib().orc_multithermal_update(U, V, self.dbeta.ctypes.data_as(dp), self.dF.ctypes.data_as(dp),
                                      self.rct.ctypes.data_as(dp), self.counter.ctypes.data_as(dp), len(self.dF),
                                      self.kT0)


class _PBMetaD:
    """PBMetaD state: one 1-D grid per CV; the last evaluation's V_i and
    softmax weights w_i feed the deposit (heights w_i h0 e^(-V_i hif))."""

    def __init__(self, D, params, iparams):
        self.N, self.h0, gamma, self.kT = D, params[0], params[1], params[2]
        self.hif = 0.0 if gamma <= 1.0 else 1.0 / ((gamma - 1.0) * self.kT)
        self.pace = iparams[0]
        c, off = [], 0
        for d in range(D):
            bins, per = iparams[1 + 2 * d], iparams[2 + 2 * d]
            sigma, lo, hi = params[3 + 3 * d], params[4 + 3 * d], params[5 + 3 * d]
            np_ = bins + (0 if per else 1)
            c += [lo, bins / (hi - lo), float(np_), float(per != 0), sigma, float(off)]
            off += np_
        self.c = np.array(c, dtype=np.float64)
        self.grid = np.zeros(off)
        self.numDeposited = 0
        self.Vi = np.zeros(D)
        self.wi = np.full(D, 1.0 / D)

    def eval(self, s):
        s, sp = P(s)
        g = np.zeros(self.N)
        V = lib().orc_pbmetad_eval(sp, self.N, self.c.ctypes.data_as(dp), self.grid.ctypes.data_as(dp), self.kT,
                                   self.Vi.ctypes.data_as(dp), self.wi.ctypes.data_as(dp), g.ctypes.data_as(dp))
        return V, g

    def deposit(self, s, Vi=None, wi=None):
        Vi = self.Vi if Vi is None else Vi
        wi = self.wi if wi is None else wi
        for i in range(self.N):
            off = int(self.c[6 * i + 5])
            h = wi[i] * self.h0 * math.exp(-Vi[i] * self.hif) if self.hif > 0.0 else wi[i] * self.h0
            sub = self.grid[off:off + int(self.c[6 * i + 2])]  # a view: deposited in place
            lib().orc_pbmetad_deposit(sub.ctypes.data_as(dp), self.c[6 * i:].ctypes.data_as(dp), float(s[i]), h)
        self.numDeposited += 1


class _Restraint:
    """Harmonic, linear and wall restraints: no state, no deposits.  The
    parameters are rounded to FP32 as the reference stores them."""

    def __init__(self, t, D, params, iparams):
        self.t, self.D = t, D
        per = {BIAS_HARMONIC: 2, BIAS_LINEAR: 1}.get(t, 4)
        self.p = np.asarray(params[:per * D], dtype=np.float32).astype(np.float64)
        self.periodic = int(iparams[0]) if (t == BIAS_HARMONIC and iparams) else 0

    def eval(self, s):
        s, sp = P(s)
        g = np.zeros(self.D)
        if self.t == BIAS_HARMONIC:
            V = lib().orc_harmonic_eval(sp, self.D, self.p.ctypes.data_as(dp), self.periodic, g.ctypes.data_as(dp))
        elif self.t == BIAS_LINEAR:
            V = lib().orc_linear_eval(sp, self.D, self.p.ctypes.data_as(dp), g.ctypes.data_as(dp))
        else:
            V = lib().orc_wall_eval(sp, self.D, self.p.ctypes.data_as(dp), int(self.t == BIAS_LOWER_WALL),
                                    g.ctypes.data_as(dp))
        return V, g


# ---------------------------------------------------------------------------
# Spec-level evaluator
# ---------------------------------------------------------------------------
class OracleContext:
    """One walker.  `force` is a gluedmd_amd.glued.Force (read as plain data:
    .cvs, .biases, pbc flag, test gradients).  Shared bias states can be
    passed in `shared` (bias index -> state object) to model walkers that
    deposit into one state."""

    def __init__(self, force, num_particles, masses=None, shared=None):
        self.force = force
        self.N = int(num_particles)
        m = np.ones(self.N) if masses is None else np.asarray(masses, dtype=np.float64)
        self.masses = m.astype(np.float32).astype(np.float64)  # the reference uploads float masses
        self.pbc = force.usesPeriodicBoundaryConditions()
        self.box = np.diag([1.0, 1.0, 1.0]).reshape(9)
        self.nV = force.getNumCollectiveVariables()
        self.states = []
        shared = shared or {}
        for b, (t, cvs, params, iparams) in enumerate(force.biases):
            D = len(cvs)
            if b in shared:
                self.states.append(shared[b])
            elif t == BIAS_METAD:
                self.states.append(_Grid(D, params, iparams))
            elif t == BIAS_METAD_HILLS:
                self.states.append(_Hills(D, params, iparams))
            elif t == BIAS_OPES:
                self.states.append(_Opes(D, params, iparams))
            elif t == BIAS_OPES_EXPANDED:
                self.states.append(_Expanded(D, params, iparams))
            elif t == BIAS_OPES_MULTITHERMAL:
                self.states.append(_Multithermal(params, iparams))
            elif t == BIAS_PBMETAD:
                self.states.append(_PBMetaD(D, params, iparams))
            elif t in (BIAS_HARMONIC, BIAS_LINEAR, BIAS_UPPER_WALL, BIAS_LOWER_WALL):
                self.states.append(_Restraint(t, D, params, iparams))
            else:
                raise ValueError(f"oracle: bias type {t} not covered")
        self.cv_ready = False
        self.last_update = -1
        self.cv = None
        self.benergy = None
        self.energy_input = None
        self.coord_pairs = []
        self.pb_last = {}

    def set_box(self, box):
        self.box = box9(box)

    def set_energy_input(self, U, F):
        """F: (N, 3) external force; the energy CV's Jacobian is -F."""
        self.energy_input = (float(U), np.asarray(F, dtype=np.float64))

    # -- CVs --
    def eval_cvs_f64(self, pos, exact_params=False):
        """CVs from FP64 positions (no FP32 rounding), for finite differences;
        exact_params also skips the float rounding of stored parameters."""
        return self.eval_cvs(pos, f64=True, exact_params=exact_params)

    def path_lambda_derivs(self, pos, spec_index, f64=True):
        """(s, z) and (ds/dlambda, dz/dlambda) of the spec_index-th CV (a path)."""
        t, atoms, params = self.force.cvs[spec_index]
        assert t == CV_PATH
        M, N = len(atoms), int(params[1])
        pos = np.asarray(pos, dtype=np.float64 if f64 else np.float32).astype(np.float64).reshape(-1)
        pos, pp = P(pos)
        a, ap = P(atoms, np.int32)
        fr, frp = P(np.array(params[2:2 + N * M * 3], dtype=np.float64))
        sz, gs, gz, dl = np.zeros(2), np.zeros(3 * M), np.zeros(3 * M), np.zeros(2)
        lib().orc_cv_path(pp, ap, M, frp, N, params[0], sz.ctypes.data_as(dp), gs.ctypes.data_as(dp),
                          gz.ctypes.data_as(dp), dl.ctypes.data_as(dp))
        return sz, dl

    def eval_cvs(self, pos, f64=False, exact_params=False):
        L = lib()
        f32 = (lambda x: float(x)) if exact_params else globals()["f32"]
        f32v = (lambda xs: np.asarray(xs, dtype=np.float64)) if exact_params else globals()["f32v"]
        pos = np.asarray(pos, dtype=np.float64 if f64 else np.float32).astype(np.float64).reshape(-1)
        pos, pp = P(pos)
        bx, bp = P(self.box)
        pbc = 1 if self.pbc else 0
        cv = np.zeros(self.nV)
        rows = []  # (value index, atoms array, grads (n,3))
        self.coord_pairs = []
        v = 0
        for t, atoms, params in self.force.cvs:
            if t == CV_DISTANCE:
                g = np.zeros(6)
                cv[v] = L.orc_cv_distance(pp, atoms[0], atoms[1], bp, pbc, g.ctypes.data_as(dp))
                rows.append((v, np.array(atoms[:2]), g.reshape(2, 3)))
                v += 1
            elif t == CV_ANGLE:
                a, ap = P(atoms[:3], np.int32)
                g = np.zeros(9)
                cv[v] = L.orc_cv_angle(pp, ap, bp, pbc, g.ctypes.data_as(dp))
                rows.append((v, np.array(atoms[:3]), g.reshape(3, 3)))
                v += 1
            elif t == CV_DIHEDRAL:
                a, ap = P(atoms[:4], np.int32)
                g = np.zeros(12)
                cv[v] = L.orc_cv_dihedral(pp, ap, bp, pbc, g.ctypes.data_as(dp))
                rows.append((v, np.array(atoms[:4]), g.reshape(4, 3)))
                v += 1
            elif t == CV_POSITION:
                c = int(params[0])
                cv[v] = pos[3 * atoms[0] + c]
                g = np.zeros((1, 3))
                g[0, c] = 1.0
                rows.append((v, np.array(atoms[:1]), g))
                v += 1
            elif t == CV_COM_DISTANCE:
                a, ap = P(atoms, np.int32)
                m, mp = P(self.masses[atoms])
                g = np.zeros(3 * len(atoms))
                n1 = int(round(params[0]))
                cv[v] = L.orc_cv_com_distance(pp, ap, n1, len(atoms), mp, bp, pbc, g.ctypes.data_as(dp))
                rows.append((v, np.array(atoms), g.reshape(-1, 3)))
                v += 1
            elif t == CV_GYRATION:
                a, ap = P(atoms, np.int32)
                m, mp = P(self.masses[atoms])
                g = np.zeros(3 * len(atoms))
                cv[v] = L.orc_cv_gyration(pp, ap, len(atoms), mp, bp, pbc, g.ctypes.data_as(dp))
                rows.append((v, np.array(atoms), g.reshape(-1, 3)))
                v += 1
            elif t == CV_COORDINATION:
                nA = atoms[0]
                A, Ap = P(atoms[1:1 + nA], np.int32)
                B, Bp = P(atoms[1 + nA:], np.int32)
                gA = np.zeros(3 * len(A))
                gB = np.zeros(3 * len(B))
                npairs = ctypes.c_longlong()
                cutoff = params[3] if len(params) > 3 else 0.0
                cv[v] = L.orc_cv_coordination(pp, Ap, len(A), Bp, len(B), f32(params[0]), f32(params[1]),
                                              f32(params[2]), cutoff, bp, pbc, gA.ctypes.data_as(dp),
                                              gB.ctypes.data_as(dp), ctypes.byref(npairs))
                self.coord_pairs.append(npairs.value)
                rows.append((v, np.concatenate([A, B]), np.concatenate([gA, gB]).reshape(-1, 3)))
                v += 1
            elif t == CV_RMSD:
                n = len(atoms)
                a, ap = P(atoms, np.int32)
                ref, rp = P(f32v(params[:3 * n]))
                fit = 1 if (len(params) > 3 * n and params[3 * n] != 0.0) else 0
                g = np.zeros(3 * n)
                cv[v] = L.orc_cv_rmsd(pp, ap, n, rp, fit, g.ctypes.data_as(dp), None)
                rows.append((v, np.array(atoms), g.reshape(-1, 3)))
                v += 1
            elif t == CV_PATH:
                M, N = len(atoms), int(params[1])
                a, ap = P(atoms, np.int32)
                fr, frp = P(f32v(params[2:2 + N * M * 3]))
                sz = np.zeros(2)
                gs = np.zeros(3 * M)
                gz = np.zeros(3 * M)
                L.orc_cv_path(pp, ap, M, frp, N, f32(params[0]), sz.ctypes.data_as(dp), gs.ctypes.data_as(dp),
                              gz.ctypes.data_as(dp), None)
                cv[v], cv[v + 1] = sz
                rows.append((v, np.array(atoms), gs.reshape(-1, 3)))
                rows.append((v + 1, np.array(atoms), gz.reshape(-1, 3)))
                v += 2
            elif t == CV_DRMSD:
                np_ = len(atoms) // 2
                a, ap = P(atoms[:2 * np_], np.int32)
                d0, d0p = P(f32v(params[:np_]))
                g = np.zeros(6 * np_)
                cv[v] = L.orc_cv_drmsd(pp, ap, np_, d0p, bp, pbc, g.ctypes.data_as(dp))
                rows.append((v, np.array(atoms[:2 * np_]), g.reshape(-1, 3)))
                v += 1
            elif t == CV_CONTACTMAP:
                np_ = len(atoms) // 2
                a, ap = P(atoms[:2 * np_], np.int32)
                pr, prp = P(f32v(params[:4 * np_]))
                g = np.zeros(6 * np_)
                cv[v] = L.orc_cv_contactmap(pp, ap, np_, prp, bp, pbc, g.ctypes.data_as(dp))
                rows.append((v, np.array(atoms[:2 * np_]), g.reshape(-1, 3)))
                v += 1
            elif t == CV_ENERGY:
                if self.energy_input is None:
                    raise ValueError("oracle: CV_ENERGY needs set_energy_input")
                U, F = self.energy_input
                cv[v] = U
                rows.append((v, np.arange(self.N), -F[:self.N]))
                v += 1
            else:
                raise ValueError(f"oracle: CV type {t} not covered")
        return cv, rows

    # -- biases --
    def eval_biases(self, cv, welford=True):
        nB = len(self.force.biases)
        be = np.zeros(nB)
        dU = np.zeros(self.nV)
        for b, (t, cvs, params, iparams) in enumerate(self.force.biases):
            st = self.states[b]
            s = cv[cvs]
            if t == BIAS_OPES_MULTITHERMAL:
                V, g = st.eval(s[0])
            else:
                if t == BIAS_OPES and st.adaptiveStride > 0 and welford:
                    st.welford(s)
                V, g = st.eval(s)
                if t == BIAS_PBMETAD:
                    self.pb_last[b] = (st.Vi.copy(), st.wi.copy())
            be[b] = V
            for d, c in enumerate(cvs):
                dU[c] += g[d]
        return be, dU

    def execute(self, pos):
        cv, rows = self.eval_cvs(pos)
        if self.force._test_grads is not None:
            be = np.zeros(len(self.force.biases))
            dU = np.array(self.force._test_grads, dtype=np.float64)
        else:
            be, dU = self.eval_biases(cv)
        F = np.zeros((self.N, 3))
        for v, atoms, g in rows:
            np.add.at(F, atoms, -dU[v] * g)
        self.cv, self.benergy, self.dU, self.rows = cv, be, dU, rows
        self.cv_ready = True
        return {"cv": cv, "bias_energies": be, "energy": float(be.sum()), "dUdCV": dU, "forces": F}

    def forces_fixed(self, padded=None):
        """Forces through the reference scatter: FP32 rows, per-row truncation."""
        padded = padded or self.N
        atom, gx, gy, gz, cvi = [], [], [], [], []
        for v, atoms, g in self.rows:
            atom.append(np.asarray(atoms, dtype=np.int32))
            g32 = g.astype(np.float32)
            gx.append(g32[:, 0])
            gy.append(g32[:, 1])
            gz.append(g32[:, 2])
            cvi.append(np.full(len(atoms), v, dtype=np.int32))
        out = np.zeros(3 * padded, dtype=np.int64)
        if atom:
            a, ap = P(np.concatenate(atom), np.int32)
            x, xp = P(np.concatenate(gx), np.float32)
            y, yp = P(np.concatenate(gy), np.float32)
            z, zp = P(np.concatenate(gz), np.float32)
            c, cp = P(np.concatenate(cvi), np.int32)
            u, up = P(self.dU)
            lib().orc_scatter_fixed(len(a), ap, xp, yp, zp, cp, up, out.ctypes.data_as(llp), padded)
        return out.reshape(3, padded)[:, :self.N].T.copy()

    def update_state(self, step):
        if not self.cv_ready or step == self.last_update:
            return
        self.last_update = step
        if step == 0:
            return
        biases = self.force.biases
        order = ([BIAS_OPES], [BIAS_OPES_EXPANDED], [BIAS_OPES_MULTITHERMAL], [BIAS_METAD], [BIAS_PBMETAD],
                 [BIAS_METAD_HILLS])
        for kinds in order:
            for b, (t, cvs, params, iparams) in enumerate(biases):
                if t not in kinds:
                    continue
                st = self.states[b]
                if step % st.pace != 0:
                    continue
                s = self.cv[cvs]
                if t == BIAS_OPES:
                    if st.depositMirror >= st.cap:
                        continue
                    will = st.adaptiveStride == 0 or st.st_i[3] >= st.adaptiveStride
                    st.deposit(s, self.benergy[b])
                    if will:
                        st.depositMirror += 1
                elif t == BIAS_OPES_EXPANDED:
                    st.update(s)
                elif t == BIAS_OPES_MULTITHERMAL:
                    st.update(s[0], self.benergy[b])
                elif t == BIAS_PBMETAD:  # this walker's V_i, w_i of its last evaluation
                    st.deposit(s, *self.pb_last[b])
                else:
                    st.deposit(s, self.benergy[b])

    def md_step(self, pos, step):
        self.update_state(step)
        return self.execute(pos)
