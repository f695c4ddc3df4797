#!/usr/bin/env python3
"""Benchmark of the GLUED enhanced-sampling step on one MI355X.

Default workload (the headline line): the batched multi-walker
configuration -- 64 walkers of uneven size (2,000-40,000 atoms at 100
atoms/nm^3), Kabsch RMSD over 500 atoms + 2 torsions, well-tempered MetaD on
(RMSD, phi) with a shared 200x200 grid and 20,000 private hills per walker.
One eval = every CV, the bias energy and dV/ds, and the fixed-point force
scatter, for one walker.

  value        batched evals/s: every walker per step, the stage kernels
               launched directly on the stream (on this ROCm a captured
               hipGraph replay of the same kernels costs ~4 us more per step;
               its rate is reported as graph_evals_per_s)
  plain        the reference's structure: one walker per launch sequence
               (one launch per CV kind, per bias, and the scatter), host loop
               over walkers
  roofline     the dominant kernel of the workload, timed live with HIP events
               on the launch stream over back-to-back launches of that stage;
               traffic = HBM bytes per launch from rocprofv3 PMC counters
  cpu_baseline the FP64 oracle on a bounded sample of the same walkers, one
               host core; its forces also give the force error vs FP64

--workload picks one of the other seeded workloads (each prints its own line):
  mixed-large        16 walkers x 100,000 atoms: contact map (4,096 pairs),
                     DRMSD (1,500 pairs), path CV (256 frames x 1,000 atoms),
                     multithermal on a dense energy CV (roofline: the dense
                     energy-CV force scatter)
  coordination-opes  50,000 atoms: coordination 2,000 x 20,000 (1.2 nm cutoff,
                     cell list) + Rg(5,000), OPES-METAD with 5,000 kernels
  path-batch         256 configurations of 300 atoms as walkers: path CV over
                     64 frames, MetaD with 20,000 shared hills
  small              22 atoms, distance + 2 torsions, WT MetaD hill list with
                     1,000 prior hills, one deposit per frame

Prints one JSON line.  Only --gpus 1 is supported (one GPU per box).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VALU_TFLOPS = 78.6
FP32_VALU_TFLOPS = 157.3
FIXED_SCALE = 4294967296.0

# HBM traffic per launch of each kernel of a workload's batched step, from two
# rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch,
# averaged over the batched steps of `bench.py --workload W --batched-only
# --steps 20 --warmup 2`).  On gfx950 FETCH_SIZE reads half the bytes that a
# copy kernel of known size moves and WRITE_SIZE reads them in full, so
# traffic = 2 * FETCH + 1 * WRITE.  The per-kernel figures and the commands
# are in the workload summaries under profiles/.
PMC_FETCH_SCALE, PMC_WRITE_SCALE = 2.0, 1.0
PMC_KIB = {
    "multiwalker": {"k_cv_fused": (442.8, 512.5), "k_step_bias_chunks<1, 2>": (7583.6, 60.0),
                    "k_step_bias_finalize<false>": (535.3, 769.0)},
    # k_path_rows<8, true> carries the DRMSD(1,500 pairs) workgroups and the contact-map sums
    "mixed-large": {"k_cv_fused": (11525.5, 4417.7), "k_path_rows<8, true>": (11817.1, 1865.8),
                    "k_step_bias_finalize<false>": (93.0, 2.0), "k_scatter_dense_rows": (31155.6, 37512.5)},
    # k_coord_rank carries the Rg(5,000) workgroup
    "coordination-opes": {"k_coord_rank": (387.9, 948.7), "k_coord_place": (450.9, 420.6),
                          "k_coord_cells<512>": (1596.7, 709.4), "k_step_bias_chunks<2, 2>": (150.1, 3.0),
                          "k_step_finalize_scatter<false>": (470.6, 516.4)},
    "path-batch": {"k_cv_fused": (3476.3, 770.7), "k_path_rows<4, false>": (3484.8, 2544.0),
                   "k_step_hill_chunks_shared<2>": (313.2, 384.0), "k_step_bias_finalize<false>": (1769.5, 1823.3)},
    "small": {"k_step_tiny<1, 2, false>": (43.1, 0.9)},
}
# the kernel(s) each workload's roofline line is about
ROOFLINE_KERNELS = {
    "multiwalker": ["k_step_bias_chunks<1, 2>"],
    "mixed-large": ["k_scatter_dense_rows"],
    "coordination-opes": ["k_coord_rank", "k_coord_place", "k_coord_cells<512>"],
    "path-batch": ["k_cv_fused", "k_path_rows<4, false>"],
    "small": ["k_step_tiny<1, 2, false>"],
}


def pmc_traffic(workload, kernels):
    t = PMC_KIB[workload]
    return round(sum(PMC_FETCH_SCALE * t[k][0] + PMC_WRITE_SCALE * t[k][1] for k in kernels) * 1024)


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads: each builds a Context and says how to time and check it
# ---------------------------------------------------------------------------
class Workload:
    name = ""
    walkers = 1

    def dominant(self):
        """One launch of the roofline kernel(s)."""
        self.ctx.execute_stage(self.stage)

    def time_dominant(self, reps):
        """Average duration of one launch of the roofline kernel(s) (s)."""
        return time_fn(self.dominant, reps)

    def roofline(self, t_launch):
        raise NotImplementedError

    def oracle_walker(self, w):
        """An OracleContext holding walker w's current state (CPU checker)."""
        raise NotImplementedError

    def positions(self, w):
        return self.ctx.posq[self.ctx.walker_slice(w)][:self.ctx.natoms[w], :3].cpu().numpy()


class MultiWalker(Workload):
    name = "multiwalker"
    stage = 3

    def __init__(self, args):
        from gluedmd_amd import workloads as wl
        from gluedmd_amd.context import Context
        self.W = args.walkers
        self.H = args.hills
        c = wl.config4(n_walkers=self.W, n_hills=self.H, extra_hills=64)
        self.c, f = c, c["force"]
        self.force = f
        ctx = Context(f, walkers=c["sizes"])
        ctx.set_grid(c["grid_bias"], wl.grid_prefill(f, c["grid_bias"], c["grid_prior"]))
        for w in range(self.W):
            ctx.set_box(c["boxes"][w], w)
            ctx.set_positions(c["positions"][w], w)
            ctx.set_hills(c["hill_bias"], c["hills"][w][0], c["hills"][w][1], w)
        self.ctx = ctx
        self.config = {
            "workload": f"batched multi-walker: {self.W} walkers, 2k-40k atoms, Kabsch RMSD(500)+2 torsions, "
                        f"WT MetaD shared 200x200 grid + {self.H // 1000}k private hills/walker",
            "walkers": self.W,
            "hills_per_walker": self.H,
            "parallelism": "batched walkers, 3 kernel launches per step (CVs, hill-sum chunks, finalize + scatter)",
        }

    def roofline(self, t_launch):
        hb = self.c["hill_bias"]
        n = sum(len(self.ctx.get_hills(hb, w)[1]) for w in range(self.W))
        D = 2
        nbytes = n * 4 * (D + 1)  # FP32 centre (D) + height per hill, streamed once
        return dict(bound="hbm", kernel="k_step_bias_chunks (hill sums, D=2)", achieved=nbytes / t_launch / 1e9,
                    peak=HBM_PEAK_GBS, unit="GB/s", per_unit="12 B per hill (2 FP32 centres + FP32 height)",
                    units_per_launch=n, bytes_per_launch=nbytes)

    def oracle_walker(self, w):
        import numpy as np
        from oracle.oracle import OracleContext, _Grid
        f, c = self.force, self.c
        gb, hb = c["grid_bias"], c["hill_bias"]
        if not hasattr(self, "_grid"):
            t, cvs, params, ip = f.biases[gb]
            self._grid = _Grid(len(cvs), params, ip)
            blob = self.ctx.getBiasState(0)
            off = 8 + 4 + 4 + 4 + 4  # 'GPUS', version, nOpes=0, nAbmd=0, nMetaD, nDeposited
            self._grid.grid = np.frombuffer(blob[off:off + 8 * self._grid.total], dtype="<f8").copy()
        o = OracleContext(f, c["sizes"][w], shared={gb: self._grid})
        o.set_box(c["boxes"][w])
        hc, hh = self.ctx.get_hills(hb, w)
        o.states[hb].centers, o.states[hb].heights = hc, hh
        return o


class MixedLarge(Workload):
    name = "mixed-large"
    stage = 4

    def __init__(self, args):
        import torch
        from gluedmd_amd import workloads as wl
        from gluedmd_amd.context import Context
        W = args.walkers if args.walkers != 64 else 16
        c = wl.config5(n_walkers=W)
        self.c, self.force, self.W = c, c["force"], W
        ctx = Context(c["force"], walkers=[c["natoms"]] * W)
        ctx.set_box(c["box"])
        self._fext = []
        for w in range(W):
            ctx.set_positions(c["positions"][w], w)
            p = ctx.padded[w]
            Fe = torch.zeros(3 * p, dtype=torch.float32, device=ctx.device)
            Fe.view(3, p)[:, :c["natoms"]] = torch.as_tensor(c["F"][w].T.copy(), device=ctx.device)
            ctx.set_energy_input(float(c["U"][w]), Fe, w)
            self._fext.append(Fe)
        self.ctx = ctx
        self.config = {
            "workload": f"mixed large: {W} walkers x 100k atoms, contact map 4096 pairs, DRMSD 1500 pairs, "
                        "path 256 frames x 1000 atoms, multithermal 32 T on a dense energy CV",
            "walkers": W,
            "parallelism": "batched walkers, direct kernel launches",
        }

    def roofline(self, t_launch):
        n = self.W * self.c["natoms"]
        rows, _ = self.ctx.plan_info()
        nr = self.W * rows
        nbytes = n * (12 + 48) + nr * 12  # F_ext + fixed-point force RMW per atom; FP32 xyz per Jacobian row
        return dict(bound="hbm", kernel="k_scatter_dense_rows (energy-CV + atom-slot force scatter, one pass)",
                    achieved=nbytes / t_launch / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    per_unit="60 B per atom (12 B FP32 F_ext + 48 B fixed-point force read+write) "
                             "+ 12 B per Jacobian row",
                    units_per_launch={"atoms": n, "rows": nr}, bytes_per_launch=nbytes)

    def oracle_walker(self, w):
        import numpy as np
        from oracle.oracle import OracleContext
        c = self.c
        o = OracleContext(self.force, c["natoms"])
        o.set_box(c["box"])
        o.set_energy_input(float(c["U"][w]), c["F"][w].astype(np.float64))
        return o


class CoordinationOpes(Workload):
    name = "coordination-opes"
    stage = 0

    def __init__(self, args):
        from gluedmd_amd import workloads as wl
        from gluedmd_amd.context import Context
        c = wl.config2_system()
        self.c, self.force, self.W = c, c["force"], 1
        ctx = Context(c["force"], c["natoms"])
        ctx.set_box(c["box"])
        ctx.set_positions(c["positions"])
        ctx.execute_batch(use_graph=False)  # the initial (CN, Rg), around which the prior kernels are drawn
        r0 = ctx.getLastCVValues(0)[:2]
        self.kernels = wl.opes_prior_kernels(c["rng"], r0, [0.05, 0.02], c["gamma"], n=5000)
        self.su = wl.sum_uprob_np(*self.kernels, wl.opes_metad_cutoff2(c["gamma"]))
        ctx.setBiasState(wl.opes_state_blob(*self.kernels, self.su))
        self.ctx = ctx
        self.config = {
            "workload": "coordination + OPES: 50k atoms, CN 2000 x 20000 (cutoff 1.2 nm, cell list), Rg(5000), "
                        "OPES-METAD 5000 kernels",
            "walkers": 1,
            "parallelism": "one walker, direct kernel launches",
        }

    def roofline(self, t_launch):
        pairs = self.ctx.coordination_pairs(0)
        flops = 45.0 * pairs  # PBC distance (21) + switching value and gradient (24) per pair inside the cutoff
        return dict(bound="fp64-valu", kernel="CV stage (coordination cell-list pair kernels + Rg)",
                    achieved=flops / t_launch / 1e12, peak=FP64_VALU_TFLOPS, unit="TFLOP/s",
                    per_unit="45 flop per pair inside the cutoff", units_per_launch=pairs, flops_per_launch=flops)

    def oracle_walker(self, w):
        from oracle.oracle import OracleContext
        o = OracleContext(self.force, self.c["natoms"])
        o.set_box(self.c["box"])
        o.states[0].load(*self.kernels, self.su)
        return o


class PathBatch(Workload):
    name = "path-batch"
    stage = 0

    def __init__(self, args):
        import numpy as np
        from gluedmd_amd import workloads as wl
        from gluedmd_amd.context import Context
        from gluedmd_amd.glued import Force
        c = wl.config3_system()
        B = len(c["confs"])
        # hill centres over the (s, z) box of the configurations: take it from
        # one GPU evaluation of the path CV alone
        f0 = Force()
        f0.add_path(list(range(c["M"])), [fr.reshape(-1).tolist() for fr in c["frames"]], c["lam"])
        g0 = Context(f0, walkers=[c["M"]] * B)
        for b in range(B):
            g0.set_positions(c["confs"][b], b)
        g0.execute_batch(use_graph=False)
        sz = np.array([g0.getLastCVValues(b) for b in range(B)])
        del g0
        hc, hh, sig = wl.config3_hills(c["rng"], (sz[:, 0].min(), sz[:, 0].max()), (sz[:, 1].min(), sz[:, 1].max()),
                                       n_hills=20000)
        f = Force(temperature=300.0)
        s = f.add_path(list(range(c["M"])), [fr.reshape(-1).tolist() for fr in c["frames"]], c["lam"])
        hb = f.add_metad_hills([s, s + 1], sigma=sig.tolist(), height=1.2, pace=1, gamma=None,
                               max_hills=len(hh) + B * 48)
        f.share_bias(hb)
        ctx = Context(f, walkers=[c["M"]] * B)
        ctx.set_hills(hb, hc, hh)
        for b in range(B):
            ctx.set_positions(c["confs"][b], b)
        self.c, self.force, self.W, self.ctx, self.hb = c, f, B, ctx, hb
        self.hills = (hc, hh)
        self.config = {
            "workload": "path batch: 256 configurations x 300 atoms, path CV 64 frames (lambda 50), "
                        "MetaD 20k hills shared",
            "walkers": B,
            "parallelism": "configurations as walkers, direct kernel launches",
        }

    def roofline(self, t_launch):
        c = self.c
        # per frame-atom: the MSD cross term x'.r' (3 multiply-adds, FP64 matrix
        # cores) and the row pass (x - r, then two weighted accumulations, FP64)
        flops = self.W * c["N"] * c["M"] * (6 + 3 + 12)
        return dict(bound="fp64 (matrix cores + VALU)",
                    kernel="CV stage (path CV, 256 configurations x 64 frames x 300 atoms)",
                    achieved=flops / t_launch / 1e12, peak=FP64_VALU_TFLOPS, unit="TFLOP/s",
                    per_unit="21 flop per frame-atom (MSD cross term 6 on v_mfma_f64_16x16x4f64, row pass 15)",
                    units_per_launch=self.W * c["N"] * c["M"], flops_per_launch=flops)

    def oracle_walker(self, w):
        from oracle.oracle import OracleContext
        o = OracleContext(self.force, self.c["M"])
        o.states[self.hb].centers, o.states[self.hb].heights = self.ctx.get_hills(self.hb, 0)
        return o


class Small(Workload):
    name = "small"
    stage = 3

    def __init__(self, args):
        from gluedmd_amd import workloads as wl
        from gluedmd_amd.context import Context
        c = wl.config1(n_frames=max(1000, args.steps + args.warmup + 64))
        ctx = Context(c["force"], c["natoms"])
        ctx.set_box(c["box"])
        ctx.set_hills(0, c["hill_centers"], c["hill_heights"])
        ctx.set_positions(c["frames"][0])
        self.c, self.force, self.W, self.ctx = c, c["force"], 1, ctx
        self.config = {
            "workload": "small: 22 atoms, distance + 2 torsions, WT MetaD hill list (1000 prior hills, "
                        "1 deposit per frame)",
            "walkers": 1,
            "parallelism": "one walker, direct kernel launches",
        }

    def dominant(self):
        """The batched step of this plan is one launch (k_step_tiny)."""
        self.ctx.execute_batch(use_graph=False)

    def roofline(self, t_launch):
        n = len(self.ctx.get_hills(0)[1])
        # 7 atoms: 12 B position + 48 B fixed-point force read+write; 12 B per
        # hill (FP32 centres and height); 24 B of CV values and bias state
        nbytes = 7 * (12 + 48) + n * 12 + 24
        return dict(bound="hbm", kernel="k_step_tiny (whole step: CVs, hill sum, bias, scatter in one launch)",
                    achieved=nbytes / t_launch / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                    per_unit="12 B per hill + 420 B per step (7 atoms x 60 B) + 24 B", units_per_launch=n,
                    bytes_per_launch=nbytes)

    def oracle_walker(self, w):
        from oracle.oracle import OracleContext
        o = OracleContext(self.force, self.c["natoms"])
        o.set_box(self.c["box"])
        o.states[0].centers, o.states[0].heights = self.ctx.get_hills(0)
        return o


WORKLOADS = {k.name: k for k in (MultiWalker, MixedLarge, CoordinationOpes, PathBatch, Small)}


# ---------------------------------------------------------------------------
# timing
# ---------------------------------------------------------------------------
def time_loop(fn, steps, sync):
    sync()
    t0 = time.perf_counter()
    for k in range(steps):
        fn(k)
    sync()
    return (time.perf_counter() - t0) / steps


def time_fn(fn, reps):
    """Average GPU duration of one call: HIP events on the stream the engine
    launches on (torch's current stream, which the Context hands to every
    C-ABI call), around `reps` back-to-back calls."""
    import torch
    stream = torch.cuda.current_stream()
    for _ in range(10):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def time_cold(fn, reps, flush_mib=512, write=False):
    """Average GPU duration of one call with nothing of its working set in the
    caches: before each call a 512 MiB buffer (more than the 256 MB
    memory-side cache) is read on the same stream, then the call is timed
    alone between two HIP events.  The flush reads, so the caches it leaves
    hold clean lines; write=True overwrites the buffer instead, and the call
    then also pays for writing back the flush's dirty lines it evicts.
    flush_mib=0: the same single-call timing without the flush."""
    import torch
    stream = torch.cuda.current_stream()
    buf = torch.ones(max(flush_mib, 1) * 2**18, dtype=torch.float32, device="cuda")
    sink = torch.zeros(reps, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    fn()
    for k in range(reps):
        if flush_mib and write:
            buf.fill_(float(k))
        elif flush_mib:
            torch.sum(buf, dim=0, out=sink[k])
        ev[k][0].record(stream)
        fn()
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    del buf
    return sum(a.elapsed_time(b) for a, b in ev) / reps * 1e-3


def time_stage(ctx, stage, reps):
    """Average duration of one launch of a stage."""
    return time_fn(lambda: ctx.execute_stage(stage), reps)


def host_info():
    """The host the CPU baseline ran on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0))}


def _cpu_worker(items, reps, start_at):
    """Process of the CPU baseline's parallel leg (started with the spawn
    method: a fresh interpreter that imports no GPU library): evaluate the
    given oracle contexts `reps` times from `start_at` on."""
    while time.time() < start_at:
        time.sleep(0.001)
    t0 = time.time()
    n = 0
    for _ in range(reps):
        for o, p in items:
            o.execute(p)
            n += 1
    return n, t0, time.time()


def cpu_baseline(wk, seconds, procs):
    """FP64 oracle on the host.

    The oracle contexts of walkers 0..n-1 (n = max(P, min(W, 4P))) are built
    first; their first evaluation checks the GPU's forces (untimed).  Leg 1,
    one thread, about seconds/2: the contexts evaluated in turn, over and
    over.  Leg 2, `procs` processes (one per core, each single-threaded), as
    many passes as leg 1's rate fills in about seconds/2: context i in
    process i mod P; the rate is all processes' evals over the span from the
    first start to the last end.  `value` is leg 2's rate."""
    import concurrent.futures
    import multiprocessing
    import numpy as np
    import torch
    ctx = wk.ctx
    ctx.zero_forces()
    ctx.execute_batch(use_graph=False)
    torch.cuda.synchronize()
    P = max(1, procs)
    items = []
    num = den = amax = cverr = 0.0
    for i in range(max(P, min(wk.W, 4 * P))):
        w = i % wk.W
        o, pos = wk.oracle_walker(w), wk.positions(w)
        if i < wk.W:
            r = o.execute(pos)
            Fg = ctx.get_forces(w)
            num += float(np.sum((Fg - r["forces"]) ** 2))
            den += float(np.sum(r["forces"] ** 2))
            amax = max(amax, float(np.abs(Fg).max()))
            cv = ctx.getLastCVValues(w)
            cverr = max(cverr, float(np.max(np.abs(cv - r["cv"]) / np.maximum(np.abs(r["cv"]), 1e-12))))
        items.append((o, pos))
    err = {"forces_rel_rms": float(math.sqrt(num / den)) if den > 0 else None, "cv_max_rel": cverr,
           "walkers": min(len(items), wk.W)}
    if den == 0:
        err["note"] = f"FP64 forces all zero (no bias acts on the atoms yet); largest GPU force {amax}"
    half = seconds / 2
    n1, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < half:
        o, pos = items[n1 % len(items)]
        o.execute(pos)
        n1 += 1
    rate1 = n1 / (time.perf_counter() - t0)
    one = {"value": round(rate1, 2), "unit": "evals/s", "cores": 1, "kind": "port",
           "sample": f"{n1} walker evals cycling over the oracle contexts of walkers 0..{min(len(items), wk.W) - 1} "
                     "of the same workload, FP64 C oracle (gcc -O3 -march=x86-64-v3) driven from Python, one thread"}
    if P == 1:
        one["host"] = host_info()
        return one, err
    work = [items[i::P] for i in range(P)]
    reps = max(1, int(math.ceil(rate1 * half / max(len(work[0]), 1))))
    saved = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in saved:
        os.environ[k] = "1"
    try:
        mp = multiprocessing.get_context("spawn")
        with concurrent.futures.ProcessPoolExecutor(P, mp_context=mp) as ex:
            start_at = time.time() + 8.0
            res = [f.result() for f in [ex.submit(_cpu_worker, work[i], reps, start_at) for i in range(P)]]
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    n_par = sum(r[0] for r in res)
    span = max(r[2] for r in res) - min(r[1] for r in res)
    out = {"value": round(n_par / span, 2), "unit": "evals/s", "cores": P, "kind": "port",
           "sample": f"{n_par} walker evals over {P} single-threaded processes ({len(items)} oracle contexts of "
                     f"walkers 0..{min(len(items), wk.W) - 1}, {reps} passes each) of the same workload, "
                     "FP64 C oracle (gcc -O3 -march=x86-64-v3) driven from Python",
           "single_thread": one, "host": host_info()}
    return out, err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--workload", default="multiwalker", choices=sorted(WORKLOADS))
    ap.add_argument("--walkers", type=int, default=64, help="multiwalker: walker count")
    ap.add_argument("--hills", type=int, default=20000, help="multiwalker: private hills per walker")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-procs", type=int, default=min(16, len(os.sched_getaffinity(0))),
                    help="processes of the CPU baseline's parallel leg (the GPU box's CPU share is 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="time only the batched step (for profiling its kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(walkers would shard across GPUs, one process per GPU; not built here)", file=sys.stderr)
        return 2

    import torch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    wk = WORKLOADS[args.workload](args)
    ctx, W = wk.ctx, wk.W
    sync = torch.cuda.synchronize
    log(f"{wk.name} ready ({W} walkers)")

    # ---- batched: every walker per step, direct launches ----
    for _ in range(args.warmup):
        ctx.execute_batch(use_graph=False)
    t_batched = time_loop(lambda k: ctx.execute_batch(use_graph=False), args.steps, sync)
    log(f"batched {t_batched * 1e6:.1f} us/step; launch shapes: {ctx.tuning()}")
    if args.batched_only:
        print(json.dumps({"metric": "batched step only", "workload": wk.name, "ms_per_step": round(t_batched * 1e3, 4),
                          "value": round(W / t_batched, 1), "unit": "evals/s", "tuning": ctx.tuning()}))
        return 0

    # ---- the same batched step replayed as one captured hipGraph (timed
    # right after the direct launches, with the same warm-up) ----
    for _ in range(args.warmup):
        ctx.execute_batch(use_graph=True)
    t_graph = time_loop(lambda k: ctx.execute_batch(use_graph=True), args.steps, sync)
    log(f"graph {t_graph * 1e6:.1f} us/step")

    # ---- roofline of the dominant kernel: back to back (its working set
    # stays in the caches across launches, as across the timed steps) and
    # with the caches flushed before every launch ----
    t_launch = wk.time_dominant(200)
    t_cold = time_cold(wk.dominant, 50)
    t_cold_w = time_cold(wk.dominant, 50, write=True)
    t_alone = time_cold(wk.dominant, 50, flush_mib=0)
    rf = wk.roofline(t_launch)
    rf_cold = wk.roofline(t_cold)
    rf["frac"] = round(rf["achieved"] / rf["peak"], 4)
    rf["achieved"] = round(rf["achieved"], 3)
    rf["avg_launch_us"] = round(t_launch * 1e6, 3)
    rf["cold"] = {"achieved": round(rf_cold["achieved"], 3), "frac": round(rf_cold["achieved"] / rf["peak"], 4),
                  "avg_launch_us": round(t_cold * 1e6, 3), "warm_alone_us": round(t_alone * 1e6, 3),
                  "after_writes_us": round(t_cold_w * 1e6, 3),
                  "how": "512 MiB read on the stream before each of 50 launches, each launch timed alone "
                         "(warm_alone_us: the same single-launch timing without the flush; after_writes_us: "
                         "the flush overwrites the 512 MiB instead, and the launch also writes back the dirty "
                         "lines it evicts)"}
    rf["traffic"] = pmc_traffic(wk.name, ROOFLINE_KERNELS[wk.name])
    rf["traffic_kernels"] = ROOFLINE_KERNELS[wk.name]
    stage_t = {}
    for st, name in ((0, "cv"), (1, "bias"), (2, "scatter")):
        stage_t[name] = round(time_stage(ctx, st, 100) * 1e6, 2)
    step_traffic = pmc_traffic(wk.name, sorted(PMC_KIB[wk.name]))
    step = {"us": round(t_batched * 1e6, 3), "kernels": sorted(PMC_KIB[wk.name]), "traffic": step_traffic,
            "traffic_GBps": round(step_traffic / t_batched / 1e9, 1),
            "hbm_frac": round(step_traffic / t_batched / 1e9 / HBM_PEAK_GBS, 4)}

    # ---- eval + deposit (pace 1: every walker deposits every step), from a
    # snapshot of the bias state that is restored afterwards ----
    snap = [ctx.getBiasState(w) for w in range(W)]
    ctx.update_state(1)
    ctx.execute_batch(use_graph=False)
    # (20 steps per form: the smallest hill capacity, path-batch's, holds 48 steps of 256 deposits)
    t_dep = time_loop(lambda k: (ctx.update_state(k + 2), ctx.execute_batch(use_graph=False)), 20, sync)
    # as one graph per deposit pattern (glued_step_batch)
    ctx.step_batch(100)
    t_dep_graph = time_loop(lambda k: ctx.step_batch(k + 101), 20, sync)
    for w in range(W):
        ctx.setBiasState(snap[w], w)
    ctx.execute_batch(use_graph=False)
    sync()

    # ---- plain: one walker per launch sequence ----
    for _ in range(2):
        ctx.execute_all_plain()
    t_plain = time_loop(lambda k: ctx.execute_all_plain(), args.plain_steps, sync)
    log(f"plain {t_plain * 1e6:.1f} us/step")

    # ---- CPU baseline and force check, on the restored pre-deposit state
    # (last, so the host legs do not overlap the GPU timings) ----
    cpu = err = None
    if not args.no_cpu_baseline:
        cpu, err = cpu_baseline(wk, args.cpu_seconds, args.cpu_procs)
        log(f"cpu baseline {cpu['value']} evals/s ({cpu['cores']} cores), error vs FP64 {err}")

    out = {
        "metric": "CV+bias+force evals/sec, batched vs plain on 1 GPU; force RMS err vs FP64",
        "value": round(W / t_batched, 1),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batched * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded SplitMix64 inputs)",
        "config": wk.config,
        "plain_evals_per_s": round(W / t_plain, 1),
        "batched_over_plain": round(t_plain / t_batched, 2),
        "graph_evals_per_s": round(W / t_graph, 1),
        "eval_plus_deposit_evals_per_s": round(W / t_dep, 1),
        "eval_plus_deposit_graph_evals_per_s": round(W / t_dep_graph, 1),
        "stage_us": stage_t,
        "tuning": ctx.tuning(),
        "roofline": rf,
        "step": step,
    }
    if cpu is not None:
        out["cpu_baseline"], out["error_vs_fp64"] = cpu, err
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
