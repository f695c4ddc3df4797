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

# HBM traffic per launch of each workload's roofline kernel(s), from two
# rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch,
# averaged over 20 batched steps of `bench.py --workload W --batched-only
# --steps 20 --warmup 2`).  On gfx950 FETCH_SIZE reads half the bytes that a
# copy kernel of known size moves and WRITE_SIZE reads them in full, so
# traffic = 2 * FETCH + 1 * WRITE.  The per-kernel figures are in the
# workload summaries under profiles/.
PMC_FETCH_SCALE, PMC_WRITE_SCALE = 2.0, 1.0
PMC_KIB = {
    "multiwalker": {"k_step_bias_chunks<1, 2>": (15104.7, 60.0)},
    "mixed-large": {"k_scatter_dense_rows": (41545.6, 49685.4)},
    "coordination-opes": {"k_coord_rank": (278.4, 792.4), "k_coord_offsets": (20.8, 38.8),
                          "k_coord_place": (357.3, 418.8), "k_coord_pairs_a": (1562.8, 250.0),
                          "k_coord_pairs<false, true>": (835.8, 590.3), "k_cv_group_big": (84.2, 156.3)},
    "path-batch": {"k_cv_fused (path MSD blocks)": (3377.2, 128.0), "k_path_rows<4>": (2116.1, 2416.0)},
    "small": {"k_step_tiny<1, 2, false>": (54.9, 0.9)},
}


def pmc_traffic(workload):
    t = PMC_KIB.get(workload)
    if not t:
        return None
    return round(sum(PMC_FETCH_SCALE * f + PMC_WRITE_SCALE * w for f, w in t.values()) * 1024)


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads: each builds a Context and says how to time and check it
# ---------------------------------------------------------------------------
class Workload:
    name = ""
    walkers = 1

    def time_dominant(self, reps):
        """Average duration of one launch of the roofline kernel(s) (s)."""
        return time_stage(self.ctx, self.stage, reps)

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
                               max_hills=len(hh) + B * 32)
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
        flops = self.W * c["N"] * c["M"] * (3 * 3 + 3 * 4)  # MSD pass (sub, fma) + gradient pass per frame-atom
        return dict(bound="fp32-valu", kernel="CV stage (path CV, 256 configurations x 64 frames x 300 atoms)",
                    achieved=flops / t_launch / 1e12, peak=FP32_VALU_TFLOPS, unit="TFLOP/s",
                    per_unit="21 flop per frame-atom (MSD 9, row gradient 12)",
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

    def time_dominant(self, reps):
        """The batched step of this plan is one launch (k_step_tiny): its duration."""
        return time_fn(lambda: self.ctx.execute_batch(use_graph=False), reps)

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


def time_stage(ctx, stage, reps):
    """Average duration of one launch of a stage."""
    return time_fn(lambda: ctx.execute_stage(stage), reps)


def cpu_baseline(wk, seconds):
    """FP64 oracle on the host: one core, walkers evaluated in turn until the
    time budget is used; also the checker of the GPU forces for those walkers."""
    import numpy as np
    import torch
    ctx = wk.ctx
    ctx.zero_forces()
    ctx.execute_batch(use_graph=False)
    torch.cuda.synchronize()
    evals, t_used, num, den, amax, w = 0, 0.0, 0.0, 0.0, 0.0, 0
    while t_used < seconds and evals < 1000 * wk.W:
        o = wk.oracle_walker(w)
        pos = wk.positions(w)
        t0 = time.perf_counter()
        r = o.execute(pos)
        t_used += time.perf_counter() - t0
        evals += 1
        Fg = ctx.get_forces(w)
        num += float(np.sum((Fg - r["forces"]) ** 2))
        den += float(np.sum(r["forces"] ** 2))
        amax = max(amax, float(np.abs(Fg).max()))
        w = (w + 1) % wk.W
    rate = evals / t_used
    return ({"value": round(rate, 2), "unit": "evals/s", "cores": 1, "kind": "port",
             "sample": f"{evals} walker evals (walkers 0..{min(evals, wk.W) - 1}, cycled) of the same workload, "
                       "FP64 C oracle driven from Python, single thread"},
            # all-zero FP64 forces (a flat bias): report the largest GPU force instead
            float(math.sqrt(num / den)) if den > 0 else {"fp64_forces": "all zero", "gpu_abs_max": amax})


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

    # ---- plain: one walker per launch sequence ----
    for _ in range(2):
        ctx.execute_all_plain()
    t_plain = time_loop(lambda k: ctx.execute_all_plain(), args.plain_steps, sync)
    log(f"plain {t_plain * 1e6:.1f} us/step")

    # ---- CPU baseline and force check, on the state before any deposit ----
    cpu = err = None
    if not args.no_cpu_baseline:
        cpu, err = cpu_baseline(wk, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']} evals/s, force rel RMS {err}")

    # ---- the same batched step replayed as one captured hipGraph ----
    for _ in range(10):
        ctx.execute_batch(use_graph=True)
    t_graph = time_loop(lambda k: ctx.execute_batch(use_graph=True), args.steps, sync)

    # ---- eval + deposit (pace 1: every walker deposits every step) ----
    ctx.update_state(1)
    ctx.execute_batch(use_graph=False)
    # (10 steps per form: the smallest hill capacity, path-batch's, holds 32 steps of 256 deposits)
    t_dep = time_loop(lambda k: (ctx.update_state(k + 2), ctx.execute_batch(use_graph=False)), 10, sync)
    # as one graph per deposit pattern (glued_step_batch)
    ctx.step_batch(100)
    t_dep_graph = time_loop(lambda k: ctx.step_batch(k + 101), 10, sync)

    # ---- roofline of the dominant kernel ----
    t_launch = wk.time_dominant(200)
    rf = wk.roofline(t_launch)
    rf["frac"] = round(rf["achieved"] / rf["peak"], 4)
    rf["achieved"] = round(rf["achieved"], 3)
    rf["traffic"] = pmc_traffic(wk.name)
    if rf["traffic"] is not None:
        rf["traffic_kernels"] = sorted(PMC_KIB[wk.name])
    rf["avg_launch_us"] = round(t_launch * 1e6, 3)
    stage_t = {}
    for st, name in ((0, "cv"), (1, "bias"), (2, "scatter")):
        stage_t[name] = round(time_stage(ctx, st, 100) * 1e6, 2)

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
    }
    if cpu is not None:
        out["cpu_baseline"], out["force_rel_rms_vs_fp64"] = cpu, err
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
