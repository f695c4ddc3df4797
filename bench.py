#!/usr/bin/env python3
"""Benchmark of the GLUED enhanced-sampling step on one MI355X.

Workload: the batched multi-walker configuration -- 64 walkers of uneven size
(2,000-40,000 atoms at 100 atoms/nm^3), Kabsch RMSD over 500 atoms + 2
torsions, well-tempered MetaD on (RMSD, phi) with a shared 200x200 grid and
20,000 private hills per walker.  One eval = every CV, the bias energy and
dV/ds, and the fixed-point force scatter, for one walker.

  value        batched evals/s: all 64 walkers per step as one captured
               hipGraph (CV kernel, hill-chunk kernel, finalize + scatter
               kernel)
  plain        the reference's structure: one walker per launch sequence
               (one launch per CV kind, per bias, and the scatter), host loop
               over walkers
  roofline     the batched hill-sum chunk kernel (HBM-bound), timed live
               with HIP events on the launch stream
  cpu_baseline the FP64 oracle on a bounded sample of the same walkers, one
               host core; its forces also give the force error vs FP64

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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--walkers", type=int, default=64)
    ap.add_argument("--hills", type=int, default=20000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="time only the batched step (for profiling its kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(walkers shard across GPUs by design only, see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from gluedmd_amd import workloads as wl
    from gluedmd_amd.context import Context

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    c = wl.config4(n_walkers=args.walkers, n_hills=args.hills, extra_hills=64)
    f = c["force"]
    W = args.walkers
    ctx = Context(f, walkers=c["sizes"])
    grid = wl.grid_prefill(f, c["grid_bias"], c["grid_prior"])
    ctx.set_grid(c["grid_bias"], grid)
    for w in range(W):
        ctx.set_box(c["boxes"][w], w)
        ctx.set_positions(c["positions"][w], w)
        ctx.set_hills(c["hill_bias"], c["hills"][w][0], c["hills"][w][1], w)
    print(f"bench: config ready ({W} walkers, {args.hills} hills/walker)", file=sys.stderr, flush=True)

    # ---- batched: one graph per step ----
    for _ in range(args.warmup):
        ctx.execute_batch(use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.execute_batch(use_graph=True)
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t0) / args.steps
    print(f"bench: batched {t_batched * 1e6:.1f} us/step", file=sys.stderr, flush=True)

    if args.batched_only:
        print(json.dumps({"metric": "batched step only", "ms_per_step": round(t_batched * 1e3, 4),
                          "value": round(W / t_batched, 1), "unit": "evals/s"}))
        return 0

    # ---- plain: one walker per launch sequence ----
    for _ in range(2):
        ctx.execute_all_plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        ctx.execute_all_plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / args.plain_steps
    print(f"bench: plain {t_plain * 1e6:.1f} us/step", file=sys.stderr, flush=True)

    # ---- eval + deposit (pace 1: a shared-grid and a hill deposit per walker) ----
    dep_steps = 20
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(dep_steps):
        ctx.update_state(k + 1)
        ctx.execute_batch(use_graph=True)
    torch.cuda.synchronize()
    t_dep = (time.perf_counter() - t0) / dep_steps

    # ---- roofline: the batched hill-sum chunk kernel, HIP events on the launch stream ----
    stream = torch.cuda.current_stream(dev)
    reps = 200
    for _ in range(10):
        ctx.execute_stage(3)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        ctx.execute_stage(3)
    e1.record(stream)
    torch.cuda.synchronize()
    t_hill = e0.elapsed_time(e1) / reps * 1e-3
    n_hills = args.hills + (dep_steps)  # hills grew by the deposits above
    D = 2
    hill_bytes = W * n_hills * 8 * (D + 1)
    achieved = hill_bytes / t_hill / 1e9
    # other stages (events too), for the record
    stage_t = {}
    for st, name in ((0, "cv"), (1, "bias"), (2, "scatter")):
        e0.record(stream)
        for _ in range(reps):
            ctx.execute_stage(st)
        e1.record(stream)
        torch.cuda.synchronize()
        stage_t[name] = round(e0.elapsed_time(e1) / reps * 1e3, 2)

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
        "data": "synthetic (seeded SplitMix64 positions, reference and hills)",
        "config": {
            "workload": "batched multi-walker: 64 walkers, 2k-40k atoms, Kabsch RMSD(500)+2 torsions, "
                        "WT MetaD shared 200x200 grid + 20k private hills/walker",
            "walkers": W,
            "hills_per_walker": args.hills,
            "parallelism": "batched walkers, one hipGraph per step (3 kernels)",
        },
        "plain_evals_per_s": round(W / t_plain, 1),
        "batched_over_plain": round(t_plain / t_batched, 2),
        "eval_plus_deposit_evals_per_s": round(W / t_dep, 1),
        "stage_us": stage_t,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_step_bias_chunks (hill sums, D=2)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_launch": hill_bytes,
            "avg_launch_us": round(t_hill * 1e6, 3),
        },
    }

    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["force_rel_rms_vs_fp64"] = cpu_baseline(ctx, c, args, f)
    print(json.dumps(out))
    return 0


def cpu_baseline(ctx, c, args, f):
    """FP64 oracle on the host: one core, walkers evaluated in turn until the
    time budget is used; also the checker of the GPU forces for those walkers."""
    import numpy as np
    import torch
    from oracle.oracle import OracleContext, _Grid

    torch.cuda.synchronize()
    gb, hb = c["grid_bias"], c["hill_bias"]
    t, cvs, params, ip = f.biases[gb]
    shared = _Grid(len(cvs), params, ip)
    # the GPU grid has taken deposits during the benchmark: read the current state
    blob = ctx.getBiasState(0)
    import struct
    # skip to the MetaD section: header(8) + nOpes(4)=0 + nAbmd(4)=0 + nMetaD(4) + nDeposited(4)
    off = 8 + 4 + 4 + 4 + 4
    shared.grid = np.frombuffer(blob[off:off + 8 * shared.total], dtype="<f8").copy()
    ctx.zero_forces()
    ctx.execute_batch(use_graph=True)
    torch.cuda.synchronize()
    evals, t_used, worst, num, den = 0, 0.0, 0.0, 0.0, 0.0
    w = 0
    while t_used < args.cpu_seconds and evals < 10 * args.walkers:
        o = OracleContext(f, c["sizes"][w], shared={gb: shared})
        o.set_box(c["boxes"][w])
        hc, hh = ctx.get_hills(hb, w)
        o.states[hb].centers, o.states[hb].heights = hc, hh
        pos = ctx.posq[ctx.walker_slice(w)][:c["sizes"][w], :3].cpu().numpy()
        t0 = time.perf_counter()
        r = o.execute(pos)
        t_used += time.perf_counter() - t0
        evals += 1
        Fg = ctx.get_forces(w)
        num += float(np.sum((Fg - r["forces"]) ** 2))
        den += float(np.sum(r["forces"] ** 2))
        w = (w + 1) % args.walkers
    rate = evals / t_used
    return ({"value": round(rate, 1), "unit": "evals/s", "cores": 1, "kind": "port",
             "sample": f"{evals} walker evals (walkers 0..{min(evals, args.walkers) - 1}, cycled) of the same "
                       "workload, FP64 C oracle driven from Python, single thread"},
            float(math.sqrt(num / den)))


if __name__ == "__main__":
    sys.exit(main())
