#!/usr/bin/env python3
"""Benchmark of the MI355X TTSH3D step on one GPU.

Workload (default --grid 128, the 128^3 single-droplet configuration): periodic 128^3 box, h = 1,
droplet phi = (1 - tanh((r - 40)/2))/2 at the box centre, P a seeded random
unit vector where phi > 0.5 (zero outside), v = 0, dt = 0.01, extensile
activity cA = 2 x the run_cuda.pl default (-2), all other parameters at the
run_cuda.pl defaults.  FP32 fields, FP64 reductions.  --grid 256 runs the 256^3
droplet (R = 80, xi = 3, noisy aster P, run_cuda.pl defaults).

A "step" is one explicit-Euler step of the reference's EEuler
(TTSH3D.cu:347-359) on the droplet, device-resident.  W warm-up steps, then
exactly K timed steps between two synchronisations, launched as a caller
launches them (ttsh_step: the step's kernels queued back to back on one
stream, no instrumentation); value = N^3 * K / wall seconds / 1e6
(Mcell-steps/s).  Then K more steps with HIP events recorded on
the step's stream between consecutive kernel kinds give each kernel's average
duration (kernel_ms; the events add launch gaps, so this run's own wall time,
ms_per_step_instrumented, is longer).

roofline: the dominant kernel's algorithmic bytes per launch (FP32 field
traffic: K_rhs and K_corr 56 B/cell, K_div 32, each FFT pass 16) divided by
its average duration, against MI355X's 8 TB/s HBM peak.  traffic: HBM bytes
per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE (calibrated against a
plain copy kernel), when bench_traffic.json beside this file holds them for
the grid size (profiles/step_<N>_hbm_traffic.md records those runs).

cpu_baseline: the FP64 oracle (OpenMP, all host threads) on the same initial
state, timed over whole steps until at least 10 s of CPU work (a bounded
sample of the same workload).

The batched-sweep line (--batch, on by default) runs 64 droplets with L drawn
from {48, 64, 96, 128} (R = 0.3 L, cA = default x logU[0.1, 10], La^2 =
default x U[0.5, 2]) for a bounded number of steps as one HIP-graph batch and
one at a time through the plain step, and reports sims/hour for 2,000-step
droplets.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "bacterial-droplet_amd"))

# Algorithmic FP32 field traffic per cell of each kernel kind (reads + writes):
# K_rhs reads phi, P, v and writes phi', v*, P*; K_div reads v*, P* and writes
# the complex s; each FFT pass reads and writes the complex buffer (K_fft_xy
# and K_fft_xy_inv are the x passes, or in large batches the fused x+y
# passes); K_corr reads p (complex), v*, P* and writes v, P.
ALG_BYTES = {"K_rhs": 56, "K_div": 32, "K_fft_xy": 16, "K_fft_y": 16, "K_fft_z_poisson": 16,
             "K_fft_y_inv": 16, "K_fft_xy_inv": 16, "K_corr": 56}
STEP_BYTES = 176  # algorithmic bytes per cell-step of the whole step (four-stage figure)
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = {
    # 128^3 single droplet: R=40, xi=2, cA = 2x default, seeded random-unit P
    128: dict(R=40.0, xi=2.0, pmode=0, cA=-2.0, seed=1002,
              desc="single droplet 128^3, R=40, xi=2, cA=-2, random-unit P, dt=0.01"),
    # 256^3 single droplet: R=80, xi=3, seeded noisy aster P, defaults otherwise
    256: dict(R=80.0, xi=3.0, pmode=1, cA=-1.0, seed=1003,
              desc="single droplet 256^3, R=80, xi=3, aster P (r_hat + 0.1 N(0,1)), dt=0.01"),
}


def workload(ttsh, N):
    w = WORKLOADS[N]
    p = ttsh.default_params(N=N, dt0=0.01, cA=w["cA"])
    phi, P, v = ttsh.init_droplet(N, w["R"], w["xi"], w["pmode"], w["seed"])
    return p, phi, P, v


def cpu_baseline(p, phi, P, v, min_s=10.0, max_steps=10000):
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    O.set_threads(threads)
    keys = ("h", "dt0", "mu", "cA", "alpha", "beta", "eta", "La", "kappa", "lambda0", "aPhi", "kPhi", "gamma", "A")
    o = O.Oracle(p.Nx, faithful=True, **{k: getattr(p, k) for k in keys})
    o.set_state(phi, P, v)
    o.step(1)  # warm-up
    # whole steps until at least min_s of CPU work has been timed
    n, t0 = 0, time.perf_counter()
    while n < max_steps:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 >= min_s:
            break
    dt = time.perf_counter() - t0
    cells = p.Nx ** 3
    return {"value": round(cells * n / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": O.threads(),
            "kind": "port",
            "sample": f"FP64 oracle (OpenMP), the same {p.Nx}^3 droplet, {n} steps after 1 warm-up step, {dt:.1f} s"}


def batch_sweep(ttsh, steps, seed=1004, ndrop=64, plain=True):
    """Batched 64-droplet sweep vs the plain step one droplet at a time
    (plain=False: the batched part only)."""
    rng = np.random.default_rng(seed)
    Ls = rng.choice([48, 64, 96, 128], size=ndrop)
    base = ttsh.default_params()
    cases = []
    for i, L in enumerate(Ls):
        cA = base.cA * math.exp(rng.uniform(math.log(0.1), math.log(10.0)))
        La2 = base.La ** 2 * rng.uniform(0.5, 2.0)
        p = ttsh.default_params(N=int(L), dt0=0.01, cA=cA, La=math.sqrt(La2))
        cases.append((p, int(L), seed * 1000 + i))
    cells = int(sum(int(L) ** 3 for L in Ls))
    # batched, graph-captured
    b = ttsh.Batch(max_slots=ndrop, use_graph=True)
    for p, L, s in cases:
        phi, P, v = ttsh.init_droplet(L, 0.3 * L, 2.0, 0, s)
        b.admit(p, phi, P, v)
    b.step(16)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    tb = time.perf_counter() - t0
    b.close()
    full = 2000
    sph_b = ndrop / (tb * full / steps) * 3600.0
    if not plain:
        return {"droplets": ndrop, "cells": cells, "steps_timed": steps, "batched_sims_per_hour": round(sph_b, 2),
                "batched_mcell_steps_per_s": round(cells * steps / tb / 1e6, 1)}
    # plain step, one droplet at a time (same droplets, same steps)
    tp = 0.0
    for p, L, s in cases:
        phi, P, v = ttsh.init_droplet(L, 0.3 * L, 2.0, 0, s)
        with ttsh.Simulation(p) as sim:
            sim.set_state(phi, P, v)
            sim.step(4)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.step(steps)
            sim.synchronize()
            tp += time.perf_counter() - t0
    sph_p = ndrop / (tp * full / steps) * 3600.0
    return {"droplets": ndrop, "cells": cells, "steps_timed": steps,
            "batched_sims_per_hour": round(sph_b, 2), "plain_sims_per_hour": round(sph_p, 2),
            "ratio": round(sph_b / sph_p, 3),
            "batched_mcell_steps_per_s": round(cells * steps / tb / 1e6, 1),
            "note": "sims/hour for 2,000-step droplets, " + (
                "timed over all 2,000 steps" if steps == full else "extrapolated from the timed steps")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sweep")
    ap.add_argument("--batch-steps", type=int, default=64)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched part of the sweep (for kernel traces of the batch) and print it")
    ap.add_argument("--grid", type=int, default=128, choices=sorted(WORKLOADS),
                    help="single-droplet workload: 128 (default) or 256")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (3 launches), the calibration "
                         "kernel for rocprofv3 FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "droplets are independent, so a multi-GPU sweep places whole droplets per GPU.", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib > 0:
        import torch
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    import ttsh3d as ttsh
    if args.batch_only:
        print(json.dumps({"batch_sweep": batch_sweep(ttsh, args.batch_steps, plain=False)}), flush=True)
        return
    p, phi, P, v = workload(ttsh, args.grid)
    N = p.Nx
    cells = N ** 3
    sim = ttsh.Simulation(p, device=0)
    sim.set_state(phi, P, v)
    sim.step(args.warmup)
    sim.synchronize()
    # timed region: K steps as a caller runs them (ttsh_step, no
    # instrumentation)
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    # per-kernel durations: K more steps with HIP events recorded on the
    # step's stream at every kernel-kind boundary (plain launches)
    t1 = time.perf_counter()
    kms = sim.profile(args.steps)
    sim.synchronize()
    wall_prof = time.perf_counter() - t1
    red, finite = sim.reductions()
    sim.close()
    if not finite:
        raise RuntimeError(f"non-finite state after the timed steps: {red}")

    ms_step = wall * 1e3 / args.steps
    value = cells * args.steps / wall / 1e6
    avg = {k: t / args.steps for k, t in kms.items()}
    dom = max(avg, key=avg.get)
    achieved = ALG_BYTES[dom] * cells / (avg[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "bench_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(str(N), {}).get(dom)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
            "alg_bytes_per_cell": ALG_BYTES[dom]}
    out = {
        "metric": f"Mcell-steps/s at {N}^3 FP32 (single droplet TTSH3D step)",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded droplet initial condition)",
        "config": {"workload": WORKLOADS[N]["desc"],
                   "grid": N, "dt": 0.01},
        "hbm_frac_step": round(value * 1e6 * STEP_BYTES / (HBM_PEAK_GBS * 1e9), 4),
        "kernel_ms": {k: round(t, 5) for k, t in avg.items()},
        "ms_per_step_instrumented": round(wall_prof * 1e3 / args.steps, 5),
        "roofline": roof,
        "reductions_fp64": {"mass": red[0], "mean_abs_P": red[1], "kinetic_energy": red[2]},
    }
    if not args.no_batch:
        log("batched sweep ...")
        out["batch_sweep"] = batch_sweep(ttsh, args.batch_steps)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, phi, P, v)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
