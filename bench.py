#!/usr/bin/env python3
"""Benchmark of the MI355X TTSH3D step on one GPU.

Metric (one JSON line): Mcell-steps/s of the 128^3 single droplet at FP32
with its share of HBM peak; the batched sweep's sims/hour; the served
stream's p50/p99 completion latency.

Single-droplet legs (ttsh3d.workloads): the 128^3 droplet (R = 40, xi = 2,
cA = -2, random-unit P) is `value`; the 256^3 droplet (R = 80, xi = 3, noisy
aster P), whose 1.55 GB working set exceeds the 256 MB Infinity Cache, is
reported beside it.  A "step" is one explicit-Euler step of the reference's
EEuler (TTSH3D.cu:347-359) on the droplet, device-resident.  W warm-up
steps, then max(K, 100) steps through ttsh_profile, whose kernels record HIP
events at their own start and end on the step's stream (kernel_ms: each
kind's average duration, as a kernel trace measures it; these steps also
bring the GPU to its sustained clock), then exactly K timed steps between
two synchronisations, launched as a caller launches them (ttsh_step: the
step's kernels queued back to back on one stream); value = N^3 * K / wall
seconds / 1e6.

roofline: the dominant kernel's algorithmic bytes per launch (FP32 field
traffic: K_rhs and K_corr 56 B/cell, K_div 32, each FFT pass 16) divided by
its average duration, against MI355X's 8 TB/s HBM peak; "kernels" lists the
same for every kernel kind.  traffic: HBM bytes per launch from rocprofv3
FETCH_SIZE + WRITE_SIZE (calibrated against a plain copy kernel), when
bench_traffic.json beside this file holds them for the grid size.

cpu_baseline: the FP64 oracle (OpenMP, all host threads) on the same 128^3
initial state, timed over whole steps until at least 10 s of CPU work (a
bounded sample of the same workload).

batch_sweep: the 64-droplet mix (L in {48, 64, 96, 128}, seed 1004) for
--batch-steps steps (default 2,000, the full sweep) as one HIP-graph batch,
and the same droplets one at a time through the plain step; sims/hour of
each and their ratio.

serve: the first --serve-requests requests of the served stream (seeded
Poisson arrivals at 80% of the measured batch capacity, L in {64, 96, 128,
192}, 1,000 to 20,000 steps each) through the continuous-batching loop;
sims/hour and p50/p99 completion latency of that bounded sample.
"""
import argparse
import json
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
STEP_BYTES = sum(ALG_BYTES.values())  # bytes per cell-step this design moves (five FFT passes)
PLAN_BYTES = 176  # the four-stage plan: one forward and one inverse transform of 16 B/cell each
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(ttsh, N):
    from ttsh3d import workloads
    return workloads.single(ttsh, N)


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


def kernel_table(avg_ms, N, traffic=None):
    """Per kernel kind: average ms, algorithmic GB/s and its share of peak,
    and the calibrated HBM bytes per launch where known."""
    cells = N ** 3
    out = {}
    for k, ms in avg_ms.items():
        gbs = ALG_BYTES[k] * cells / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[k] = {"ms": round(ms, 5), "alg_bytes_per_cell": ALG_BYTES[k], "achieved_GBs": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 4),
                  "traffic": (traffic or {}).get(k)}
    return out


def single_leg(ttsh, N, steps, warmup):
    """One droplet of the N^3 workload: instrumented and timed steps."""
    p, phi, P, v = workload(ttsh, N)
    cells = N ** 3
    sim = ttsh.Simulation(p, device=0)
    sim.set_state(phi, P, v)
    sim.step(warmup)
    sim.synchronize()
    # per-kernel durations: max(K, 100) steps whose kernels record HIP events
    # at their own start and end on the step's stream (plain launches).  They
    # also keep the GPU busy long enough (~20 ms at 128^3) to reach its
    # sustained clock before the timed region: after an idle spell the first
    # ~15 ms of steps run up to 7% slower, which a few warm-up steps do not cover.
    nprof = max(steps, 100)
    t1 = time.perf_counter()
    kms = sim.profile(nprof)
    sim.synchronize()
    wall_prof = time.perf_counter() - t1
    # timed region: K steps as a caller runs them (ttsh_step, no
    # instrumentation)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    red, finite = sim.reductions()
    sim.close()
    if not finite:
        raise RuntimeError(f"non-finite state after the timed steps at {N}^3: {red}")
    avg = {k: t / nprof for k, t in kms.items()}
    traffic = None
    tpath = os.path.join(ROOT, "bench_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(str(N))
    value = cells * steps / wall / 1e6
    return {"value": value, "ms_per_step": wall * 1e3 / steps, "ms_per_step_instrumented": wall_prof * 1e3 / nprof,
            "kernels": kernel_table(avg, N, traffic), "avg": avg, "traffic": traffic,
            "reductions_fp64": {"mass": red[0], "mean_abs_P": red[1], "kinetic_energy": red[2]}, "params": p,
            "state": (phi, P, v)}


def roofline(leg, N):
    dom = max(leg["avg"], key=leg["avg"].get)
    k = leg["kernels"][dom]
    return {"bound": "hbm", "achieved": k["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": k["frac"], "traffic": k["traffic"], "kernel": dom, "alg_bytes_per_cell": ALG_BYTES[dom],
            "grid": N}


def batch_sweep(ttsh, steps, plain=True):
    """The 64-droplet sweep mix as one graph batch against the plain step one
    droplet at a time (plain=False: the batched part only), both timed over
    `steps` steps per droplet."""
    from ttsh3d import workloads
    cases = workloads.sweep_mix(ttsh)
    ndrop = len(cases)
    cells = int(sum(L ** 3 for _, L, _ in cases))
    states = [workloads.mix_state(ttsh, L, s) for _, L, s in cases]
    b = ttsh.Batch(max_slots=ndrop, use_graph=True)
    for (p, L, s), st in zip(cases, states):
        b.admit(p, *st)
    b.step(16)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    tb = time.perf_counter() - t0
    finite = all(b.reductions(s)[1] for s in range(ndrop))
    b.close()
    full = 2000
    sph_b = ndrop / (tb * full / steps) * 3600.0
    out = {"droplets": ndrop, "cells": cells, "steps_timed": steps, "batched_sims_per_hour": round(sph_b, 2),
           "batched_s": round(tb, 3), "batched_mcell_steps_per_s": round(cells * steps / tb / 1e6, 1),
           "finite": finite}
    if not plain:
        return out
    # plain step, one droplet at a time (same droplets, same steps)
    tp = 0.0
    for (p, L, s), st in zip(cases, states):
        with ttsh.Simulation(p) as sim:
            sim.set_state(*st)
            sim.step(4)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.step(steps)
            sim.synchronize()
            tp += time.perf_counter() - t0
    sph_p = ndrop / (tp * full / steps) * 3600.0
    out.update({"plain_sims_per_hour": round(sph_p, 2), "plain_s": round(tp, 3), "ratio": round(sph_b / sph_p, 3),
                "note": "sims/hour for 2,000-step droplets, " + (
                    "timed over all 2,000 steps" if steps == full else "extrapolated from the timed steps")})
    return out


def serve_leg(nreq, rate_frac=0.8, budget_gib=64.0, max_slots=64):
    """The first nreq requests of the served stream through the request loop."""
    from ttsh3d import serve, workloads
    sizes, steps = workloads.SERVE_SIZES, (1000, 20000)
    budget = int(budget_gib * (1 << 30))
    cap = serve.measure_capacity(sizes, steps, budget, max_slots)
    mean_work = np.mean([N ** 3 for N in sizes]) * (steps[0] + steps[1]) / 2.0
    rate = rate_frac * cap / mean_work
    reqs = serve.make_requests(nreq, rate, sizes, steps, workloads.SERVE_SEED)
    be = serve.GpuBackend(max_slots, budget)
    done, wall = serve.Server(be, macro=16).run(reqs, log=log)
    be.close()
    out = serve.summarize(done, wall)
    out.update({"requests": nreq, "of_stream": 500, "rate_per_s": round(rate, 4), "rate_frac": rate_frac,
                "capacity_mcell_steps_per_s": round(cap / 1e6, 1), "sizes": list(sizes), "steps": list(steps),
                "note": f"the first {nreq} requests of the 500-request stream (seed {workloads.SERVE_SEED})"})
    return out


def main():
    from ttsh3d import workloads
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sweep")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--no-256", action="store_true", help="skip the 256^3 single-droplet leg")
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched part of the sweep (for kernel traces of the batch) and print it")
    ap.add_argument("--grid", type=int, default=128, choices=sorted(workloads.SINGLE),
                    help="single-droplet workload of `value`: 128 (default), 64 or 256")
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
    N = args.grid
    log(f"{N}^3 single droplet ...")
    leg = single_leg(ttsh, N, args.steps, args.warmup)
    value = leg["value"]
    out = {
        "metric": f"Mcell-steps/s at {N}^3 FP32 (single droplet TTSH3D step)",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(leg["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded droplet initial condition)",
        "config": {"workload": workloads.SINGLE[N]["desc"], "grid": N, "dt": 0.01},
        # share of HBM peak at the bytes this step moves (224 B/cell: five FFT
        # passes), and at the 176 B/cell of the four-stage plan
        "hbm_frac_step": round(value * 1e6 * STEP_BYTES / (HBM_PEAK_GBS * 1e9), 4),
        "hbm_frac_step_plan176": round(value * 1e6 * PLAN_BYTES / (HBM_PEAK_GBS * 1e9), 4),
        "kernel_ms": {k: v["ms"] for k, v in leg["kernels"].items()},
        "kernels": leg["kernels"],
        "ms_per_step_instrumented": round(leg["ms_per_step_instrumented"], 5),
        "roofline": roofline(leg, N),
        "reductions_fp64": leg["reductions_fp64"],
    }
    p, (phi, P, v) = leg["params"], leg["state"]
    if N == 128 and not args.no_256:
        log("256^3 single droplet ...")
        l2 = single_leg(ttsh, 256, args.steps, args.warmup)
        out["grid_256"] = {
            "workload": workloads.SINGLE[256]["desc"], "value": round(l2["value"], 2), "unit": "Mcell-steps/s",
            "ms_per_step": round(l2["ms_per_step"], 5), "steps": args.steps, "warmup": args.warmup,
            "hbm_frac_step": round(l2["value"] * 1e6 * STEP_BYTES / (HBM_PEAK_GBS * 1e9), 4),
            "hbm_frac_step_plan176": round(l2["value"] * 1e6 * PLAN_BYTES / (HBM_PEAK_GBS * 1e9), 4),
            "kernels": l2["kernels"], "roofline": roofline(l2, 256),
            "ms_per_step_instrumented": round(l2["ms_per_step_instrumented"], 5),
            "reductions_fp64": l2["reductions_fp64"]}
        del l2
    if not args.no_batch:
        log("batched sweep ...")
        out["batch_sweep"] = batch_sweep(ttsh, args.batch_steps)
    if not args.no_serve and args.serve_requests > 0:
        log("served stream ...")
        out["serve"] = serve_leg(args.serve_requests)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(p, phi, P, v)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
