"""Autotuner for the TTSH3D step kernels.

For each grid size N it times, with HIP events through ttsh_profile, every
K_rhs tile variant that fits N at several z-depths, and several z-depths of
the divergence/corrector stencils, on a droplet of that size; then writes the
fastest choices as the tuning table libttsh3d.so reads (tuned.txt beside the
library, lines "N rhs_variant rhs_tz proj_tz").

With --batch it tunes the layout of a batch that fills the device instead:
on the 64-droplet sweep mix (L in {48, 64, 96, 128}) as one graph batch, it
times whole batch steps for every K_rhs tile that fits all sizes, minimum
K_rhs z-depth and divergence/corrector z-depth, through the table's "batch"
line (re-read between candidates; the batch is laid out again at its next
step), and keeps the fastest.  The FFT passes' lines per work-group are
compile-time constants and not part of the search.

    python -m ttsh3d.autotune [--sizes 48 64 96 128 192 256] [--steps 30] [--out PATH]
    python -m ttsh3d.autotune --batch [--steps 48] [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import F64, TTSHError, Batch, Simulation, default_params, init_droplet, lib, LIB_PATH, tuning_reload

RHS_VARIANTS = {0: (32, 8), 1: (16, 16), 2: (64, 4), 3: (16, 8), 4: (32, 8), 5: (32, 16), 6: (64, 8)}
DEFAULT_TABLE = os.path.join(os.path.dirname(LIB_PATH), "tuned.txt")


def _time(p, state, steps, warm=5):
    phi, P, v = state
    with Simulation(p) as s:
        s.set_state(phi, P, v, dtype=F64)
        s.step(warm)
        s.synchronize()
        return {k: ms / steps for k, ms in s.profile(steps).items()}


def tune_size(N: int, steps: int = 30, log=print):
    state = init_droplet(N, 0.3125 * N, 2.0, 0, 4242 + N)
    base = dict(N=N, dt0=0.01, cA=-2.0)
    best_rhs, best_proj = None, None
    results = []
    # z-depths: the engine splits z into ceil(N/tz) near-equal chunks, so
    # depths like 11 or 43 set the number of work items per CU
    tzs = sorted({t for t in (8, 11, 16, 22, 32, 43, 64) if t <= N})
    for var, (tx, ty) in RHS_VARIANTS.items():
        if N < tx + 2 or N < ty + 2:
            continue
        for tz in tzs:
            try:
                k = _time(default_params(rhs_variant=var, rhs_tz=tz, corr_tz=8, **base), state, steps)
            except TTSHError as e:
                log(f"N={N} variant {var} tz {tz}: {e}")
                continue
            results.append({"kind": "rhs", "variant": var, "tz": tz, "ms": k["K_rhs"]})
            if best_rhs is None or k["K_rhs"] < best_rhs[2]:
                best_rhs = (var, tz, k["K_rhs"])
            log(f"N={N} K_rhs variant {var} ({tx}x{ty}) tz {tz}: {k['K_rhs'] * 1e3:.1f} us")
    for ptz in [t for t in (4, 6, 8, 11, 16, 32) if t <= N]:
        k = _time(default_params(rhs_variant=best_rhs[0], rhs_tz=best_rhs[1], corr_tz=ptz, **base), state, steps)
        t = k["K_div"] + k["K_corr"]
        results.append({"kind": "proj", "tz": ptz, "ms": t})
        if best_proj is None or t < best_proj[1]:
            best_proj = (ptz, t)
        log(f"N={N} K_div+K_corr tz {ptz}: {t * 1e3:.1f} us")
    return (best_rhs[0], best_rhs[1], best_proj[0]), results


def read_table(path: str):
    """(per-size entries, batch entry or None) of a tuning table file."""
    entries, batch = {}, None
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                t = line.split("#")[0].split()
                if len(t) < 4:
                    continue
                if t[0] == "batch":
                    batch = tuple(int(x) for x in t[1:4])
                else:
                    entries[int(t[0])] = tuple(int(x) for x in t[1:4])
    return entries, batch


def write_table(entries: dict, path: str, batch=None):
    with open(path, "w") as f:
        f.write("# N rhs_variant rhs_tz proj_tz  (written by ttsh3d.autotune)\n")
        for N in sorted(entries):
            v, rtz, ptz = entries[N]
            f.write(f"{N} {v} {rtz} {ptz}\n")
        if batch is not None:
            f.write("# device-filling batch: rhs_variant min_rhs_tz proj_tz\n")
            f.write("batch {} {} {}\n".format(*batch))


def tune_batch(table: str, steps: int = 48, log=print):
    """Fastest device-filling batch layout on the 64-droplet sweep mix.
    Candidates are written one at a time into a scratch copy of the table
    (its per-size lines kept) and re-read; returns ((variant, rhs_tz,
    corr_tz), ms per batch step, all results)."""
    from . import workloads
    entries, _ = read_table(table)
    cases = workloads.sweep_mix(__import__("ttsh3d"))
    sizes = sorted({L for _, L, _ in cases})
    variants = [v for v, (tx, ty) in RHS_VARIANTS.items() if all(L >= tx + 2 and L >= ty + 2 for L in sizes)]
    scratch = table + ".tuning"
    results = []
    best = None
    b = Batch(max_slots=len(cases), use_graph=True)
    try:
        for p, L, s in cases:
            b.admit(p, *workloads.mix_state(__import__("ttsh3d"), L, s))
        for var in variants:
            for rtz in (16, 24, 32, 48, 64):
                for ctz in (16, 32, 64):
                    write_table(entries, scratch, (var, rtz, ctz))
                    tuning_reload(scratch)
                    b.step(16)  # re-layout and graph capture
                    b.synchronize()
                    t0 = time.perf_counter()
                    b.step(steps)
                    b.synchronize()
                    ms = (time.perf_counter() - t0) * 1e3 / steps
                    results.append({"variant": var, "rhs_tz": rtz, "corr_tz": ctz, "ms": ms})
                    log(f"batch variant {var} {RHS_VARIANTS[var]} rhs_tz {rtz} corr_tz {ctz}: {ms:.3f} ms/step")
                    if best is None or ms < best[1]:
                        best = ((var, rtz, ctz), ms)
    finally:
        b.close()
        if os.path.exists(scratch):
            os.remove(scratch)
        tuning_reload(table if os.path.exists(table) else None)
    return best[0], best[1], results


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--sizes", type=int, nargs="+", default=[48, 64, 96, 128, 192, 256])
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--json", default=None, help="also write all timings here")
    ap.add_argument("--batch", action="store_true", help="tune the device-filling batch layout instead")
    a = ap.parse_args(argv)
    lib()
    if a.batch:
        t0 = time.time()
        entries, _ = read_table(a.out)
        best, ms, res = tune_batch(a.out, a.steps if a.steps != 30 else 48,
                                   log=lambda *x: print(*x, flush=True))
        write_table(entries, a.out, best)
        if a.json:
            with open(a.json, "w") as f:
                json.dump({"batch": res, "best": best, "ms": ms}, f, indent=1)
        print(f"batch layout: variant {best[0]} rhs_tz {best[1]} corr_tz {best[2]} ({ms:.3f} ms per batch step); "
              f"wrote {a.out} in {time.time() - t0:.1f} s", flush=True)
        tuning_reload(a.out)
        return 0
    entries, allres = {}, {}
    t0 = time.time()
    for N in a.sizes:
        best, res = tune_size(N, a.steps, log=lambda *x: print(*x, flush=True))
        entries[N] = best
        allres[N] = res
        print(f"N={N}: best rhs variant {best[0]} tz {best[1]}, proj tz {best[2]}", flush=True)
    # sizes not tuned in this run keep their entries
    kept, batch = read_table(a.out)
    kept.update(entries)
    write_table(kept, a.out, batch)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(allres, f, indent=1)
    print(f"wrote {a.out} in {time.time() - t0:.1f} s", flush=True)
    lib().ttsh_tuning_reload(a.out.encode())
    return 0


if __name__ == "__main__":
    sys.exit(main())
