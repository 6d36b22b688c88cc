"""Autotuner for the TTSH3D step kernels.

For each grid size N it times, with HIP events through ttsh_profile, every
K_rhs tile variant that fits N at several z-depths, and several z-depths of
the divergence/corrector stencils, on a droplet of that size; then writes the
fastest choices as the tuning table libttsh3d.so reads (tuned.txt beside the
library, lines "N rhs_variant rhs_tz proj_tz").

    python -m ttsh3d.autotune [--sizes 48 64 96 128 192 256] [--steps 30] [--out PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import F64, TTSHError, Simulation, default_params, init_droplet, lib, LIB_PATH

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


def write_table(entries: dict, path: str):
    with open(path, "w") as f:
        f.write("# N rhs_variant rhs_tz proj_tz  (written by ttsh3d.autotune)\n")
        for N in sorted(entries):
            v, rtz, ptz = entries[N]
            f.write(f"{N} {v} {rtz} {ptz}\n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--sizes", type=int, nargs="+", default=[48, 64, 96, 128, 192, 256])
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--json", default=None, help="also write all timings here")
    a = ap.parse_args(argv)
    lib()
    entries, allres = {}, {}
    t0 = time.time()
    for N in a.sizes:
        best, res = tune_size(N, a.steps, log=lambda *x: print(*x, flush=True))
        entries[N] = best
        allres[N] = res
        print(f"N={N}: best rhs variant {best[0]} tz {best[1]}, proj tz {best[2]}", flush=True)
    write_table(entries, a.out)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(allres, f, indent=1)
    print(f"wrote {a.out} in {time.time() - t0:.1f} s", flush=True)
    lib().ttsh_tuning_reload(a.out.encode())
    return 0


if __name__ == "__main__":
    sys.exit(main())
