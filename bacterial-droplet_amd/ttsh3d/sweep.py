"""Parameter sweeps: forEach.pl's knob grid as one batch of droplets.

forEach.pl:12-35 holds one list per knob (mu, cA, A, eta, La, alpha, beta,
lambda0, kappa, aPhi, kPhi, gamma) and global settings (Tend, dT, dTe, init0,
N, h).  forEach.pl:37-133 loops over the cartesian product, outermost mu,
innermost gamma.  For each point it rewrites run_cuda.pl and runs it, one
simulation after another.  run_cuda.pl:89-146 then writes input.dat and
readme.dat into a run directory named by forEach.pl:70, runs the step
binary, and moves the directory to conf-data/<name>.

Here the same grid, in the same order and with the same directory names,
runs as one batch.  Every point shares the global settings, so all points
export at the same steps.  The batch steps to each export and writes every
point's conf_<m>.dat.  Each point's result is bitwise equal to its run
alone, because the batch is deterministic per droplet.

forEach.pl:56 sets kPhi to the aPhi value (and :70 names the directory with
it).  That is reproduced by default; fix_kphi=True uses the kPhi list.

    python -m ttsh3d.sweep --out conf-data --N 64 --Tend 1000 --dT 0.002 --dTe 10 \\
        --cA -1 -0.5 --La 1 2 [--fix-kphi] [--seed 1] [--dry-run]
"""
from __future__ import annotations

import argparse
import itertools
import math
import os
import sys

from . import Batch, default_params, init_reference, write_conf_arrays, write_input_dat

# forEach.pl:12-23, in loop order (outermost first), with the run_cuda.pl
# parameter each one sets
KNOBS = ("mu", "cA", "A", "eta", "La", "alpha", "beta", "lambda0", "kappa", "aPhi", "kPhi", "gamma")
DEFAULT_LISTS = {"mu": [1], "cA": [-1], "A": [0], "eta": [-0.187], "La": [1], "alpha": [-0.02], "beta": [4.366],
                 "lambda0": [9], "kappa": [-0.364], "aPhi": [1], "kPhi": [1], "gamma": [1]}
# forEach.pl:25-31
DEFAULT_GLOBALS = {"Tend": 1000, "dT": 0.002, "dTe": 10, "init0": 0.0001, "N": 64, "h": 1}


def perl_num(x) -> str:
    """A number as Perl interpolates it into a string (%.15g)."""
    return "%.15g" % x


def run_name(k: dict, g: dict, fix_kphi: bool = False) -> str:
    """The run directory name of forEach.pl:70 (kPhi printed from aPhi unless fixed)."""
    kphi = k["kPhi"] if fix_kphi else k["aPhi"]
    n = perl_num
    return (f"eta_{n(k['eta'])}-LA_{n(k['La'])}-alpha_{n(k['alpha'])}-beta_{n(k['beta'])}-mu_{n(k['mu'])}"
            f"-cA_{n(k['cA'])}-A_{n(k['A'])}-lambda0_{n(k['lambda0'])}-aPhi_{n(k['aPhi'])}-kPhi_{n(kphi)}"
            f"-gamma_{n(k['gamma'])}-N_{n(g['N'])}-h_{n(g['h'])}-T_{n(g['Tend'])}-DT_{n(g['dTe'])}")


def grid(lists: dict | None = None, globals_: dict | None = None, fix_kphi: bool = False):
    """[(name, params)] for every point of the knob grid, in forEach.pl's loop order."""
    lists = {**DEFAULT_LISTS, **(lists or {})}
    g = {**DEFAULT_GLOBALS, **(globals_ or {})}
    out = []
    for combo in itertools.product(*(lists[k] for k in KNOBS)):
        k = dict(zip(KNOBS, combo))
        p = default_params(N=int(g["N"]), h=float(g["h"]), dt0=float(g["dT"]), Ts=float(g["Tend"]),
                           dte=float(g["dTe"]), init0=float(g["init0"]))
        for name in KNOBS:
            setattr(p, name, float(k[name]))
        if not fix_kphi:
            p.kPhi = float(k["aPhi"])  # forEach.pl:56
        out.append((run_name(k, g, fix_kphi), p))
    return out


def write_readme(path: str, p) -> None:
    """readme.dat of run_cuda.pl:117-128: T0, Ts, dte, Nx, Ny."""
    with open(path, "w") as f:
        for x in (p.T0, p.Ts, p.dte, p.Nx, p.Ny):
            f.write(perl_num(x) + "\n")


def run_sweep(points, out_root: str, seed: int = 1, device: int = 0, log=print) -> int:
    """Run every (name, params) point as one batch; returns the steps taken.

    out_root/<name>/ receives input.dat, readme.dat and conf_<m>.dat on the
    schedule of evolve() (TTSH3D.cu:311-331).  The initial state of point i is
    InitCond 1 (TTSH3D.cu:798-825) from the SplitMix64 stream of seed + i."""
    if not points:
        return 0
    p0 = points[0][1]
    for _, p in points:
        if (p.Nx, p.dt0, p.Ts, p.dte, p.T0) != (p0.Nx, p0.dt0, p0.Ts, p0.dte, p0.T0):
            raise ValueError("sweep points must share N, dt0, Ts, dte and T0 (forEach.pl's globals)")
    dirs = []
    with Batch(device=device, max_slots=len(points)) as b:
        slots = []
        for i, (name, p) in enumerate(points):
            d = os.path.join(out_root, name)
            os.makedirs(d, exist_ok=True)
            write_input_dat(os.path.join(d, "input.dat"), p)
            write_readme(os.path.join(d, "readme.dat"), p)
            phi, P, v, vs, ps = init_reference(p, seed + i)
            slots.append(b.admit(p, phi, P, v, vs, ps))
            dirs.append(d)

        def export(m: int):
            for s, d in zip(slots, dirs):
                phi, P, v = b.get_state(s)
                write_conf_arrays(os.path.join(d, f"conf_{m}.dat"), phi, P, v, expo_data=int(p0.ExpoData))

        if p0.T0 == 0:
            export(0)
        T, Dt = float(p0.T0), float(p0.dt0)
        total = pending = 0
        while True:
            T = T + Dt
            pending += 1
            expo = math.floor(T / p0.dte) > math.floor((T - Dt) / p0.dte)
            stop = T > p0.Ts
            if expo or stop:
                b.step(pending)
                total += pending
                pending = 0
                if expo:
                    m = int(math.floor(T / p0.dte))
                    export(m)
                    log(f"export {m} at T = {T:.6g} ({total} steps)")
                if stop:
                    break
        b.synchronize()
    return total


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    for k in KNOBS:
        ap.add_argument(f"--{k}", type=float, nargs="+", default=DEFAULT_LISTS[k])
    for k, v in DEFAULT_GLOBALS.items():
        ap.add_argument(f"--{k}", type=type(v) if k != "N" else int, default=v)
    ap.add_argument("--out", default="conf-data")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--fix-kphi", action="store_true", help="use the kPhi list (forEach.pl:56 uses aPhi)")
    ap.add_argument("--dry-run", action="store_true", help="print the grid and exit")
    a = ap.parse_args(argv)
    lists = {k: getattr(a, k) for k in KNOBS}
    g = {k: getattr(a, k) for k in DEFAULT_GLOBALS}
    pts = grid(lists, g, fix_kphi=a.fix_kphi)
    for name, _ in pts:
        print(name)
    if a.dry_run:
        return 0
    steps = run_sweep(pts, a.out, seed=a.seed, device=a.device)
    print(f"{len(pts)} runs, {steps} steps each")
    return 0


if __name__ == "__main__":
    sys.exit(main())
