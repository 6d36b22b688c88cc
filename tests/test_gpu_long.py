"""Long parity runs at the workload sizes, and the process drop-in.

  - 64^3 droplet (R=20, xi=2, random-unit P, dt=0.01, run_cuda.pl defaults),
    1,000 steps: FP32 GPU vs FP64 oracle, relative RMS <= 1e-4 per field;
  - 128^3 droplet (R=40, xi=2, cA=-2), 1,000 steps: same bar;
  - 64^3 aster (R=20, xi=3, radial P, cA=-1: the 256^3 workload's initial
    condition at a quarter of the size), 1,000 steps: same bar; its div P*
    exceeds div v* by two orders of magnitude, the case the power-of-two
    scale of the packed projection is for;
  - the ttsh3d CLI on a run_cuda.pl-style data_<n>/input.dat: the snapshot
    schedule of evolve() (TTSH3D.cu:311-331) and the conf_<m>.dat format.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import max_rel, rel_rms

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = ("h", "dt0", "mu", "cA", "alpha", "beta", "eta", "La", "kappa", "lambda0", "aPhi", "kPhi", "gamma", "A")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N,R,xi,pmode,cA,seed", [(64, 20.0, 2.0, 0, -1.0, 1001), (128, 40.0, 2.0, 0, -2.0, 1002),
                                                 (64, 20.0, 3.0, 1, -1.0, 1003)])
def test_thousand_steps_droplet(ttsh, oracle_mod, N, R, xi, pmode, cA, seed):
    log(f"[thousand_steps_droplet N={N} pmode={pmode}]")
    p = ttsh.default_params(N=N, dt0=0.01, cA=cA)
    phi, P, v = ttsh.init_droplet(N, R, xi, pmode, seed)
    o = oracle_mod.Oracle(N, faithful=True, **{k: getattr(p, k) for k in KEYS})
    o.set_state(phi, P, v)
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v)
        sim.step(1000)
        g = sim.get_state()
        red, ok = sim.reductions()
    o.step(1000)
    ref = o.get_state()
    ored = o.reductions()
    errs = [rel_rms(a, b) for a, b in zip(g, ref)]
    log(f"N={N} 1000 steps rel RMS phi {errs[0]:.2e} P {errs[1]:.2e} v {errs[2]:.2e}; "
        f"max rel {max_rel(g[0], ref[0]):.2e} {max_rel(g[1], ref[1]):.2e} {max_rel(g[2], ref[2]):.2e}")
    log("reductions gpu", red, "oracle", ored)
    assert ok
    assert max(errs) <= 1e-4
    assert red[0] == pytest.approx(ored[0], rel=1e-6)
    assert red[1] == pytest.approx(ored[1], rel=1e-4)
    assert red[2] == pytest.approx(ored[2], rel=1e-3)


def test_cli_run_dir(ttsh, tmp_path):
    """bin/ttsh3d reads data_<n>/input.dat and writes conf_<m>.dat on the
    floor(T/dte) schedule with T += Dt in double, stopping once T > Ts."""
    log("[cli_run_dir]")
    exe = os.path.join(ROOT, "bacterial-droplet_amd", "bin", "ttsh3d")
    p = ttsh.default_params(N=32, dt0=0.01, Ts=0.5, dte=0.13)
    d = tmp_path / "data_3"
    d.mkdir()
    ttsh.write_input_dat(str(d / "input.dat"), p)
    r = subprocess.run([exe, "3", "--seed", "7"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    log(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    assert "Simulation finished." in r.stdout
    # replay the reference loop (TTSH3D.cu:311-331) in double
    T, Dt, expected, steps = 0.0, 0.01, {0}, 0
    while True:
        T = T + Dt
        steps += 1
        if math.floor(T / 0.13) > math.floor((T - Dt) / 0.13):
            expected.add(int(math.floor(T / 0.13)))
        if T > 0.5:
            break
    got = {int(f[5:-4]) for f in os.listdir(d) if f.startswith("conf_")}
    assert got == expected
    assert f"({steps} steps)" in r.stdout
    lines = open(d / f"conf_{max(expected)}.dat").read().splitlines()
    assert len(lines) == 32 ** 3 and len(lines[0].split(" ")) == 7
