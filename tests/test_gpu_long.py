"""Long parity runs at the workload sizes, and the process drop-in.

  - 64^3 droplet (R=20, xi=2, random-unit P, dt=0.01, run_cuda.pl defaults),
    1,000 steps: FP32 GPU vs FP64 oracle, relative RMS <= 1e-4 per field;
  - 128^3 droplet (R=40, xi=2, cA=-2), about 1,000 steps through ttsh_run_ex
    with dte = 5: the reference's time loop (TTSH3D.cu:311-331), whose text
    snapshots conf_1.dat and conf_2.dat (the 500-step exports of the 10,000-
    step 128^3 run) are parsed back and checked against the oracle at their
    steps, and the final state after the loop's last step: same bar;
  - 64^3 aster (R=20, xi=3, radial P, cA=-1: the 256^3 workload's initial
    condition at a quarter of the size), 1,000 steps: same bar; its div P*
    exceeds div v* by two orders of magnitude, the case the power-of-two
    scale of the packed projection is for;
  - the 256^3 aster workload itself (R=80, xi=3, noisy aster P, cA=-1),
    40 steps: same bar;
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


def export_schedule(Ts, dte, Dt):
    """Steps after which the reference loop (TTSH3D.cu:311-331) writes
    conf_<m>.dat, m >= 1, and the number of steps it takes: T += Dt in
    double, an export when floor(T/dte) grows, a stop once T > Ts."""
    T, steps, exports = 0.0, 0, {}
    while True:
        T = T + Dt
        steps += 1
        if math.floor(T / dte) > math.floor((T - Dt) / dte):
            exports[int(math.floor(T / dte))] = steps
        if T > Ts:
            return exports, steps


def read_conf_text(path, N):
    """phi, P, v of a conf_<m>.dat (ExpoData 1): one line per (y, x, z), z
    fastest, columns vx vy vz Px Py Pz phi (TTSH3D.cu:975-998), as [z][y][x]."""
    import pandas as pd
    a = pd.read_csv(path, sep=" ", header=None, dtype=np.float64, engine="c").to_numpy()
    assert a.shape == (N ** 3, 7)
    f = a.reshape(N, N, N, 7).transpose(3, 2, 0, 1)  # (field, z, y, x)
    return f[6], f[3:6], f[0:3]


def check_fields(g, ref, what, bar=1e-4):
    errs = [rel_rms(a, b) for a, b in zip(g, ref)]
    log(f"{what}: rel RMS phi {errs[0]:.2e} P {errs[1]:.2e} v {errs[2]:.2e}; "
        f"max rel {max_rel(g[0], ref[0]):.2e} {max_rel(g[1], ref[1]):.2e} {max_rel(g[2], ref[2]):.2e}")
    assert max(errs) <= bar, f"{what}: rel RMS {max(errs):.2e} above {bar}"


@pytest.mark.parametrize("N,R,xi,pmode,cA,seed", [(64, 20.0, 2.0, 0, -1.0, 1001), (128, 40.0, 2.0, 0, -2.0, 1002),
                                                 (64, 20.0, 3.0, 1, -1.0, 1003)])
def test_thousand_steps_droplet(ttsh, oracle_mod, N, R, xi, pmode, cA, seed, tmp_path):
    log(f"[thousand_steps_droplet N={N} pmode={pmode}]")
    text = N == 128
    p = ttsh.default_params(N=N, dt0=0.01, cA=cA, Ts=10.0, dte=5.0)
    phi, P, v = ttsh.init_droplet(N, R, xi, pmode, seed)
    o = oracle_mod.Oracle(N, faithful=True, **{k: getattr(p, k) for k in KEYS})
    o.set_state(phi, P, v)
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v)
        if text:
            # the 128^3 run goes through the reference's time loop with text
            # snapshots at T = 5 and 10
            exports, steps = export_schedule(10.0, 5.0, 0.01)
            assert sim.run(str(tmp_path), text=True) == steps
            os.remove(tmp_path / "conf_0.dat")
        else:
            exports, steps = {}, 1000
            sim.step(steps)
        g = sim.get_state()
        red, ok = sim.reductions()
    done = 0
    for m in sorted(exports):
        o.step(exports[m] - done)
        done = exports[m]
        path = tmp_path / f"conf_{m}.dat"
        check_fields(read_conf_text(path, N), o.get_state(), f"N={N} conf_{m}.dat (step {done}) vs oracle")
        os.remove(path)
    if steps > done:
        o.step(steps - done)
    ref = o.get_state()
    ored = o.reductions()
    check_fields(g, ref, f"N={N} {steps} steps")
    log("reductions gpu", red, "oracle", ored)
    assert ok
    assert red[0] == pytest.approx(ored[0], rel=1e-6)
    assert red[1] == pytest.approx(ored[1], rel=1e-4)
    assert red[2] == pytest.approx(ored[2], rel=1e-3)


def test_aster_256_forty_steps(ttsh, oracle_mod):
    """The 256^3 aster workload (R=80, xi=3, noisy aster P, cA=-1) with its
    tuned tiles, 40 steps against the oracle on all host threads."""
    log("[aster_256_forty_steps]")
    oracle_mod.set_threads(min(16, os.cpu_count() or 1))
    N, steps = 256, 40
    p = ttsh.default_params(N=N, dt0=0.01, cA=-1.0)
    phi, P, v = ttsh.init_droplet(N, 80.0, 3.0, 1, 1003)
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v)
        sim.step(steps)
        g = sim.get_state()
        red, ok = sim.reductions()
    o = oracle_mod.Oracle(N, faithful=True, **{k: getattr(p, k) for k in KEYS})
    o.set_state(phi, P, v)
    del phi, P, v
    o.step(steps)
    check_fields(g, o.get_state(), f"N={N} {steps} steps")
    ored = o.reductions()
    log("reductions gpu", red, "oracle", ored)
    assert ok
    assert red[0] == pytest.approx(ored[0], rel=1e-6)
    assert red[2] == pytest.approx(ored[2], rel=1e-4)


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
    exports, steps = export_schedule(0.5, 0.13, 0.01)
    expected = {0} | set(exports)
    got = {int(f[5:-4]) for f in os.listdir(d) if f.startswith("conf_")}
    assert got == expected
    assert f"({steps} steps)" in r.stdout
    lines = open(d / f"conf_{max(expected)}.dat").read().splitlines()
    assert len(lines) == 32 ** 3 and len(lines[0].split(" ")) == 7
