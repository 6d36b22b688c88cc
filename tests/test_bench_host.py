"""Host-side checks of bench.py's contract (no GPU): one GPU only, the two
single-droplet workloads, the algorithmic byte counts, and the CPU-baseline
leg's record."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _run(args, env=None):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                       timeout=120, env={**os.environ, **(env or {})})
    return r.returncode, r.stdout, r.stderr


@pytest.mark.parametrize("args,env", [(["--gpus", "2"], None), ([], {"WORLD_SIZE": "2"})])
def test_more_than_one_gpu_is_refused(args, env):
    rc, out, err = _run(args, env)
    assert rc == 2
    assert out == ""
    assert "one GPU" in err


def test_workloads_match_the_documented_configurations():
    from ttsh3d import workloads
    w = workloads.SINGLE
    assert sorted(w) == [64, 128, 256]
    assert (w[64]["R"], w[64]["xi"], w[64]["pmode"], w[64]["cA"]) == (20.0, 2.0, 0, -1.0)
    assert (w[128]["R"], w[128]["xi"], w[128]["pmode"], w[128]["cA"]) == (40.0, 2.0, 0, -2.0)  # random-unit P
    assert (w[256]["R"], w[256]["xi"], w[256]["pmode"], w[256]["cA"]) == (80.0, 3.0, 1, -1.0)  # aster P


def test_algorithmic_bytes():
    a = bench.ALG_BYTES
    # K_rhs reads phi, P, v and writes phi', v*, P* (14 FP32 fields); K_corr
    # reads p (complex), v*, P* and writes v, P; K_div reads v*, P* and writes
    # the complex s; each FFT pass reads and writes the complex buffer
    assert a["K_rhs"] == 14 * 4 and a["K_corr"] == (2 + 6 + 6) * 4 and a["K_div"] == (6 + 2) * 4
    assert all(a[k] == 16 for k in a if k.startswith("K_fft"))
    assert sum(a.values()) == 224 == bench.STEP_BYTES  # five FFT passes of 16 B/cell
    assert bench.PLAN_BYTES == 176  # the four-stage figure: one forward and one inverse transform
    assert bench.HBM_PEAK_GBS == 8000.0


def test_cpu_baseline_record(ttsh, oracle_mod):
    N = 32
    p = ttsh.default_params(N=N, dt0=0.01, cA=-2.0)
    phi, P, v = ttsh.init_droplet(N, 10.0, 2.0, 0, 7)
    r = bench.cpu_baseline(p, phi, P, v, min_s=0.05, max_steps=3)
    assert set(r) == {"value", "unit", "cores", "kind", "sample"}
    assert r["unit"] == "Mcell-steps/s" and r["kind"] == "port" and r["value"] > 0
    assert r["cores"] >= 1 and "32^3" in r["sample"]
    assert np.isfinite(r["value"])


def test_sweep_mix_is_the_documented_one(ttsh):
    """64 droplets, L in {48, 64, 96, 128}, cA = -1 x logU[0.1, 10],
    La^2 = U[0.5, 2], seeded (seed 1004): the same list on every call."""
    from ttsh3d import workloads
    a = workloads.sweep_mix(ttsh)
    b = workloads.sweep_mix(ttsh)
    assert len(a) == 64 and {L for _, L, _ in a} == {48, 64, 96, 128}
    assert [(p.Nx, p.cA, p.La, s) for p, _, s in a] == [(p.Nx, p.cA, p.La, s) for p, _, s in b]
    for p, L, s in a:
        assert p.Nx == L and -10.0 <= p.cA <= -0.1 and 0.5 <= p.La ** 2 <= 2.0 and p.dt0 == 0.01


def test_kernel_table_and_roofline():
    avg = {k: 0.01 for k in bench.ALG_BYTES}
    avg["K_rhs"] = 0.05
    t = bench.kernel_table(avg, 128, {"K_rhs": 123})
    assert t["K_rhs"]["achieved_GBs"] == pytest.approx(56 * 128 ** 3 / 0.05e-3 / 1e9, rel=1e-4)
    assert t["K_rhs"]["frac"] == pytest.approx(t["K_rhs"]["achieved_GBs"] / 8000.0, abs=1e-4)
    assert t["K_rhs"]["traffic"] == 123 and t["K_div"]["traffic"] is None
    r = bench.roofline({"avg": avg, "kernels": t}, 128)
    assert r["kernel"] == "K_rhs" and r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
