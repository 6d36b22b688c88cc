"""GPU parity at every grid size with a compile-time FFT plan beyond those of
test_gpu_parity.py (32, 48, 64, 128): the radix-3 sizes 36, 72, 96, 144, 192
(plans 6x6, 12x6, 8x12, 12x12, 16x12), the sizes with a factor 5 (40, 50,
60, 80, 90, 100, 120, 150, 160, 180, 200, 240: register DFTs of length 5,
10, 15 and 20) and the 256^3 droplet (the workload beyond the Infinity
Cache), each against the FP64 oracle after a few steps.  36, 50, 60, 90, 100,
150 and 180 are not multiples of 8, so their K_rhs, K_div and K_corr tiles
are partial in y as well as in x.  These sizes also run the tuned K_rhs tiles
and z-chunkings of tuned.txt.  Every size up to 128 also runs in a batch
holding at least two z-planes per CU, which switches it to the fused x+y FFT
pass, and must equal the same droplets stepped alone bitwise."""
import os
import sys

import numpy as np
import pytest

from conftest import max_rel, rel_rms

pytestmark = pytest.mark.gpu

KEYS = ("h", "dt0", "mu", "cA", "alpha", "beta", "eta", "La", "kappa", "lambda0", "aPhi", "kPhi", "gamma", "A")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N,steps,pmode,cA", [(36, 10, 0, -1.0), (40, 10, 1, -1.5), (50, 10, 0, -2.0),
                                              (60, 10, 1, -1.0), (72, 10, 0, -1.5), (80, 10, 0, -1.0),
                                              (90, 8, 1, -1.0), (96, 10, 0, -1.0), (100, 8, 0, -1.5),
                                              (120, 8, 1, -1.0), (144, 8, 1, -2.0), (150, 6, 0, -1.0),
                                              (160, 6, 1, -1.5), (180, 6, 0, -1.0), (192, 6, 0, -1.0),
                                              (200, 4, 1, -1.0), (240, 4, 0, -1.0), (256, 4, 1, -1.0)])
def test_sizes_against_oracle(ttsh, oracle_mod, N, steps, pmode, cA):
    log(f"[sizes N={N} steps={steps}]")
    oracle_mod.set_threads(min(16, os.cpu_count() or 1))
    p = ttsh.default_params(N=N, dt0=0.01, cA=cA)
    phi, P, v = ttsh.init_droplet(N, 0.3125 * N, 2.0 if N < 256 else 3.0, pmode, 2000 + N)
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v)
        sim.step(steps)
        g = sim.get_state()
        red, ok = sim.reductions()
    o = oracle_mod.Oracle(N, faithful=True, **{k: getattr(p, k) for k in KEYS})
    o.set_state(phi, P, v)
    o.step(steps)
    ref = o.get_state()
    ored = o.reductions()
    errs = [rel_rms(a, b) for a, b in zip(g, ref)]
    log(f"N={N} {steps} steps rel RMS phi {errs[0]:.2e} P {errs[1]:.2e} v {errs[2]:.2e}; "
        f"max rel {max_rel(g[0], ref[0]):.2e} {max_rel(g[1], ref[1]):.2e} {max_rel(g[2], ref[2]):.2e}")
    assert ok
    assert max(errs) <= 2e-5
    assert red[0] == pytest.approx(ored[0], rel=1e-6)
    assert red[2] == pytest.approx(ored[2], rel=1e-4)


@pytest.mark.parametrize("N", [36, 40, 50, 60, 72, 80, 90, 100, 120])
def test_fused_xy_batch_equals_alone(ttsh, N):
    """ceil(512 / N) droplets of size N (at least 512 z-planes, two per CU)
    make the engine run the fused x+y FFT pass for that size; alone each
    runs the single-axis passes.  Same per-line arithmetic, so the first and
    the last droplet are bitwise equal to themselves stepped alone."""
    log(f"[fused_xy_batch N={N}]")
    steps, k = 4, -(-512 // N)
    f = lambda a: np.asarray(a, np.float32)
    cases = [(-1.0 - 0.1 * i, 1.0 + 0.02 * i, 800 + i) for i in range(k)]
    with ttsh.Batch(max_slots=k, use_graph=True) as b:
        slots = []
        for cA, La, seed in cases:
            p = ttsh.default_params(N=N, dt0=0.01, cA=cA, La=La)
            slots.append(b.admit(p, *[f(x) for x in ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)], dtype=ttsh.F32))
        b.step(steps)
        got = {i: b.get_state(slots[i]) for i in (0, k - 1)}
    for i, g in got.items():
        cA, La, seed = cases[i]
        p = ttsh.default_params(N=N, dt0=0.01, cA=cA, La=La)
        with ttsh.Simulation(p) as sim:
            sim.set_state(*[f(x) for x in ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)], dtype=ttsh.F32)
            sim.step(steps)
            ref = sim.get_state()
        assert all(np.isfinite(a).all() for a in g)
        for a, c in zip(ref, g):
            assert np.array_equal(a, c), f"{N}^3 droplet seed {seed} differs between batch and alone"
