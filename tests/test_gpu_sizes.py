"""GPU parity at every grid size with a compile-time FFT plan beyond those of
test_gpu_parity.py: the radix-3 sizes 96, 144, 192 (plans 8x12, 12x12,
16x12), the radix-5 sizes 80, 160, 240 (plans 10x8, 10x16, 16x15: register
DFTs of length 5, 10 and 15) and the 256^3 droplet (the workload beyond the
Infinity Cache), each against the FP64 oracle after a few steps.  These sizes
also run the tuned K_rhs tiles and z-chunkings the single-droplet workloads
use (tuned.txt).  80^3 also runs in a batch of eight, which switches it to the
fused x+y FFT pass."""
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


@pytest.mark.parametrize("N,steps,pmode,cA", [(80, 10, 0, -1.0), (96, 10, 0, -1.0), (144, 8, 1, -2.0),
                                              (160, 6, 1, -1.5), (192, 6, 0, -1.0), (240, 4, 0, -1.0),
                                              (256, 4, 1, -1.0)])
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


def test_radix5_fused_xy_batch_equals_alone(ttsh):
    """Eight 80^3 droplets (640 z-planes) make the engine run the fused x+y
    FFT pass for that size; alone each runs the single-axis passes.  Same
    per-line arithmetic, so the results are bitwise equal."""
    log("[radix5_fused_xy_batch]")
    N, steps = 80, 6
    f = lambda a: np.asarray(a, np.float32)
    cases = [(-1.0 - 0.2 * i, 1.0 + 0.05 * i, 800 + i) for i in range(8)]
    with ttsh.Batch(max_slots=8, use_graph=True) as b:
        slots = []
        for cA, La, seed in cases:
            p = ttsh.default_params(N=N, dt0=0.01, cA=cA, La=La)
            slots.append(b.admit(p, *[f(x) for x in ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)], dtype=ttsh.F32))
        b.step(steps)
        got = [b.get_state(s) for s in slots]
    for (cA, La, seed), g in zip(cases, got):
        p = ttsh.default_params(N=N, dt0=0.01, cA=cA, La=La)
        with ttsh.Simulation(p) as sim:
            sim.set_state(*[f(x) for x in ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)], dtype=ttsh.F32)
            sim.step(steps)
            ref = sim.get_state()
        for a, c in zip(ref, g):
            assert np.array_equal(a, c), f"80^3 droplet seed {seed} differs between batch and alone"
