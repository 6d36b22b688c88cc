"""The batched-sweep and served workloads at their real sizes, checked against
the FP64 oracle and against each droplet run alone.

  - The sweep mix (64 droplets, L in {48, 64, 96, 128}, seed 1004) runs 100
    steps as one graph batch; it fills the device, so K_rhs runs the batch
    layout (32x8 tiles, z-chunks of 32) and the small sizes the fused x+y
    FFT pass.  One droplet per size matches the oracle (relative RMS 1e-4,
    the bar of every 100-step parity case) and its own plain-step run
    bitwise; every droplet stays finite.
  - A served stream over L in {64, 96, 128, 192} (12 requests, two or more
    of each size, short step counts) completes; every request equals its run
    alone bitwise, and the shortest request of each size matches the oracle
    (relative RMS 2e-5, the bar of the short multi-size cases).
"""
import sys

import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu

ORACLE_KEYS = ("h", "dt0", "mu", "cA", "alpha", "beta", "eta", "La", "kappa", "lambda0",
               "aPhi", "kPhi", "gamma", "A")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_run(O, p, phi, P, v, steps):
    o = O.Oracle(p.Nx, faithful=True, **{k: getattr(p, k) for k in ORACLE_KEYS})
    o.set_state(phi, P, v)
    o.step(steps)
    return o.get_state()


def alone(ttsh, p, phi, P, v, steps):
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v, dtype=ttsh.F32)
        sim.step(steps)
        return sim.get_state()


def errs(got, ref):
    return {k: rel_rms(g, r) for k, g, r in zip(("phi", "P", "v"), got, ref)}


def test_sweep_mix_64_droplets(ttsh, oracle_mod):
    from ttsh3d import workloads
    steps = 100
    cases = workloads.sweep_mix(ttsh)
    log(f"[sweep_mix_64_droplets steps={steps}]", "sizes", sorted({L for _, L, _ in cases}))
    f32 = lambda a: np.asarray(a, np.float32)
    states = [tuple(f32(x) for x in workloads.mix_state(ttsh, L, s)) for _, L, s in cases]
    with ttsh.Batch(max_slots=len(cases), use_graph=True) as b:
        slots = [b.admit(p, *st, dtype=ttsh.F32) for (p, _, _), st in zip(cases, states)]
        b.step(steps)
        got = [b.get_state(s) for s in slots]
        for s in slots:
            assert b.steps_done(s) == steps
    for i, g in enumerate(got):
        assert all(np.isfinite(x).all() for x in g), f"droplet {i} not finite"
    # one droplet per size: oracle and plain-step run alone
    picked = {}
    for i, (_, L, _) in enumerate(cases):
        picked.setdefault(L, i)
    for L, i in sorted(picked.items()):
        p, _, dseed = cases[i]
        phi, P, v = workloads.mix_state(ttsh, L, dseed)
        ref = oracle_run(oracle_mod, p, phi, P, v, steps)
        e = errs(got[i], ref)
        log(f"L={L} droplet {i} cA={p.cA:.3f} La={p.La:.3f} vs oracle", {k: f"{x:.2e}" for k, x in e.items()})
        for k, x in e.items():
            assert x <= 1e-4, f"L={L} droplet {i}: {k} relative RMS {x:.3e}"
        a = alone(ttsh, p, *states[i], steps)
        for name, x, y in zip(("phi", "P", "v"), got[i], a):
            assert np.array_equal(x, y), f"L={L} droplet {i}: {name} differs from its run alone"


def test_served_stream_sizes(ttsh, oracle_mod):
    from ttsh3d import serve
    reqs = serve.make_requests(12, rate=100.0, sizes=(64, 96, 128, 192), steps=(8, 40), seed=2000)
    log("[served_stream_sizes]", [(r.N, r.steps) for r in reqs])
    assert {r.N for r in reqs} == {64, 96, 128, 192}
    be = serve.GpuBackend(max_slots=12, hbm_budget=16 << 30, keep_states=True)
    done, wall = serve.Server(be, macro=16).run(reqs)
    be.close()
    s = serve.summarize(done, wall)
    log("summary", s)
    assert s["completed"] == 12 and s["failed"] == 0 and s["rejected"] == 0
    shortest = {}
    for r in reqs:
        if r.N not in shortest or r.steps < shortest[r.N].steps:
            shortest[r.N] = r
    for r in reqs:
        p = ttsh.default_params(N=r.N, dt0=0.01, cA=r.cA, La=r.La)
        phi, P, v = ttsh.init_droplet(r.N, 0.3 * r.N, 2.0, 0, r.seed)
        f32 = [np.asarray(x, np.float32) for x in (phi, P, v)]
        got = be.final[r.rid]
        ref = alone(ttsh, p, *f32, r.steps)
        for name, x, y in zip(("phi", "P", "v"), got, ref):
            assert np.array_equal(x, y), f"request {r.rid} (N={r.N}): {name} differs from its run alone"
        if shortest[r.N] is r:
            e = errs(got, oracle_run(oracle_mod, p, phi, P, v, r.steps))
            log(f"request {r.rid} N={r.N} steps={r.steps} vs oracle", {k: f"{x:.2e}" for k, x in e.items()})
            for k, x in e.items():
                assert x <= 2e-5, f"request {r.rid} N={r.N}: {k} relative RMS {x:.3e}"
