"""GPU tests of batching and serving: results are the same alone and in a
mix (bitwise: the kernels are deterministic and per-droplet), the step
limits hold, and a diverging droplet is flagged without touching its
neighbours."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def alone(ttsh, N, cA, La, seed, steps, dt0=0.01):
    p = ttsh.default_params(N=N, dt0=dt0, cA=cA, La=La)
    phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi.astype(np.float32), P.astype(np.float32), v.astype(np.float32), dtype=ttsh.F32)
        sim.step(steps)
        return sim.get_state()


def test_served_results_equal_alone(ttsh):
    from ttsh3d import serve
    log("[served_results_equal_alone]")
    reqs = serve.make_requests(10, rate=40.0, sizes=(32, 48), steps=(20, 90), seed=11)
    be = serve.GpuBackend(max_slots=6, hbm_budget=3 * ttsh.droplet_bytes(48), keep_states=True)
    done, wall = serve.Server(be, macro=16).run(reqs)
    be.close()
    s = serve.summarize(done, wall)
    log("summary", s)
    assert s["completed"] == 10 and s["failed"] == 0
    for r in reqs:
        ref = alone(ttsh, r.N, r.cA, r.La, r.seed, r.steps)
        got = be.final[r.rid]
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), f"request {r.rid} differs from its run alone"


def test_step_limits_and_divergence_isolation(ttsh):
    log("[step_limits_and_divergence_isolation]")
    good = ttsh.default_params(N=32, dt0=0.01)
    bad = ttsh.default_params(N=32, dt0=0.05)  # beyond the explicit stability bound ~0.019
    phi, P, v = ttsh.init_droplet(32, 0.3 * 32, 2.0, 0, 5)
    f = lambda a: a.astype(np.float32)
    with ttsh.Batch(max_slots=3, use_graph=True) as b:
        sg = b.admit(good, f(phi), f(P), f(v), dtype=ttsh.F32, limit=37)
        sb = b.admit(bad, f(phi), f(P), f(v), dtype=ttsh.F32, limit=300)
        b.step(320)
        assert b.steps_done(sg) == 37 and b.steps_done(sb) == 300
        red_bad, ok_bad = b.reductions(sb)
        red_good, ok_good = b.reductions(sg)
        log("bad", red_bad, ok_bad, "good", red_good, ok_good)
        assert not ok_bad and ok_good
        got = b.get_state(sg)
    ref = alone(ttsh, 32, good.cA, good.La, 5, 37)
    for a, c in zip(ref, got):
        assert np.array_equal(a, c)


def test_large_batch_fused_xy_equals_alone(ttsh):
    """Eight 64^3 droplets give 512 z-planes, enough for the engine to switch
    that size to the fused x+y FFT pass; alone, each droplet runs the
    single-axis passes.  Both perform the same per-line arithmetic, so the
    results are bitwise equal."""
    log("[large_batch_fused_xy_equals_alone]")
    N, steps = 64, 12
    cases = [(-1.0 * (1 + 0.25 * i), 1.0 + 0.1 * i, 300 + i) for i in range(8)]
    f = lambda a: a.astype(np.float32)
    with ttsh.Batch(max_slots=8, use_graph=True) as b:
        slots = []
        for cA, La, seed in cases:
            p = ttsh.default_params(N=N, dt0=0.01, cA=cA, La=La)
            phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)
            slots.append(b.admit(p, f(phi), f(P), f(v), dtype=ttsh.F32))
        b.step(steps)
        got = [b.get_state(s) for s in slots]
    for (cA, La, seed), g in zip(cases, got):
        ref = alone(ttsh, N, cA, La, seed, steps)
        for a, c in zip(ref, g):
            if not np.array_equal(a, c):
                d = np.abs(a.astype(np.float64) - c) / (np.abs(a).max() + 1e-30)
                log(f"seed {seed}: max scaled diff {d.max():.3e}")
            assert np.array_equal(a, c), f"droplet seed {seed} differs between batch and alone"


def test_sweep_equals_runs_alone(ttsh, tmp_path):
    """A forEach.pl grid of 3 points run as one batch writes, per point, the
    same conf_<m>.dat bytes as that point's own ttsh_run."""
    from ttsh3d import sweep
    log("[sweep_equals_runs_alone]")
    pts = sweep.grid({"cA": [-1, -0.5], "La": [1, 1.5]}, {"N": 32, "Tend": 0.3, "dT": 0.01, "dTe": 0.1})[:3]
    steps = sweep.run_sweep(pts, str(tmp_path / "sweep"), seed=40, log=log)
    for i, (name, p) in enumerate(pts):
        d = tmp_path / "alone" / name
        d.mkdir(parents=True)
        phi, P, v, vs, ps = ttsh.init_reference(p, 40 + i)
        with ttsh.Simulation(p) as sim:
            sim.set_state(phi, P, v, vs, ps)
            assert sim.run(str(d)) == steps
        for f in sorted(os.listdir(d)):
            assert open(d / f, "rb").read() == open(tmp_path / "sweep" / name / f, "rb").read(), (name, f)
        assert os.path.exists(tmp_path / "sweep" / name / "input.dat")


def test_device_filling_batch_layout_equals_alone(ttsh):
    """Six 128^3 and two 96^3 droplets make 1,752 K_rhs items at the batch
    layout (32x8 tiles, z-chunks of 32), more than two rounds of three groups
    on 256 CUs, so the engine runs K_rhs that way and K_corr with chunks of 32.
    Alone, each droplet runs its tuned 32x16 tile with chunks of 8 or 16
    planes and K_corr with 8 or 16.  Neither kernel reduces anything, so the
    results are bitwise equal."""
    log("[device_filling_batch_layout_equals_alone]")
    steps = 6
    cases = [(128, -1.0 * (1 + 0.3 * i), 1.0 + 0.1 * i, 700 + i) for i in range(6)]
    cases += [(96, -1.5, 1.2, 710), (96, -0.7, 0.9, 711)]
    f = lambda a: a.astype(np.float32)
    with ttsh.Batch(max_slots=len(cases), use_graph=True) as b:
        slots = []
        for N, cA, La, seed in cases:
            p = ttsh.default_params(N=N, dt0=0.01, cA=cA, La=La)
            phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)
            slots.append(b.admit(p, f(phi), f(P), f(v), dtype=ttsh.F32))
        b.step(steps)
        got = [b.get_state(s) for s in slots]
    for (N, cA, La, seed), g in zip(cases, got):
        ref = alone(ttsh, N, cA, La, seed, steps)
        for a, c in zip(ref, g):
            assert np.array_equal(a, c), f"{N}^3 droplet seed {seed} differs between batch and alone"


def test_layout_switch_mid_run_equals_alone(ttsh):
    """Six 128^3 and two 96^3 droplets start in the device-filling layout
    (1,752 K_rhs items at depth 32).  Three 128^3 droplets stop after 8 steps,
    leaving 984 items, so the engine returns to the tuned layout (and
    re-captures its graphs) while the others run on to 40 steps.  Every
    droplet still equals its run alone, bitwise."""
    log("[layout_switch_mid_run_equals_alone]")
    cases = [(128, -1.0 - 0.2 * i, 1.0 + 0.05 * i, 720 + i, 8 if i < 3 else 40) for i in range(6)]
    cases += [(96, -1.2, 1.1, 730, 40), (96, -0.8, 0.9, 731, 40)]
    f = lambda a: a.astype(np.float32)
    with ttsh.Batch(max_slots=len(cases), use_graph=True) as b:
        slots = []
        for N, cA, La, seed, lim in cases:
            p = ttsh.default_params(N=N, dt0=0.01, cA=cA, La=La)
            phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)
            slots.append(b.admit(p, f(phi), f(P), f(v), dtype=ttsh.F32, limit=lim))
        b.step(40)
        for s, case in zip(slots, cases):
            assert b.steps_done(s) == case[4]
        got = [b.get_state(s) for s in slots]
    for (N, cA, La, seed, lim), g in zip(cases, got):
        ref = alone(ttsh, N, cA, La, seed, lim)
        for a, c in zip(ref, g):
            assert np.array_equal(a, c), f"{N}^3 droplet seed {seed} ({lim} steps) differs between batch and alone"


def test_ready_wait_and_admission_while_steps_run(ttsh):
    """ttsh_batch_ready turns true once a droplet's limit is reached and those
    steps are done; its state then stays untouched by later steps.  A droplet
    admitted while steps are queued (uploaded on the side stream, published
    at the next step call) and a retired droplet's reused memory both give
    the run-alone result, bitwise."""
    log("[ready_wait_and_admission_while_steps_run]")
    f = lambda a: a.astype(np.float32)
    pa = ttsh.default_params(N=48, dt0=0.01, cA=-1.5)
    pb = ttsh.default_params(N=32, dt0=0.01)
    pc = ttsh.default_params(N=48, dt0=0.01, La=1.2)
    sa, sb, sc = (ttsh.init_droplet(48, 14.4, 2.0, 0, 91), ttsh.init_droplet(32, 9.6, 2.0, 0, 92),
                  ttsh.init_droplet(48, 14.4, 2.0, 0, 93))
    with ttsh.Batch(max_slots=4, use_graph=True) as b:
        a = b.admit(pa, *map(f, sa), dtype=ttsh.F32, limit=20)
        bb = b.admit(pb, *map(f, sb), dtype=ttsh.F32)
        b.step(32)
        b.wait(0)
        assert b.ready(a) and not b.ready(bb)
        got_a = b.get_state(a)
        b.step(64)           # queued; droplet a is past its limit
        assert b.ready(a)
        assert all(np.array_equal(x, y) for x, y in zip(got_a, b.get_state(a)))
        b.retire(a)          # its memory is kept for the next 48^3 droplet
        c = b.admit(pc, *map(f, sc), dtype=ttsh.F32, limit=24)  # admitted while steps run
        b.step(32)
        b.wait(1)
        b.synchronize()
        assert b.ready(c) and b.steps_done(c) == 24 and b.steps_done(bb) == 128
        got_c = b.get_state(c)
        got_b = b.get_state(bb)
    for got, n, p, st in ((got_a, 20, pa, sa), (got_c, 24, pc, sc), (got_b, 128, pb, sb)):
        for x, y in zip(alone_f32(ttsh, p, st, n), got):
            assert np.array_equal(x, y), f"N={p.Nx}, {n} steps"


def alone_f32(ttsh, p, st, steps):
    with ttsh.Simulation(p) as sim:
        sim.set_state(*[x.astype(np.float32) for x in st], dtype=ttsh.F32)
        sim.step(steps)
        return sim.get_state()
