"""GPU parity over seeded random parameter sets: every knob of the step
(mu, cA, La, alpha, beta, eta, lambda0, aPhi, kPhi, gamma, A, h) away from its
run_cuda.pl default at once, random-unit and aster P, a non-zero initial flow,
20 steps against the FP64 oracle.  A coefficient slip in any term whose
default value hides it (A = 0, h = 1, ...) shows here.

The ranges keep the explicit step stable at dt = 0.01: the Cahn-Hilliard
bound scales as h^4 / (gamma kPhi), 0.0189 at the defaults, so
gamma * kPhi / h^4 <= 1.5 keeps it above 0.0126."""
import sys

import numpy as np
import pytest

from conftest import max_rel, rel_rms

pytestmark = pytest.mark.gpu

KEYS = ("h", "dt0", "mu", "cA", "alpha", "beta", "eta", "La", "kappa", "lambda0", "aPhi", "kPhi", "gamma", "A")
# relative RMS bar per field after 20 steps (the step's target is 1e-4 after 1,000)
BAR = 5e-5


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_case(ttsh, k):
    rng = np.random.default_rng(5000 + k)
    N = (32, 48)[k % 2]
    h = float(rng.choice([1.0, 1.2]))
    gamma = float(rng.uniform(0.5, 1.5))
    kPhi = float(rng.uniform(0.5, min(2.0, 1.5 * h ** 4 / gamma)))
    kw = dict(h=h, gamma=gamma, kPhi=kPhi, mu=float(rng.uniform(0.5, 2.0)), cA=float(rng.uniform(-3.0, 1.0)),
              La=float(rng.uniform(0.5, 2.0)), alpha=float(rng.uniform(-0.1, 0.1)), beta=float(rng.uniform(1.0, 6.0)),
              eta=float(rng.uniform(-0.5, -0.05)), lambda0=float(rng.uniform(0.0, 12.0)),
              aPhi=float(rng.uniform(0.5, 2.0)), A=0.0 if k % 3 == 0 else float(rng.uniform(0.2, 0.8)))
    p = ttsh.default_params(N=N, dt0=0.01, **kw)
    phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 1 if k % 3 == 1 else 0, 6000 + k)  # aster for k = 1, 4
    idx = np.arange(v.size, dtype=np.float64).reshape(v.shape)
    v = 0.02 * np.sin(idx * 0.0013 + k)
    return p, phi, P, v, kw


@pytest.mark.parametrize("k", range(6))
def test_random_parameters(ttsh, oracle_mod, k):
    p, phi, P, v, kw = random_case(ttsh, k)
    log(f"[random_parameters {k}] N={p.Nx}", {a: round(b, 3) for a, b in kw.items()})
    o = oracle_mod.Oracle(p.Nx, faithful=True, **{a: getattr(p, a) for a in KEYS})
    o.set_state(phi, P, v)
    o.step(20)
    ref = o.get_state()
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v)
        sim.step(20)
        g = sim.get_state()
        red, ok = sim.reductions()
    errs = [rel_rms(a, b) for a, b in zip(g, ref)]
    log(f"rel RMS phi {errs[0]:.2e} P {errs[1]:.2e} v {errs[2]:.2e}; "
        f"max rel {max_rel(g[0], ref[0]):.2e} {max_rel(g[1], ref[1]):.2e} {max_rel(g[2], ref[2]):.2e}")
    assert ok
    assert max(errs) <= BAR
    assert red[0] == pytest.approx(o.reductions()[0], rel=1e-6)
