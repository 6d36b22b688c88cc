"""The benchmark workloads, shared by bench.py and the tests.

Single droplets (periodic box, h = 1, dt = 0.01, run_cuda.pl:30-44 defaults
otherwise):
  - 64^3:  R = 20, xi = 2, seeded random-unit P inside phi > 0.5;
  - 128^3: R = 40, xi = 2, cA = 2 x default (-2), random-unit P;
  - 256^3: R = 80, xi = 3, noisy aster P (r_hat + 0.1 N(0,1), renormalised).
The sweep mix: 64 droplets, L uniform over {48, 64, 96, 128}, R = 0.3 L,
cA = default x logU[0.1, 10], La^2 = default x U[0.5, 2], random-unit P, all
from seed 1004.  The served stream: seeded Poisson arrivals of requests with
L uniform over {64, 96, 128, 192}, 1,000 to 20,000 steps, cA and La^2 as in
the mix (serve.make_requests).
"""
from __future__ import annotations

import math

import numpy as np

SINGLE = {
    64: dict(R=20.0, xi=2.0, pmode=0, cA=-1.0, seed=1001,
             desc="single droplet 64^3, R=20, xi=2, random-unit P, dt=0.01"),
    128: dict(R=40.0, xi=2.0, pmode=0, cA=-2.0, seed=1002,
              desc="single droplet 128^3, R=40, xi=2, cA=-2, random-unit P, dt=0.01"),
    256: dict(R=80.0, xi=3.0, pmode=1, cA=-1.0, seed=1003,
              desc="single droplet 256^3, R=80, xi=3, aster P (r_hat + 0.1 N(0,1)), dt=0.01"),
}

MIX_SIZES = (48, 64, 96, 128)
MIX_SEED = 1004
SERVE_SIZES = (64, 96, 128, 192)
SERVE_SEED = 1005


def single(ttsh, N):
    """(params, phi, P, v) of the single-droplet workload of size N (FP64 host arrays)."""
    w = SINGLE[N]
    p = ttsh.default_params(N=N, dt0=0.01, cA=w["cA"])
    phi, P, v = ttsh.init_droplet(N, w["R"], w["xi"], w["pmode"], w["seed"])
    return p, phi, P, v


def sweep_mix(ttsh, ndrop=64, seed=MIX_SEED):
    """The batched-sweep mix: a list of (params, L, droplet seed); the state of
    droplet i is init_droplet(L, 0.3 L, 2, 0, droplet seed)."""
    rng = np.random.default_rng(seed)
    Ls = rng.choice(list(MIX_SIZES), size=ndrop)
    base = ttsh.default_params()
    cases = []
    for i, L in enumerate(Ls):
        cA = base.cA * math.exp(rng.uniform(math.log(0.1), math.log(10.0)))
        La2 = base.La ** 2 * rng.uniform(0.5, 2.0)
        p = ttsh.default_params(N=int(L), dt0=0.01, cA=cA, La=math.sqrt(La2))
        cases.append((p, int(L), seed * 1000 + i))
    return cases


def mix_state(ttsh, L, dseed):
    return ttsh.init_droplet(L, 0.3 * L, 2.0, 0, dseed)
