"""Resume from a state: N steps equal N/2 steps, a copy of the state out and
into a fresh simulation, and N/2 more, bitwise.

The state is phi, P, v and the last step's predictors v*, P* (the stale halo
of the next divergence reads them; TTSH3D.cu:366-371,405-407).  Nothing else
carries over between steps: the scale of the packed projection is a function
of the current step, and the edge columns of the previous predictors are
rebuilt from the uploaded ones.  The reference's own restart path cannot do
this (its snapshots lack v*, P* and the exact T; TTSH3D.cu:768-797)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("faithful,pmode", [(1, 1), (0, 0)])
def test_resume_is_bitwise(ttsh, faithful, pmode):
    log(f"[resume faithful={faithful} pmode={pmode}]")
    N, half = 48, 12  # 48 = 1.5 tiles of 32 in x: partial tiles and both edge columns in one tile
    p = ttsh.default_params(N=N, dt0=0.01, cA=-2.0)
    p.faithful_halo = faithful
    phi, P, v = ttsh.init_droplet(N, 15.0, 2.0, pmode, 4242)
    with ttsh.Simulation(p) as a:
        a.set_state(phi, P, v)
        a.step(2 * half)
        ref = a.get_state()
    with ttsh.Simulation(p) as b:
        b.set_state(phi, P, v)
        b.step(half)
        mid = b.get_state()
        vs, ps = b.aux("vstar"), b.aux("pstar")
    with ttsh.Simulation(p) as c:
        c.set_state(*mid, vs, ps, dtype=ttsh.F32)
        c.step(half)
        got = c.get_state()
    for name, x, y in zip(("phi", "P", "v"), got, ref):
        log(name, "max |diff|", float(np.max(np.abs(x - y))))
        assert np.array_equal(x, y), name
    # the same resume into a batch slot
    with ttsh.Batch(max_slots=2) as bt:
        s = bt.admit(p, *mid, vs, ps, dtype=ttsh.F32)
        bt.step(half)
        got_b = bt.get_state(s)
    for name, x, y in zip(("phi", "P", "v"), got_b, ref):
        assert np.array_equal(x, y), f"batch {name}"
