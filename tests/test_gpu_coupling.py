"""The closed forms of the right-hand side's coupling terms on the HIP step.

The cases of test_oracle_coupling.py (advection, Cahn-Hilliard chemical
potential and capillary force, active stress with a non-uniform phi, the
molecular field, lambda_0 self-advection, anchoring, the zero co-rotation
terms; TTSH3D.cu:561-632) are built on a 32^3 grid and stepped once through
ttsh_step.  The predictors v* = v + Dt dt v and P* = P + Dt dt P (ttsh_get_aux)
and phi^{n+1} = phi + Dt dt phi then give the GPU's dt v, dt P and dt phi,
which must match the exact values.

Tolerance: the state is uploaded in FP32 and the step computes in FP32, so
(x* - x)/Dt carries the rounding of x (about 6e-8 |x|) divided by Dt = 0.01;
the bar is 1e-4 of max |dt x| + 2e-6 max |x| / Dt (about 30 times that
rounding), while a wrong sign, coefficient or index changes dt x by O(1) of
its size.  The stale halo is off: it only enters the divergence, not these
stage outputs."""
import sys

import numpy as np
import pytest

import test_oracle_coupling as T

pytestmark = pytest.mark.gpu

NG = 32


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("build", T.BUILDERS, ids=[b.__name__[len("build_"):] for b in T.BUILDERS])
def test_coupling_closed_form_on_gpu(ttsh, build):
    name = build.__name__[len("build_"):]
    log(f"[gpu coupling {name}]")
    c = build(NG)
    p = ttsh.default_params(N=NG, dt0=T.DT, h=T.H, **c["params"])
    p.faithful_halo = 0
    f32 = lambda a: np.asarray(a, np.float32)
    phi, P, v = f32(c["phi"]), f32(c["P"]), f32(c["v"])
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v, dtype=ttsh.F32)
        sim.step(1)
        phi1, _, _ = sim.get_state()
        vs, ps = sim.aux("vstar"), sim.aux("pstar")
    got = {"dtphi": (phi1.astype(np.float64) - phi) / T.DT,
           "dtv": (vs.astype(np.float64) - v) / T.DT,
           "dtP": (ps.astype(np.float64) - P) / T.DT}
    base = {"dtphi": phi, "dtv": v, "dtP": P}
    for key in ("dtphi", "dtv", "dtP"):
        exp = c[key]
        err = float(np.abs(got[key] - exp).max())
        scale = float(np.abs(exp).max())
        bar = 1e-4 * scale + 2e-6 * float(np.abs(base[key]).max()) / T.DT + 1e-6
        log(f"{name} {key}: max |dt - exact| {err:.3e}, max |exact| {scale:.3e}, bar {bar:.3e}")
        assert err <= bar, f"{name}: {key} differs from its closed form by {err:.3e} (bar {bar:.3e})"
