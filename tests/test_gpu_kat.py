"""Closed-form checks of the FP32 GPU step, with no oracle in the loop.

Each expected value is fixed by the mathematics of the reference's
discretisation: the stencils of TTSH3D.cu:3280-3343 (D_a = (1/2h) delta_a
(x) S (x) S with S = [1,4,1]/6, and the 15-point Laplacian), the right-hand
side of :537-632 and the projection of :635-691.  In Fourier terms a mode
e^{i k.x} has
    D_a -> i d_a,  d_a = sin(k_a h)/h S(k_b) S(k_c),  S(q) = (2 + cos qh)/3
    L   -> Lhat = ((4/3)(cx + cy + cz) + (2/3) cx cy cz - 14/3)/h^2,  c = cos(k h)
and the projection (p = FFT^-1(-FFT(D.X*/dt0)/|k|^2), X = X* - Dt D p)
removes (Dt/dt0) d (d.X*)/|k|^2.

* A single mode of P with phi = v = 0 and beta = lambda0 = A = 0:
  dtP = eta La^2 L L P, so P* = P (1 + Dt eta La^2 Lhat^2), then projected.
  v* = 0, so the packed projection has no real part and its scale takes
  its largest value.
* A transverse single mode of v (real amplitude a with a.d = 0) with
  phi = P = 0: v.Dv = -(a.d) a cos sin = 0 and D.v* = 0, so
  v^{n+1} = v (1 + Dt mu Lhat).
* A uniform state: every stencil of a constant is zero, so
  P^{n+1} = P + Dt (-alpha phi P - beta |P|^2 P), and v and phi keep their
  values.
* Invariances: periodic shifts, the x/y swap, kappa (the integer-division
  zeros of :579-585 make it inert), v* affine in cA, phi mass from rest.

The stale halo is off here (faithful_halo = 0): it reads the previous
predictor across the box faces, which breaks translation invariance by
design.  The parity tests cover it against the oracle."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, H, DT = 32, 1.1, 0.01


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def grid(n):
    return np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")  # z, y, x


def symbols(m, n=N, h=H):
    k = np.array([2 * np.pi * mi / (n * h) for mi in m])
    S = lambda q: (2 + np.cos(q * h)) / 3
    d = np.array([np.sin(k[0] * h) / h * S(k[1]) * S(k[2]),
                  np.sin(k[1] * h) / h * S(k[0]) * S(k[2]),
                  np.sin(k[2] * h) / h * S(k[0]) * S(k[1])])
    c = np.cos(k * h)
    lap = ((4 / 3) * c.sum() + (2 / 3) * c.prod() - 14 / 3) / h ** 2
    z, y, x = grid(n)
    th = k[0] * x * h + k[1] * y * h + k[2] * z * h
    return k, d, lap, th


def params(ttsh, n=N, **kw):
    p = ttsh.default_params(N=n, h=H, dt0=DT, **kw)
    p.faithful_halo = 0
    return p


def step(ttsh, p, phi, P, v, steps=1):
    with ttsh.Simulation(p) as s:
        s.set_state(phi, P, v)
        s.step(steps)
        out = s.get_state(ttsh.F64)
        vs = s.aux("vstar")
        red, ok = s.reductions()
    assert ok
    return out, vs, red


@pytest.mark.parametrize("m", [(1, 2, 3), (5, 0, -2), (7, -7, 1)])
def test_single_mode_polarization(ttsh, m):
    log(f"[single_mode_polarization m={m}]")
    k, d, lap, th = symbols(m)
    rng = np.random.default_rng(sum(abs(q) for q in m))
    a = 0.2 * (rng.standard_normal(3) + 1j * rng.standard_normal(3))
    P = np.stack([np.real(a[c] * np.exp(1j * th)) for c in range(3)])
    eta, La = -0.187, 1.3
    p = params(ttsh, eta=eta, La=La, beta=0.0, lambda0=0.0, A=0.0)
    zero = np.zeros((N, N, N))
    (phi1, P1, v1), _, _ = step(ttsh, p, zero, P, np.zeros((3, N, N, N)))
    astar = a * (1 + DT * eta * La ** 2 * lap ** 2)
    anew = astar - d * (d @ astar) / float(k @ k)
    exp = np.stack([np.real(anew[c] * np.exp(1j * th)) for c in range(3)])
    amp = np.abs(P).max()
    err = np.abs(P1 - exp).max() / amp
    log(f"max |P - closed form| / amplitude {err:.2e}; step change {np.abs(exp - P).max() / amp:.2e}; "
        f"max |v| {np.abs(v1).max():.1e}")
    assert err < 1e-5
    assert np.abs(exp - P).max() / amp > 1e-2  # the step does change P
    assert np.all(phi1 == 0.0)
    assert np.abs(v1).max() < 1e-9


@pytest.mark.parametrize("m", [(8, -6, 5), (3, 1, -9)])
def test_transverse_velocity_mode(ttsh, m):
    log(f"[transverse_velocity_mode m={m}]")
    k, d, lap, th = symbols(m)
    a = np.cross(d, np.array([0.3, 0.9, -0.4]))
    a = 0.05 * a / np.linalg.norm(a)
    v = np.stack([a[c] * np.cos(th) for c in range(3)])
    mu = 5.0
    p = params(ttsh, mu=mu)
    zero = np.zeros((N, N, N))
    (phi1, P1, v1), _, _ = step(ttsh, p, zero, np.zeros((3, N, N, N)), v)
    exp = v * (1 + DT * mu * lap)
    amp = np.abs(v).max()
    err = np.abs(v1 - exp).max() / amp
    log(f"max |v - closed form| / amplitude {err:.2e}; step change {np.abs(exp - v).max() / amp:.2e}")
    assert err < 1e-5
    assert np.abs(exp - v).max() / amp > 1e-2
    assert np.all(phi1 == 0.0)
    assert np.abs(P1).max() < 1e-9


def test_uniform_state(ttsh):
    log("[uniform_state]")
    phi0, P0, v0 = 0.7, np.array([0.3, -0.5, 0.2]), np.array([0.1, 0.02, -0.04])
    alpha, beta = -0.02, 4.366
    p = params(ttsh, alpha=alpha, beta=beta, A=0.4)
    phi = np.full((N, N, N), phi0)
    P = np.ones((3, N, N, N)) * P0[:, None, None, None]
    v = np.ones((3, N, N, N)) * v0[:, None, None, None]
    (phi1, P1, v1), _, _ = step(ttsh, p, phi, P, v)
    P0f = P0.astype(np.float32).astype(np.float64)
    phi0f = float(np.float32(phi0))
    expP = P0f + DT * (-alpha * phi0f * P0f - beta * (P0f @ P0f) * P0f)
    errs = (np.abs(P1 - expP[:, None, None, None]).max(), np.abs(v1 - v0[:, None, None, None]).max(),
            np.abs(phi1 - phi0).max())
    log(f"max errors P {errs[0]:.1e} v {errs[1]:.1e} phi {errs[2]:.1e}")
    assert errs[0] < 1e-6 and errs[1] < 1e-7 and errs[2] < 1e-6


def droplet(ttsh, n, seed):
    phi, P, _ = ttsh.init_droplet(n, 0.3 * n, 2.0, 0, seed)
    rng = np.random.default_rng(seed)
    v = 0.02 * rng.standard_normal((3, n, n, n))
    return phi, P, v


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


def test_shift_equivariance(ttsh):
    n = 48  # 1.5 tiles of 32 in x: shifted cells change tiles and tile positions
    log(f"[shift_equivariance N={n}]")
    phi, P, v = droplet(ttsh, n, 31)
    sh = (5, -3, 7)
    roll = lambda f: np.roll(f, sh, axis=(-3, -2, -1))
    p = params(ttsh, n=n, cA=-2.0)
    a, _, _ = step(ttsh, p, phi, P, v, steps=3)
    b, _, _ = step(ttsh, p, roll(phi), roll(P), roll(v), steps=3)
    errs = [rel_rms(y, roll(x)) for x, y in zip(a, b)]
    log("rel RMS phi %.1e P %.1e v %.1e" % tuple(errs))
    assert max(errs) < 1e-5


def test_axis_swap_symmetry(ttsh):
    n = 48
    log(f"[axis_swap_symmetry N={n}]")
    phi, P, v = droplet(ttsh, n, 32)
    sw = lambda f: np.swapaxes(f, -1, -2)
    swv = lambda F: np.stack([sw(F[1]), sw(F[0]), sw(F[2])])
    p = params(ttsh, n=n, cA=-2.0, A=0.3)
    a, _, _ = step(ttsh, p, phi, P, v, steps=3)
    b, _, _ = step(ttsh, p, sw(phi), swv(P), swv(v), steps=3)
    errs = [rel_rms(b[0], sw(a[0])), rel_rms(b[1], swv(a[1])), rel_rms(b[2], swv(a[2]))]
    log("rel RMS phi %.1e P %.1e v %.1e" % tuple(errs))
    assert max(errs) < 1e-5


def test_kappa_is_inert(ttsh):
    log("[kappa_is_inert]")
    phi, P, v = droplet(ttsh, N, 33)
    a, _, _ = step(ttsh, params(ttsh, kappa=-0.364), phi, P, v, steps=2)
    b, _, _ = step(ttsh, params(ttsh, kappa=7.5), phi, P, v, steps=2)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_vstar_affine_in_activity(ttsh):
    log("[vstar_affine_in_activity]")
    phi, P, v = droplet(ttsh, N, 34)
    vs = {cA: step(ttsh, params(ttsh, cA=cA), phi, P, v)[1].astype(np.float64) for cA in (0.0, -1.0, 2.5)}
    d1 = vs[-1.0] - vs[0.0]
    d2 = vs[2.5] - vs[0.0]
    ulp = np.spacing(np.float32(np.abs(vs[0.0]).max()))
    dev = np.abs(d2 + 2.5 * d1).max()
    log(f"max |d2 + 2.5 d1| = {dev:.2e} ({dev / ulp:.1f} ulp of max |v*|); max |d1| = {np.abs(d1).max():.2e}")
    assert dev < 8 * ulp
    assert np.abs(d1).max() > 100 * ulp


def test_mass_conserved_from_rest(ttsh):
    log("[mass_conserved_from_rest]")
    phi, P, _ = droplet(ttsh, N, 35)
    v = np.zeros((3, N, N, N))
    _, _, red = step(ttsh, params(ttsh), phi, P, v)
    m0 = float(phi.astype(np.float32).astype(np.float64).sum()) * H ** 3  # mass = sum(phi) h^3
    log(f"mass {m0:.9f} -> {red[0]:.9f} (rel {abs(red[0] - m0) / m0:.1e})")
    assert red[0] == pytest.approx(m0, rel=1e-6)
