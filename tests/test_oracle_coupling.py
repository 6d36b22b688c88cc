"""Closed-form pins of the oracle's right-hand side (CPU only).

getTempVari_2 and getDtF (TTSH3D.cu:561-632) couple the fields through
advection, the Cahn-Hilliard chemical potential, the capillary force, the
active stress, the polarization's molecular field, the lambda_0 self-advection
and the anchoring term.  Each case below builds a state on which every term
of dt v, dt P and dt phi reduces to a pointwise closed form, using only the
Fourier symbols of the reference's stencils (pinned on their own in
test_oracle_kat.py::test_stencil_symbols):

    d1x cos(th) = -d_x sin(th),  d_x = sin(k_x h)/h * S(k_y) S(k_z),
    S(q) = (2 + cos(q h))/3,  lap cos(th) = lam(k) cos(th),
    lam(k) = ((4/3)(c_x + c_y + c_z) + (2/3) c_x c_y c_z - 14/3)/h^2.

Products of modes are expanded into sum and difference wavevectors, and powers
of cos into harmonics, so no expected value uses a stencil or an oracle
output.  The builders take the grid size, so tests/test_gpu_coupling.py
checks the same cases on the HIP step; the check functions take the oracle
module as an argument, so tests/test_oracle_mutations.py can run them against
copies of the oracle with a planted mistake and show that each such mistake
fails one of them.
"""
import numpy as np
import pytest

N = 16
H = 1.1
DT = 0.01


def grid(n):
    return np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")  # z, y, x


def kvec(m, n, h=H):
    return np.array([2 * np.pi * mi / (n * h) for mi in m], dtype=np.float64)


def theta(k, phase, n, h=H):
    z, y, x = grid(n)
    return k[0] * x * h + k[1] * y * h + k[2] * z * h + phase


def S(q, h=H):
    return (2 + np.cos(q * h)) / 3


def dsym(k, h=H):
    """Symbol of (d1x, d1y, d1z): d1_a e^{i k.x} = i d_a e^{i k.x}."""
    return np.array([np.sin(k[0] * h) / h * S(k[1], h) * S(k[2], h),
                     np.sin(k[1] * h) / h * S(k[0], h) * S(k[2], h),
                     np.sin(k[2] * h) / h * S(k[0], h) * S(k[1], h)])


def lsym(k, h=H):
    """Symbol of the 15-point Laplacian."""
    cx, cy, cz = np.cos(k * h)
    return ((4 / 3) * (cx + cy + cz) + (2 / 3) * cx * cy * cz - 14 / 3) / h ** 2


def vec(a, f):
    """(3,) amplitude times a scalar field -> (3, n, n, n)."""
    return np.asarray(a, dtype=np.float64)[:, None, None, None] * f[None]


def const_vec(a, n):
    return np.asarray(a, dtype=np.float64)[:, None, None, None] * np.ones((3, n, n, n))


def chemical_potential(phi0, eps, k, th, aPhi, kPhi):
    """mu0 = aPhi*2*(phi - 3 phi^2 + 2 phi^3) - kPhi*lap(phi) for
    phi = phi0 + eps cos(th), and lap(mu0), both from harmonics n <= 3."""
    f0 = phi0 - 3 * phi0 ** 2 + 2 * phi0 ** 3
    f1 = 1 - 6 * phi0 + 6 * phi0 ** 2
    f2 = -3 + 6 * phi0
    f3 = 2.0
    # f(phi0 + e c) = f0 + f1 e c + f2 e^2 c^2 + f3 e^3 c^3, with
    # c^2 = (1 + cos 2th)/2 and c^3 = (3 cos th + cos 3th)/4
    A0 = f0 + f2 * eps ** 2 / 2
    A1 = f1 * eps + 0.75 * f3 * eps ** 3
    A2 = f2 * eps ** 2 / 2
    A3 = f3 * eps ** 3 / 4
    l1, l2, l3 = lsym(k), lsym(2 * k), lsym(3 * k)
    c1, c2, c3 = np.cos(th), np.cos(2 * th), np.cos(3 * th)
    mu0 = aPhi * 2 * (A0 + A1 * c1 + A2 * c2 + A3 * c3) - kPhi * eps * l1 * c1
    lap_mu0 = aPhi * 2 * (A1 * l1 * c1 + A2 * l2 * c2 + A3 * l3 * c3) - kPhi * eps * l1 * l1 * c1
    return mu0, lap_mu0


def case(phi, P, v, prm, dtphi, dtv, dtP, **extra):
    """A state, its parameters, the exact dt phi, dt v, dt P, and optional
    exact intermediates (mu0, mu, P2) that only the oracle exposes."""
    return dict(phi=phi, P=P, v=v, params=prm, dtphi=dtphi, dtv=dtv, dtP=dtP, extra=extra)


# ---------------------------------------------------------------------------
# The cases, on an n^3 grid
# ---------------------------------------------------------------------------

def build_advection(n):
    """Transport by the flow, TTSH3D.cu:610-616,631: with P = 0 and
    aPhi = kPhi = 0 (so mu0 = 0), dt phi = -v.D phi and
    dt v_b = -v.D v_b + mu lap v_b for v = v0 + a cos(th2) (a not transverse:
    the self-advection (a.d) a_b sin cos term is included), phi = phi0 + e cos(th1)."""
    k1, k2 = kvec((3, 1, 2), n), kvec((1, 2, 5), n)
    th1, th2 = theta(k1, 0.4, n), theta(k2, -0.9, n)
    d1, d2 = dsym(k1), dsym(k2)
    l2 = lsym(k2)
    phi0, eps = 0.6, 0.2
    v0 = np.array([0.3, -0.2, 0.45])
    a = np.array([0.15, 0.25, -0.1])
    mu = 0.7
    phi = phi0 + eps * np.cos(th1)
    v = const_vec(v0, n) + vec(a, np.cos(th2))
    P = np.zeros((3, n, n, n))
    prm = dict(aPhi=0.0, kPhi=0.0, mu=mu, cA=-1.3, alpha=-0.02, beta=4.366, eta=-0.187, lambda0=9.0, A=0.5)
    dtphi = eps * np.sin(th1) * ((v0 @ d1) + (a @ d1) * np.cos(th2))
    adv = np.sin(th2) * ((v0 @ d2) + (a @ d2) * np.cos(th2))
    dtv = vec(a, adv + mu * l2 * np.cos(th2))
    return case(phi, P, v, prm, dtphi, dtv, np.zeros((3, n, n, n)), mu0=np.zeros((n, n, n)))


def build_cahn_hilliard_capillary(n):
    """Chemical potential and capillary force, TTSH3D.cu:565,610-616,631:
    v = P = 0, phi = phi0 + e cos(th): mu0 and gamma*lap(mu0) through the
    harmonics cos(n th), n <= 3, and dt v_b = mu0 * D_b phi."""
    k = kvec((2, 3, 1), n)
    th = theta(k, 0.25, n)
    d = dsym(k)
    phi0, eps = 0.45, 0.3
    aPhi, kPhi, gamma = 0.8, 1.3, 0.9
    phi = phi0 + eps * np.cos(th)
    zero = np.zeros((3, n, n, n))
    prm = dict(aPhi=aPhi, kPhi=kPhi, gamma=gamma, mu=1.1, cA=-2.0, A=0.7)
    mu0, lap_mu0 = chemical_potential(phi0, eps, k, th, aPhi, kPhi)
    return case(phi, zero, zero, prm, gamma * lap_mu0, vec(-eps * d, mu0 * np.sin(th)), zero, mu0=mu0)


def build_active_stress(n):
    """Active stress and molecular field, TTSH3D.cu:573-575,587-595,610-630:
    v = 0, aPhi = kPhi = 0, lambda0 = A = 0, phi = phi0 + e cos(th2) and a
    transverse polarization mode P = a cos(th1) (a.d1 = 0).
    dt v_b = cA sum_a D_a(phi D_a P_b) expands into the modes k1, k1 +- k2;
    dt P_b = -alpha phi P_b - beta |P|^2 P_b + eta lap(phi P_b + La^2 lap P_b)."""
    k1, k2 = kvec((2, 1, 3), n), kvec((1, 3, 1), n)
    th1, th2 = theta(k1, 0.3, n), theta(k2, 1.1, n)
    d1, dp, dm = dsym(k1), dsym(k1 + k2), dsym(k1 - k2)
    l1, lp, lm = lsym(k1), lsym(k1 + k2), lsym(k1 - k2)
    # a transverse to d1
    t = np.cross(d1, np.array([0.3, -0.7, 0.4]))
    a = 0.35 * t / np.linalg.norm(t)
    assert abs(a @ d1) < 1e-15
    phi0, eps = 0.7, 0.2
    cA, alpha, beta, eta, La = -1.7, -0.05, 2.5, -0.3, 1.2
    phi = phi0 + eps * np.cos(th2)
    P = vec(a, np.cos(th1))
    v = np.zeros((3, n, n, n))
    prm = dict(aPhi=0.0, kPhi=0.0, cA=cA, alpha=alpha, beta=beta, eta=eta, La=La, lambda0=0.0, A=0.0, mu=0.9)
    tp, tm = th1 + th2, th1 - th2
    # sum_a d1_a [phi0 d1_a cos th1 + e/2 (dp_a cos tp + dm_a cos tm)]
    s = phi0 * (d1 @ d1) * np.cos(th1) + eps / 2 * ((d1 @ dp) * np.cos(tp) + (d1 @ dm) * np.cos(tm))
    c1 = np.cos(th1)
    molecular = phi0 * l1 * c1 + eps / 2 * (lp * np.cos(tp) + lm * np.cos(tm)) + La ** 2 * l1 * l1 * c1
    dtP = vec(a, -alpha * phi * c1 - beta * (a @ a) * c1 ** 3 + eta * molecular)
    return case(phi, P, v, prm, np.zeros((n, n, n)), vec(-cA * a, s), dtP,
                mu=vec(a, phi * c1 + La ** 2 * l1 * c1))


def build_self_advection(n):
    """lambda0 term and active stress with uniform phi, TTSH3D.cu:618-630:
    v = 0, phi = phi0, P = P0 + a cos(th) (a not transverse).
    dt P_b = -alpha phi0 P_b - beta |P|^2 P_b + eta(phi0 lam + La^2 lam^2) a_b cos
             - lambda0 (P.D) P_b,  with -(P.D)P_b = a_b sin (P0.d + (a.d) cos);
    dt v_b = -cA phi0 |d|^2 a_b cos;  dt phi = 0; anchoring vanishes (D phi = 0)."""
    k = kvec((1, 4, 2), n)
    th = theta(k, -0.6, n)
    d = dsym(k)
    lam = lsym(k)
    phi0 = 0.85
    P0 = np.array([0.4, -0.3, 0.5])
    a = np.array([0.2, 0.1, -0.25])
    cA, alpha, beta, eta, La, lambda0 = 1.4, -0.03, 1.9, -0.25, 0.9, 6.5
    c, sn = np.cos(th), np.sin(th)
    phi = np.full((n, n, n), phi0)
    P = const_vec(P0, n) + vec(a, c)
    v = np.zeros((3, n, n, n))
    prm = dict(cA=cA, alpha=alpha, beta=beta, eta=eta, La=La, lambda0=lambda0, A=0.8, aPhi=1.0, kPhi=1.0)
    P2 = sum((P0[b] + a[b] * c) ** 2 for b in range(3))
    dtP = (-alpha * phi0 * P - beta * P2[None] * P
           + vec(a, eta * (phi0 * lam + La ** 2 * lam * lam) * c)
           + lambda0 * vec(a, sn * ((P0 @ d) + (a @ d) * c)))
    return case(phi, P, v, prm, np.zeros((n, n, n)), vec(-cA * phi0 * (d @ d) * a, c), dtP, P2=P2)


def build_anchoring(n):
    """Anchoring, TTSH3D.cu:597-602,624-630: v = 0, P = P0 uniform,
    phi = phi0 + e cos(th).  D phi = -e d sin, |D phi|^2 = e^2 sin^2 |d|^2, and
    A_bc = D_b phi D_c phi (the reference's 1/3 is an integer zero), so
      dt P_b = -alpha phi P0_b - beta |P0|^2 P0_b + eta e lam cos P0_b
               - A e^2 sin^2 (|d|^2 |P0|^2 P0_b + |P0| d_b (d.P0));
    dt v_b = mu0 D_b phi and dt phi = gamma lap mu0 (as in the Cahn-Hilliard case)."""
    k = kvec((3, 2, 1), n)
    th = theta(k, 0.8, n)
    d = dsym(k)
    lam = lsym(k)
    phi0, eps = 0.55, 0.35
    P0 = np.array([0.5, 0.2, -0.6])
    alpha, beta, eta, La, A, lambda0 = -0.04, 3.1, -0.2, 1.3, 2.3, 7.0
    aPhi, kPhi, gamma = 1.2, 0.9, 1.1
    c, sn = np.cos(th), np.sin(th)
    phi = phi0 + eps * c
    P = const_vec(P0, n)
    v = np.zeros((3, n, n, n))
    prm = dict(alpha=alpha, beta=beta, eta=eta, La=La, A=A, lambda0=lambda0, aPhi=aPhi, kPhi=kPhi, gamma=gamma,
               cA=-1.0, mu=1.0)
    nP2 = P0 @ P0
    anch = A * eps ** 2 * sn ** 2
    dtP = (-alpha * phi[None] * P - beta * nP2 * P + vec(P0, eta * eps * lam * c)
           - vec((d @ d) * nP2 * P0 + np.sqrt(nP2) * (d @ P0) * d, anch))
    mu0, lap_mu0 = chemical_potential(phi0, eps, k, th, aPhi, kPhi)
    return case(phi, P, v, prm, gamma * lap_mu0, vec(-eps * d, mu0 * sn), dtP)


def build_vorticity_coupling_is_zero(n):
    """The co-rotation and flow-alignment terms OP, DP are integer zeros
    (1/2 * (...), TTSH3D.cu:579-585): a velocity mode with vorticity leaves
    dt P at its value without flow, apart from advection, which is zero for
    uniform P.  v = a cos(th) (transverse), P = P0, phi = phi0."""
    k = kvec((2, 2, 3), n)
    th = theta(k, 0.1, n)
    d = dsym(k)
    t = np.cross(d, np.array([1.0, 0.2, -0.5]))
    a = 0.4 * t / np.linalg.norm(t)
    phi0 = 0.9
    P0 = np.array([0.3, 0.6, -0.2])
    alpha, beta, mu = -0.02, 4.366, 1.0
    phi = np.full((n, n, n), phi0)
    P = const_vec(P0, n)
    v = vec(a, np.cos(th))
    prm = dict(alpha=alpha, beta=beta, kappa=-0.364, lambda0=9.0, A=0.0, mu=mu)
    dtP = const_vec(-alpha * phi0 * P0 - beta * (P0 @ P0) * P0, n)
    # v is transverse: -v.D v = (a.d) a sin cos = 0; mu lap v; no phi gradient
    dtv = vec(a, mu * lsym(k) * np.cos(th))
    return case(phi, P, v, prm, np.zeros((n, n, n)), dtv, dtP)


BUILDERS = [build_advection, build_cahn_hilliard_capillary, build_active_stress, build_self_advection,
            build_anchoring, build_vorticity_coupling_is_zero]


# ---------------------------------------------------------------------------
# Checks against the oracle (each takes the oracle module)
# ---------------------------------------------------------------------------

def one_step(O, c, n=N):
    o = O.Oracle(n, faithful=False, h=H, dt0=DT, **c["params"])
    o.set_state(c["phi"], c["P"], c["v"])
    o.step(1)
    return o


def close(got, exp, what, rtol=1e-12):
    scale = max(1.0, float(np.abs(exp).max()))
    err = float(np.abs(np.asarray(got) - exp).max())
    assert err <= rtol * scale, f"{what}: max error {err:.3e} (scale {scale:.3e})"


def check_case(O, build):
    c = build(N)
    o = one_step(O, c)
    name = build.__name__[len("build_"):]
    close(o.aux("dtphi"), c["dtphi"], f"{name}: dt phi")
    close(o.aux("dtv"), c["dtv"], f"{name}: dt v")
    close(o.aux("dtP"), c["dtP"], f"{name}: dt P")
    for key, exp in c["extra"].items():
        close(o.aux(key), exp, f"{name}: {key}")
    # the predictor and phi^{n+1} (getVStar, getVNew for phi, TTSH3D.cu:471-491,513)
    close(o.aux("vstar"), c["v"] + DT * c["dtv"], f"{name}: v* = v + Dt dt v")
    close(o.aux("pstar"), c["P"] + DT * c["dtP"], f"{name}: P* = P + Dt dt P")
    close(o.get_state()[0], c["phi"] + DT * c["dtphi"], f"{name}: phi^(n+1) = phi + Dt dt phi")


def _check(build):
    def f(O):
        check_case(O, build)
    f.__name__ = "check_" + build.__name__[len("build_"):]
    return f


CHECKS = [_check(b) for b in BUILDERS]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__ for c in CHECKS])
def test_coupling_closed_form(oracle_mod, check):
    check(oracle_mod)
