"""Known-answer tests that pin the FP64 oracle (CPU only).

The reference has no tests, fixtures or golden vectors, so each part of the
oracle is pinned by what the mathematics fixes:
  - FFT: brute-force DFT, numpy fixtures (tests/golden/fft_golden.npz),
    Parseval, linearity, delta <-> constant, round trip; radix 2 and 3.
  - stencils d1x/d1y/d1z/lap (TTSH3D.cu:3280-3343): exact Fourier symbols.
  - projection (TTSH3D.cu:635-691, 505-511): single-mode formula
    v = v* - d (d.v*)/|k|^2, d_a = sin(k_a h)/h S(k_b) S(k_c).
  - the step: uniform-state closed form, shift and axis-permutation
    equivariance, kappa independence (integer-division zeros, TTSH3D.cu:579-602),
    Sigma = phi*G exactly, v* affine in cA, phi mass conserved from v = 0,
    stale-halo divergence confined to the boundary shell (TTSH3D.cu:665-667).
"""
import os

import numpy as np
import pytest

from conftest import rel_rms

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fft_golden.npz")


def brute_dft(a, sign):
    N = a.shape[0]
    k = np.arange(N)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / N)
    return W @ a


# ---------------------------------------------------------------- FFT ------
@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 8, 9, 12, 16, 18, 24])
def test_fft1d_brute_force(oracle_mod, N):
    rng = np.random.default_rng(N)
    a = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    assert np.allclose(oracle_mod.fft1d(a), brute_dft(a, -1), rtol=0, atol=1e-12 * N)
    assert np.allclose(oracle_mod.fft1d(a, inverse=True), brute_dft(a, +1), rtol=0, atol=1e-12 * N)


@pytest.mark.parametrize("N", [2, 3, 4, 6])
def test_fft3d_brute_force(oracle_mod, N):
    rng = np.random.default_rng(100 + N)
    a = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    ref = a.astype(np.complex128)
    for ax in range(3):
        ref = np.apply_along_axis(lambda v: brute_dft(v, -1), ax, ref)
    assert np.abs(oracle_mod.fft3d(a) - ref).max() < 1e-11


def test_fft_golden(oracle_mod):
    g = np.load(GOLD)
    for N in (8, 12, 16, 48, 96):
        x = g[f"x1_{N}"]
        assert np.abs(oracle_mod.fft1d(x) - g[f"f1_{N}"]).max() < 1e-12 * N
        assert np.abs(oracle_mod.fft1d(x, inverse=True) - g[f"i1_{N}"]).max() < 1e-12 * N
    for N in (6, 8, 12):
        x = g[f"x3_{N}"]
        assert np.abs(oracle_mod.fft3d(x) - g[f"f3_{N}"]).max() < 1e-10
        assert np.abs(oracle_mod.fft3d(x, inverse=True) - g[f"i3_{N}"]).max() < 1e-10


@pytest.mark.parametrize("N", [12, 16])
def test_fft_properties(oracle_mod, N):
    rng = np.random.default_rng(7)
    a = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    b = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    A = oracle_mod.fft3d(a)
    # Parseval
    assert np.sum(np.abs(A) ** 2) == pytest.approx(N ** 3 * np.sum(np.abs(a) ** 2), rel=1e-12)
    # linearity
    assert np.abs(oracle_mod.fft3d(2.5 * a - 1j * b) - (2.5 * A - 1j * oracle_mod.fft3d(b))).max() < 1e-10
    # round trip (unnormalised)
    assert np.abs(oracle_mod.fft3d(A, inverse=True) / N ** 3 - a).max() < 1e-12
    # delta <-> constant
    d = np.zeros((N, N, N), complex)
    d[0, 0, 0] = 1.0
    assert np.abs(oracle_mod.fft3d(d) - 1.0).max() < 1e-14
    c = np.ones((N, N, N), complex)
    C = oracle_mod.fft3d(c)
    assert C[0, 0, 0] == pytest.approx(N ** 3)
    C[0, 0, 0] = 0
    assert np.abs(C).max() < 1e-9


# ----------------------------------------------------------- stencils ------
def grid(N):
    return np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")  # z, y, x


@pytest.mark.parametrize("m", [(3, 5, 2), (1, 0, 7), (8, 8, 8), (0, 4, 0)])
def test_stencil_symbols(oracle_mod, m):
    N, h = 16, 1.3
    z, y, x = grid(N)
    kx, ky, kz = [2 * np.pi * mi / (N * h) for mi in m]
    th = kx * x * h + ky * y * h + kz * z * h + 0.3
    f = np.cos(th)
    S = lambda k: (2 + np.cos(k * h)) / 3
    exp = {
        "d1x": -np.sin(th) * np.sin(kx * h) / h * S(ky) * S(kz),
        "d1y": -np.sin(th) * np.sin(ky * h) / h * S(kx) * S(kz),
        "d1z": -np.sin(th) * np.sin(kz * h) / h * S(kx) * S(ky),
    }
    cx, cy, cz = np.cos(kx * h), np.cos(ky * h), np.cos(kz * h)
    exp["lap"] = ((4 / 3) * (cx + cy + cz) + (2 / 3) * cx * cy * cz - 14 / 3) / h ** 2 * f
    for k, e in exp.items():
        assert np.abs(oracle_mod.stencil(k, f, h) - e).max() < 1e-13, k


def test_stencil_axis_roles(oracle_mod):
    """d1x differentiates along x (the fastest index), d1z along z."""
    N = 12
    z, y, x = grid(N)
    f = np.sin(2 * np.pi * x / N)
    assert np.abs(oracle_mod.stencil("d1y", f)).max() < 1e-15
    assert np.abs(oracle_mod.stencil("d1z", f)).max() < 1e-15
    assert np.abs(oracle_mod.stencil("d1x", f)).max() > 0.1


# --------------------------------------------------------- projection ------
@pytest.mark.parametrize("m", [(1, 2, 3), (5, 0, 2), (7, 7, 1)])
def test_projection_single_mode(oracle_mod, m):
    N, h, dt0 = 16, 1.1, 0.01
    z, y, x = grid(N)
    k = np.array([2 * np.pi * mi / (N * h) for mi in m])
    th = k[0] * x * h + k[1] * y * h + k[2] * z * h
    rng = np.random.default_rng(sum(m))
    a = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    Xs = np.stack([np.real(a[c] * np.exp(1j * th)) for c in range(3)])
    S = lambda q: (2 + np.cos(q * h)) / 3
    d = np.array([np.sin(k[0] * h) / h * S(k[1]) * S(k[2]),
                  np.sin(k[1] * h) / h * S(k[0]) * S(k[2]),
                  np.sin(k[2] * h) / h * S(k[0]) * S(k[1])])
    k2 = float(np.sum(k * k))
    ahat = a - d * (d @ a) / k2
    exp = np.stack([np.real(ahat[c] * np.exp(1j * th)) for c in range(3)])
    p, X = oracle_mod.projection(Xs, h=h, dt0=dt0)
    assert np.abs(X - exp).max() < 1e-12
    # p = -(i d.a)/(dt0 k2) e^{i th}
    pexp = np.real(-(1j * (d @ a)) / (dt0 * k2) * np.exp(1j * th))
    assert np.abs(p - pexp).max() < 1e-10 * max(1.0, np.abs(pexp).max())


def test_projection_constant_field_zero_pressure(oracle_mod):
    N = 8
    Xs = np.ones((3, N, N, N)) * np.array([0.3, -1.0, 2.0])[:, None, None, None]
    p, X = oracle_mod.projection(Xs)
    assert np.abs(p).max() < 1e-12
    assert np.abs(X - Xs).max() < 1e-12


# --------------------------------------------------------------- step ------
def random_state(N, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    z, y, x = grid(N)
    r = np.sqrt((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2)
    phi = 0.5 * (1 - np.tanh((r - N / 4) / 2.0))
    P = amp * rng.standard_normal((3, N, N, N)) * (phi > 0.5)
    v = 0.05 * rng.standard_normal((3, N, N, N))
    return phi, P, v


def run(O, N, phi, P, v, steps=1, faithful=False, vs=None, ps=None, **kw):
    o = O.Oracle(N, faithful=faithful, **kw)
    o.set_state(phi, P, v, vs, ps)
    o.step(steps)
    return o


def test_uniform_state_closed_form(oracle_mod):
    N = 8
    phi0, P0, v0 = 0.7, np.array([0.3, -0.5, 0.2]), np.array([0.1, 0.02, -0.04])
    phi = np.full((N, N, N), phi0)
    P = np.ones((3, N, N, N)) * P0[:, None, None, None]
    v = np.ones((3, N, N, N)) * v0[:, None, None, None]
    prm = dict(alpha=-0.02, beta=4.366, dt0=0.01)
    o = run(oracle_mod, N, phi, P, v, **prm)
    ophi, oP, ov = o.get_state()
    dtP = -prm["alpha"] * phi0 * P0 - prm["beta"] * (P0 @ P0) * P0
    assert np.abs(oP - (P0 + prm["dt0"] * dtP)[:, None, None, None]).max() < 1e-14
    assert np.abs(ov - v).max() < 1e-14
    assert np.abs(ophi - phi0).max() < 1e-14
    assert np.abs(o.aux("p_v")).max() < 1e-10 and np.abs(o.aux("p_p")).max() < 1e-10


def test_kappa_has_no_effect(oracle_mod):
    """DP = OP = 0 through integer division (TTSH3D.cu:579-585), so kappa is inert."""
    N = 12
    phi, P, v = random_state(N, 1)
    a = run(oracle_mod, N, phi, P, v, steps=2, kappa=-0.364).get_state()
    b = run(oracle_mod, N, phi, P, v, steps=2, kappa=7.5).get_state()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_sigma_is_phi_times_gradient(oracle_mod):
    """The 1/28 Laplacian correction of Sigma is integer zero (TTSH3D.cu:587-595)."""
    N = 12
    phi, P, v = random_state(N, 2)
    o = run(oracle_mod, N, phi, P, v)
    G = o.aux("G")      # dxPx dyPx dzPx dxPy ...: G[3b+a] = D_a P_b
    Sig = o.aux("Sigma")  # xx xy xz yx ...: Sig[3a+b]
    for a in range(3):
        for b in range(3):
            assert np.array_equal(Sig[3 * a + b], phi * G[3 * b + a])


def test_shift_equivariance(oracle_mod):
    N = 12
    phi, P, v = random_state(N, 3)
    sh = (3, -2, 5)
    roll = lambda f: np.roll(f, sh, axis=(-3, -2, -1))
    a = run(oracle_mod, N, phi, P, v, steps=2).get_state()
    b = run(oracle_mod, N, roll(phi), roll(P), roll(v), steps=2).get_state()
    for x, y in zip(a, b):
        assert np.abs(roll(x) - y).max() < 1e-12


def test_axis_permutation_symmetry(oracle_mod):
    """Swapping x and y (and the x/y components) commutes with the step."""
    N = 12
    phi, P, v = random_state(N, 4)
    sw = lambda f: np.swapaxes(f, -1, -2)
    swv = lambda F: np.stack([sw(F[1]), sw(F[0]), sw(F[2])])
    a = run(oracle_mod, N, phi, P, v, steps=2).get_state()
    b = run(oracle_mod, N, sw(phi), swv(P), swv(v), steps=2).get_state()
    assert np.abs(sw(a[0]) - b[0]).max() < 1e-12
    assert np.abs(swv(a[1]) - b[1]).max() < 1e-12
    assert np.abs(swv(a[2]) - b[2]).max() < 1e-12


def test_vstar_affine_in_activity(oracle_mod):
    N = 12
    phi, P, v = random_state(N, 5)
    vs = {cA: run(oracle_mod, N, phi, P, v, cA=cA).aux("vstar") for cA in (0.0, -1.0, 2.5)}
    d1 = vs[-1.0] - vs[0.0]
    d2 = vs[2.5] - vs[0.0]
    assert np.abs(d2 - (-2.5) * d1).max() < 1e-12 * max(1.0, np.abs(d2).max())
    assert np.abs(d1).max() > 1e-6


def test_mass_conserved_from_rest(oracle_mod):
    N = 16
    phi, P, _ = random_state(N, 6)
    v = np.zeros((3, N, N, N))
    o = run(oracle_mod, N, phi, P, v)
    assert o.get_state()[0].sum() == pytest.approx(phi.sum(), rel=1e-13)


def test_stale_halo_confined_to_shell(oracle_mod):
    """With everything at rest, the faithful divergence of step 1 sees only the
    previous predictors placed in the halo: it is non-zero only on the
    boundary shell, and linear in them (TTSH3D.cu:366-371, 665-667)."""
    N = 10
    z, y, x = grid(N)
    phi = np.zeros((N, N, N))
    P = np.zeros((3, N, N, N))
    v = np.zeros((3, N, N, N))
    rng = np.random.default_rng(9)
    g = rng.standard_normal((3, N, N, N))
    o1 = run(oracle_mod, N, phi, P, v, faithful=True, vs=g, ps=2 * g)
    s = o1.aux("sdiv_v")
    shell = (x == 0) | (x == N - 1) | (y == 0) | (y == N - 1) | (z == 0) | (z == N - 1)
    assert np.abs(s[~shell]).max() == 0.0
    assert np.abs(s[shell]).max() > 1e-3
    assert np.abs(o1.aux("sdiv_p") - 2 * s).max() < 1e-12 * np.abs(s).max()
    clean = run(oracle_mod, N, phi, P, v, faithful=False, vs=g, ps=2 * g)
    assert np.abs(clean.aux("sdiv_v")).max() == 0.0


def test_reductions(oracle_mod):
    N = 8
    phi, P, v = random_state(N, 11)
    o = oracle_mod.Oracle(N, h=1.5)
    o.set_state(phi, P, v)
    r = o.reductions()
    assert r[0] == pytest.approx(phi.sum() * 1.5 ** 3, rel=1e-12)
    assert r[1] == pytest.approx(np.sqrt((P ** 2).sum(0)).mean(), rel=1e-12)
    assert r[2] == pytest.approx(0.5 * (v ** 2).sum() * 1.5 ** 3, rel=1e-12)


def test_oracle_against_droplet_mass_closed_form(oracle_mod):
    """Lattice mass of the tanh droplet vs (4 pi/3)(R^3 + pi^2 xi^2 R/4)."""
    N, R, xi = 64, 20.0, 2.0
    z, y, x = grid(N)
    r = np.sqrt((x - N / 2) ** 2 + (y - N / 2) ** 2 + (z - N / 2) ** 2)
    phi = 0.5 * (1 - np.tanh((r - R) / xi))
    est = 4 * np.pi / 3 * (R ** 3 + np.pi ** 2 * xi ** 2 * R / 4)
    assert phi.sum() == pytest.approx(est, rel=1e-5)
    _ = rel_rms
