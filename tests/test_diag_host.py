"""Host-side diagnostics: the contour radius of ttsh_contour_radius against
the reference's algorithm (characteristicVelocity.py:75-122 on matplotlib's
contour paths: committed fixtures from tests/golden/make_contour_golden.py,
and matplotlib itself where it imports) and against shapes with a known
contour.  No GPU needed."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from oracle import postproc  # noqa: E402


def droplet_slice(N, R0, xi, cx, cy, ax=1.0, ay=1.0):
    y, x = np.mgrid[0:N, 0:N].astype(np.float64)
    r = np.sqrt(((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2)
    return 0.5 * (1.0 - np.tanh((r - R0) / xi))


GOLD = os.path.join(ROOT, "tests", "golden", "contour_golden.npz")


def test_contour_matches_golden(ttsh):
    """R(theta) and the centre against the reference algorithm's output on
    one droplet, an ellipse, two droplets (centre of the last path), a
    droplet cut by the border (open lines), an annulus and a random field."""
    g = np.load(GOLD)
    for name in g["names"]:
        f = g[f"field_{name}"]
        R, c = ttsh.contour_radius(f)
        assert np.abs(R - g[f"R_{name}"]).max() < 1e-9, name
        assert np.abs(np.array(c) - g[f"center_{name}"]).max() < 1e-9, name


def test_contour_matches_matplotlib_live(ttsh):
    """The same against matplotlib in this process, on seeded random fields
    of many sizes (open lines, closed loops, saddles, points on the level)."""
    pytest.importorskip("matplotlib")
    rng = np.random.default_rng(11)
    checked = 0
    for t in range(120):
        N = int(rng.integers(3, 40))
        if t % 3 == 0:
            f = rng.integers(0, 3, (N, N)) * 0.25 + 0.25  # values on the level
        else:
            f = rng.random((N, N))
            for _ in range(t % 4):
                f = (f + np.roll(f, 1, 0) + np.roll(f, 1, 1) + np.roll(f, (1, 1), (0, 1))) / 4
        f32 = f.astype(np.float32)
        try:
            R_ref, c_ref = postproc.contour_radius(f32.astype(np.float64))
        except ValueError:
            with pytest.raises(ttsh.TTSHError):
                ttsh.contour_radius(f32)
            continue
        R, c = ttsh.contour_radius(f32)
        assert np.abs(np.array(c) - np.array(c_ref)).max() < 1e-9, (t, N)
        # Where two vertices lie at the same polar angle (fields with values
        # on the level make such ties), which one np.argmin takes depends on
        # the last bit of arctan2; accept any vertex within 1e-12 of the
        # nearest angle there.
        pts = np.concatenate(postproc.contour_paths(f32.astype(np.float64)))
        dx, dy = pts[:, 0] - c_ref[0], pts[:, 1] - c_ref[1]
        pr, pt = np.sqrt(dx ** 2 + dy ** 2), np.arctan2(dy, dx)
        for a, deg in enumerate(range(-180, 180, 3)):
            d = np.abs(pt - deg / 180 * np.pi)
            ok = pr[d <= d.min() + 1e-12]
            assert np.abs(ok - R[a]).min() < 1e-9, (t, N, deg)
        assert abs(R[np.argmax(np.abs(R - R_ref))] - R_ref[np.argmax(np.abs(R - R_ref))]) < 1e-9 or t % 3 == 0
        checked += 1
    assert checked > 60


def test_contour_of_a_circle():
    """A tanh droplet of radius R0 centred on a grid point or a cell centre:
    R(theta) ~ R0 at all 120 angles.  The centre is the mean of the closed
    path's vertices, whose first vertex is repeated at its end (as in
    matplotlib's paths), so it sits (v0 - c)/(n + 1) off the droplet centre
    c, where v0 is the first vertex and n the number of crossings."""
    import ttsh3d as ttsh
    N, R0 = 96, 30.0
    for cx, cy in ((48.0, 48.0), (47.5, 47.5)):
        f = droplet_slice(N, R0, 2.0, cx, cy).astype(np.float32)
        R, c = ttsh.contour_radius(f)
        assert R.shape == (120,)
        assert abs(c[0] - cx) < 0.2 and abs(c[1] - cy) < 0.2
        assert np.max(np.abs(R - R0)) < 0.2
    f = droplet_slice(N, R0, 2.0, 47.25, 48.6).astype(np.float32)
    R, c = ttsh.contour_radius(f)
    assert abs(c[0] - 47.25) < 0.3 and abs(c[1] - 48.6) < 0.3
    assert np.max(np.abs(R - R0)) < 0.3


def test_contour_of_an_ellipse_angles():
    """Semi-axes a (x) and b (y): R(0) = R(+-180) ~ a and R(+-90) ~ b; the
    angle grid starts at -180 degrees in steps of 3."""
    import ttsh3d as ttsh
    N, a, b = 128, 40.0, 25.0
    f = droplet_slice(N, 1.0, 2.0 / 30.0, 64.0, 64.0, ax=a, ay=b).astype(np.float32)
    R, _ = ttsh.contour_radius(f)
    deg = np.arange(-180, 180, 3)
    assert abs(R[deg == 0][0] - a) < 0.1 and abs(R[0] - a) < 0.1
    assert abs(R[deg == 90][0] - b) < 0.1 and abs(R[deg == -90][0] - b) < 0.1
    # in between, R(theta) of an ellipse about its centre
    th = np.deg2rad(deg)
    Re = a * b / np.sqrt((b * np.cos(th)) ** 2 + (a * np.sin(th)) ** 2)
    assert np.max(np.abs(R - Re)) < 0.6  # nearest crossing, not an interpolated angle


def test_contour_no_crossing_is_an_error():
    import ttsh3d as ttsh
    with pytest.raises(ttsh.TTSHError):
        ttsh.contour_radius(np.zeros((16, 16), np.float32))


def test_centroid_restatement_closed_form():
    """postproc.centroid on a shifted box function: the mass centre is the box
    centre, returned in the (y, x, z) order of the reference's arrays."""
    N = 32
    phi = np.zeros((N, N, N))
    phi[4:10, 11:21, 2:8] = 1.0  # z in 4..9, y in 11..20, x in 2..7
    idx, c, m = postproc.centroid(phi)
    assert m == 6 * 10 * 6
    assert c == pytest.approx((15.5, 4.5, 6.5))
    assert idx == (15, 4, 6)


def test_vorticity_restatement_single_mode():
    """d1xO4 of sin(kx) is (8 sin(kh) - sin(2kh)) / (6h) cos(kx)."""
    N, h = 32, 1.0
    k = 2 * math.pi * 3 / N
    x = np.arange(N)
    u = np.tile(np.sin(k * x), (N, 1))  # [y][x]
    d = postproc.d1xO4(u, h)
    sym = (8 * math.sin(k * h) - math.sin(2 * k * h)) / (6 * h)
    assert np.allclose(d, sym * np.tile(np.cos(k * x), (N, 1)), atol=1e-12)
    assert np.allclose(postproc.d1yO4(u.T, h), d.T, atol=1e-12)
