"""Host-side diagnostics: the contour radius of ttsh_contour_radius against
the restatement in oracle/postproc.py and against shapes with a known
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


def test_contour_matches_restatement(ttsh):
    rng = np.random.default_rng(7)
    for N, R0, cx, cy in ((64, 20.0, 31.3, 32.7), (48, 11.0, 20.2, 25.9), (128, 40.0, 64.5, 60.1)):
        f = droplet_slice(N, R0, 2.0, cx, cy) + 0.02 * rng.standard_normal((N, N))
        f32 = f.astype(np.float32)
        R, c = ttsh.contour_radius(f32)
        R_ref, c_ref = postproc.contour_radius(f32.astype(np.float64))
        assert np.allclose(R, R_ref, rtol=0, atol=1e-11)
        assert abs(c[0] - c_ref[0]) < 1e-11 and abs(c[1] - c_ref[1]) < 1e-11


def test_contour_of_a_circle():
    """A tanh droplet of radius R0 centred on a grid point or a cell centre:
    the crossings are symmetric, so their mean is the droplet's centre and
    R(theta) ~ R0 at all 120 angles within the interpolation error.  Off such
    a centre the mean of the edge crossings (like the mean of contour
    vertices the reference takes) is biased by a fraction of a cell."""
    import ttsh3d as ttsh
    N, R0 = 96, 30.0
    for cx, cy in ((48.0, 48.0), (47.5, 47.5)):
        f = droplet_slice(N, R0, 2.0, cx, cy).astype(np.float32)
        R, c = ttsh.contour_radius(f)
        assert R.shape == (120,)
        assert abs(c[0] - cx) < 1e-9 and abs(c[1] - cy) < 1e-9
        assert np.max(np.abs(R - R0)) < 0.01
    f = droplet_slice(N, R0, 2.0, 47.25, 48.6).astype(np.float32)
    R, c = ttsh.contour_radius(f)
    assert abs(c[0] - 47.25) < 0.2 and abs(c[1] - 48.6) < 0.2
    assert np.max(np.abs(R - R0)) < 0.2


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
