"""GPU tests of the on-device diagnostics and of the run loop's exports:
- centroid and centroid slice (with the 4th-order vorticities) against the
  restatement of the reference's post-processing on the downloaded state;
- ttsh_run_ex: text snapshots written by the host thread are byte-identical
  to the synchronous writer's, the binary sidecar holds the state, and
  diag.dat / a_R.txt hold the centroid and contour radii of each export."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from oracle import postproc  # noqa: E402

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def evolved(ttsh, N=48, steps=30, seed=21, h=1.0):
    p = ttsh.default_params(N=N, dt0=0.01, cA=-2.0, h=h)
    phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 0, seed)
    # move the droplet off the box centre so the centroid is not trivial
    phi = np.roll(phi, (3, -5, 7), axis=(0, 1, 2))
    P = np.roll(P, (3, -5, 7), axis=(1, 2, 3))
    sim = ttsh.Simulation(p)
    sim.set_state(phi, P, v)
    sim.step(steps)
    return sim, p


def test_centroid_and_slice(ttsh):
    log("[centroid_and_slice]")
    sim, p = evolved(ttsh, h=1.3)
    with sim:
        phi, P, v = sim.get_state()
        idx, c, m = sim.centroid()
        idx_r, c_r, m_r = postproc.centroid(phi)
        log("centroid", idx, c, m, "restated", idx_r, c_r, m_r)
        assert idx == idx_r
        assert np.allclose(c, c_r, rtol=1e-10) and m == pytest.approx(m_r, rel=1e-10)
        f, z = sim.slice()
        assert z == idx[2]
        ref = postproc.slice_fields(phi, P, v, z, p.h)
        for k in ("phi", "vx", "vy", "vz", "px", "py", "pz"):
            assert np.array_equal(f[k], ref[k].astype(np.float32)), k
        for k in ("vort_v", "vort_p"):
            scale = np.abs(ref[k]).max()
            err = np.abs(f[k] - ref[k]).max() / scale
            log(k, "max scaled error", err)
            assert err < 1e-5, k
        f2, z2 = sim.slice(5)
        assert z2 == 5 and np.array_equal(f2["phi"], phi[5])


def test_run_exports(ttsh, tmp_path):
    """Text through the export thread equals the synchronous writer's bytes;
    the sidecar equals the state; diag lines follow the exports."""
    log("[run_exports]")
    N = 32
    p = ttsh.default_params(N=N, dt0=0.01, Ts=0.3, dte=0.1)
    phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 0, 5)
    runs = {}
    for name, kw in (("async", dict(binary=True, diag=True)), ("sync", dict(sync=True))):
        d = tmp_path / name
        d.mkdir()
        with ttsh.Simulation(p) as sim:
            sim.set_state(phi, P, v)
            steps = sim.run(str(d), **kw)
            final = sim.get_state()
            runs[name] = (d, steps, final)
    (da, sa, fa), (ds, ss, fs) = runs["async"], runs["sync"]
    assert sa == ss
    confs = sorted(f for f in os.listdir(ds) if f.endswith(".dat"))
    log("exports", confs, "steps", sa)
    assert confs == sorted(f for f in os.listdir(da) if f.startswith("conf_") and f.endswith(".dat"))
    for f in confs:
        assert open(da / f, "rb").read() == open(ds / f, "rb").read(), f
    # binary sidecar of the last export holds the final state
    last = max(int(f[5:-4]) for f in confs)
    fields, m, T, steps = ttsh.read_conf_bin(str(da / f"conf_{last}.bin"))
    assert m == last and steps == sa
    phi_f, P_f, v_f = fa
    assert np.array_equal(fields["phi"], phi_f)
    for a, k in enumerate(("px", "py", "pz")):
        assert np.array_equal(fields[k], P_f[a])
    for a, k in enumerate(("vx", "vy", "vz")):
        assert np.array_equal(fields[k], v_f[a])
    # diagnostics of the last export agree with the restatement on that state
    diag = np.loadtxt(da / "diag.dat", ndmin=2)
    aR = np.loadtxt(da / "a_R.txt", ndmin=2)
    assert diag.shape[0] == len(confs) and aR.shape == (len(confs), 120)
    assert int(diag[-1, 0]) == last
    idx_r, c_r, m_r = postproc.centroid(phi_f)
    assert tuple(int(x) for x in diag[-1, 2:5]) == idx_r
    R_ref, _ = postproc.contour_radius(phi_f[idx_r[2]].astype(np.float64))
    assert np.allclose(aR[-1], R_ref, rtol=0, atol=1e-9)


def test_run_diag_only(ttsh, tmp_path):
    """--no-text --diag: diagnostics without full-field snapshots."""
    log("[run_diag_only]")
    N = 32
    p = ttsh.default_params(N=N, dt0=0.01, Ts=0.2, dte=0.05)
    phi, P, v = ttsh.init_droplet(N, 0.3 * N, 2.0, 1, 9)
    with ttsh.Simulation(p) as sim:
        sim.set_state(phi, P, v)
        sim.run(str(tmp_path), text=False, diag=True)
    names = os.listdir(tmp_path)
    assert not any(n.startswith("conf_") for n in names)
    diag = np.loadtxt(tmp_path / "diag.dat", ndmin=2)
    log("diag rows", diag.shape[0])
    assert diag.shape[0] >= 4 and np.all(np.diff(diag[:, 0]) == 1)
