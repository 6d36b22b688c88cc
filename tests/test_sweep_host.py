"""forEach.pl's knob grid (ttsh3d.sweep): order, run-directory names and the
kPhi = aPhi behaviour.  No GPU needed."""
import pytest


def test_default_point_name_matches_run_cuda(ttsh):
    """run_cuda.pl:90 holds the name forEach.pl:70 gives its own default
    point, so the naming is pinned by the reference's file."""
    from ttsh3d import sweep
    pts = sweep.grid()
    assert len(pts) == 1
    name, p = pts[0]
    assert name == ("eta_-0.187-LA_1-alpha_-0.02-beta_4.366-mu_1-cA_-1-A_0-lambda0_9-aPhi_1-kPhi_1-gamma_1"
                    "-N_64-h_1-T_1000-DT_10")
    assert (p.Nx, p.dt0, p.Ts, p.dte, p.init0) == (64, 0.002, 1000.0, 10.0, 0.0001)


def test_grid_order_is_forEach_loop_order(ttsh):
    from ttsh3d import sweep
    pts = sweep.grid({"mu": [1, 2], "cA": [-1, -0.5], "gamma": [1, 3]})
    keys = [(p.mu, p.cA, p.gamma) for _, p in pts]
    assert keys == [(1, -1, 1), (1, -1, 3), (1, -0.5, 1), (1, -0.5, 3),
                    (2, -1, 1), (2, -1, 3), (2, -0.5, 1), (2, -0.5, 3)]
    assert len({n for n, _ in pts}) == 8


def test_kphi_takes_aphi_unless_fixed(ttsh):
    """forEach.pl:56 writes kPhi = $aPhi, and :70 names the run with it."""
    from ttsh3d import sweep
    (n1, p1), = sweep.grid({"aPhi": [2], "kPhi": [5]})
    assert p1.kPhi == 2 and "-aPhi_2-kPhi_2-" in n1
    (n2, p2), = sweep.grid({"aPhi": [2], "kPhi": [5]}, fix_kphi=True)
    assert p2.kPhi == 5 and "-aPhi_2-kPhi_5-" in n2


def test_perl_number_format():
    from ttsh3d import sweep
    assert [sweep.perl_num(x) for x in (1.0, -0.187, 0.0001, 1e-5, 4.366, 1000.0)] == \
        ["1", "-0.187", "0.0001", "1e-05", "4.366", "1000"]


def test_run_sweep_rejects_mixed_globals(ttsh):
    from ttsh3d import sweep
    a = sweep.grid(globals_={"N": 32})
    b = sweep.grid(globals_={"N": 48})
    with pytest.raises(ValueError):
        sweep.run_sweep(a + b, "/nonexistent")
