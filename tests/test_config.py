"""Problem set-up: the workloads' inputs match their stated physics."""
import math

import numpy as np

from qdeqcycle_fdm_amd import config


def test_bp1_prestress_and_initial_state():
    pb = config.make_bp1_problem(129)
    p = pb.params
    assert (p.ni, p.nk, p.n_levels) == (129, 129, 4)
    assert p.dh == 375.0
    tau0 = config.steady_prestress(50.0, 0.025, 0.015, 0.6, 1e-6, 1e-9, 32 / (2 * 3.464))
    assert abs(tau0 - 26.546) < 1e-3
    # linear U0: slip0 = 0, SlipShift = tau0 * 48 km / 32e3
    assert np.all(pb.U0[1:-1, 1] == 0.0)
    np.testing.assert_allclose(pb.U0[1:-1, p.ni], tau0 * 48000.0 / 32e3, rtol=1e-14)
    # ghost layer follows extrapolation / mirror rules
    np.testing.assert_allclose(pb.U0[:, 0], 2 * pb.U0[:, 1] - pb.U0[:, 2])
    np.testing.assert_array_equal(pb.U0[0], pb.U0[2])


def test_bp1_friction_profile():
    pb = config.make_bp1_problem(129)
    X, Z = config.coords(129, 129, pb.params.dh)
    d = -Z[1:-1, 1]
    a = pb.a[1:-1]
    assert np.all(a[d <= 15.0] == 0.010)
    assert np.all(a[d >= 18.0] == 0.025)
    mid = (d > 15.0) & (d < 18.0)
    np.testing.assert_allclose(a[mid], 0.010 + 0.005 * (d[mid] - 15.0))
    assert np.all(pb.b[1:-1] == 0.015)
    assert np.all(pb.tau_n0[1:-1] == -50.0)
    np.testing.assert_allclose(pb.eta[1:-1], 32.0 / (2 * 3.464))
    # coordinates: row nk is the free surface, column 1 the fault
    assert Z[-2, 1] == 0.0 and X[1, 1] == 0.0


def test_basin_medium():
    pb = config.make_bp1_problem(1025, basin=True)
    assert pb.params.n_levels == 7
    X, Z = config.coords(1025, 1025, pb.params.dh)
    mu = pb.mu[1:-1, 1:-1]
    x, d = X[1:-1, 1:-1], -Z[1:-1, 1:-1]
    assert abs(mu[(x < 1) & (d < 0.5)].mean() - 20.0) < 0.2
    assert np.all(np.abs(mu[(x > 15) | (d > 8)] - 32.0) < 1e-6)
    assert mu.min() >= 20.0 - 1e-9 and mu.max() <= 32.0


def test_ensemble_draws_are_seeded():
    a = config.ensemble_members(64, seed=0)
    b = config.ensemble_members(64, seed=0)
    assert a == b and len(a) == 64
    assert {m["n"] for m in a} <= {257, 513, 1025}
    for m in a:
        assert 0.004 <= m["Dc"] <= 0.016
        assert 25 <= m["sigma_n"] <= 75
        assert 0.013 <= m["b_seis"] <= 0.018


def test_reference_tmax():
    assert config.tmax_reference() == 60.0 + 60.0 * 60.0 * 24.0 * 30.0 * 12.0 * 3000
