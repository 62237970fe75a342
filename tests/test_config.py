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


def test_read_reference_u0_layout(tmp_path):
    """The reference reads ni*nk doubles, x fastest, bottom row first
    (odeSolver.cpp:96-128); a file of known values comes back in place."""
    ni, nk = 7, 5
    vals = np.array([100.0 * k + i for k in range(nk) for i in range(ni)])
    path = tmp_path / "U0.bin"
    vals.astype("<f8").tofile(path)
    u = config.read_reference_u0(str(path), ni, nk)
    assert u.shape == (nk, ni)
    assert u[0, 0] == 0.0 and u[0, ni - 1] == ni - 1 and u[nk - 1, 2] == 100.0 * (nk - 1) + 2
    # placed as the interior of the (nz, nx) field: row 1 = bottom, column 1 = fault
    pb = config.make_reference_problem(ni, nk, 100.0, U0=u)
    np.testing.assert_array_equal(pb.U0[1:-1, 1:-1], u)
    assert pb.U0[1, 1] == 0.0 and pb.U0[nk, ni] == 100.0 * (nk - 1) + ni - 1
    # ghosts by extend_crew's rule
    np.testing.assert_array_equal(pb.U0[0], pb.U0[2])
    np.testing.assert_array_equal(pb.U0[:, 0], 2 * pb.U0[:, 1] - pb.U0[:, 2])
    import pytest
    vals[:-1].tofile(path)
    with pytest.raises(ValueError):
        config.read_reference_u0(str(path), ni, nk)


def test_reference_problem_set_up():
    """The reference's hard-coded problem (odeSolver.cpp:294-319,
    fault.cpp:85-120): elliptical basin mu 8..36 GPa, a = 0.015, b = 0.020
    tapering to 0 between 12 and 17 km, eta = sqrt(rho mu)/2."""
    pb = config.make_reference_problem(201, 201, 120.0)
    X, Z = config.coords(201, 201, 120.0)
    mu = pb.mu[1:-1, 1:-1]
    assert mu.min() >= 8.0 and mu.max() <= 36.0
    assert abs(mu[-1, 0] - 8.0) < 0.5          # at the fault, at the surface
    assert abs(mu[0, -1] - 36.0) < 1e-6        # far corner
    d = -Z[1:-1, 1]
    b = pb.b[1:-1]
    assert np.all(b[d <= 12.0] == 0.020) and np.all(b[d >= 17.0] == 0.0)
    assert np.all(pb.a[1:-1] == 0.015)
    assert pb.eta[-2] < pb.eta[1]              # softer near the surface


def test_workload_files_match_the_builders():
    import dataclasses
    pb = config.workload("basin_1025")
    ref = config.make_bp1_problem(1025, basin=True, tol_mode=config.TOL_REL, tol=1e-10)
    for k in ("mu", "a", "b", "eta", "tau_n0", "U0"):
        np.testing.assert_array_equal(getattr(pb, k), getattr(ref, k))
    p = pb.params
    assert (p.n_levels, p.tol_mode, p.tol, p.cycle, p.n_pre, p.n_post, p.n_coarse,
            p.relax) == (7, config.TOL_REL, 1e-10, config.CYCLE_FAST, 1, 3, 0, 1.15)
    assert p.tmax == config.tmax_reference() and p.max_steps == 3000
    bp = config.workload("bp1_129")
    assert dataclasses.asdict(bp.params) == dataclasses.asdict(
        dataclasses.replace(config.make_bp1_problem(129).params, max_steps=2000))
    r = config.workload("ref_basin_801")
    assert (r.params.ni, r.params.dh, r.params.n_levels) == (801, 30.0, 6)
    assert config.workload("basin_2049").params.precision == config.PREC_FP32_IR
    assert set(config.WORKLOADS) == {"bp1_129", "basin_1025", "basin_2049", "ref_basin_801"}


def test_params_file_keys_and_errors(tmp_path):
    import json
    import pytest

    def load(d, name="p.json"):
        path = tmp_path / name
        path.write_text(json.dumps(d))
        return config.load_params_file(str(path))

    # SetParams.cpp's time fields give its tmax; the reference's other names
    pb = load({"problem": "bp1", "ni": 65, "t_year": 1, "t_min": 2, "totTol": 1e-8,
               "maxNumSteps": 50, "Dc": 0.004, "sigma_n": 30.0, "tol_mode": "rel",
               "tol": 1e-11, "cycle": "fast", "n_pre": 2, "n_post": 2, "n_coarse": 0})
    p = pb.params
    assert p.tmax == 120.0 + 60.0 * 60.0 * 24.0 * 30.0 * 12.0
    assert (p.totTol, p.max_steps, p.L, p.tol_mode, p.tol) == (1e-8, 50, 0.004,
                                                              config.TOL_REL, 1e-11)
    assert np.all(pb.tau_n0[1:-1] == -30.0)
    # a U0 file in the reference's format, relative to the parameter file
    u = np.arange(9 * 9, dtype=np.float64).reshape(9, 9)
    u.tofile(tmp_path / "u0.bin")
    r = load({"problem": "reference", "ni": 9, "nk": 9, "dh": 500.0, "n_levels": 2,
              "u0_file": "u0.bin"})
    np.testing.assert_array_equal(r.U0[1:-1, 1:-1], u)
    bad = [({"ni": 65}, "required"), ({"problem": "bp1", "nii": 65}, "unknown"),
           ({"problem": "bp1", "ni": "65"}, "integer"),
           ({"problem": "bp1", "tol_mode": "relative"}, "one of"),
           ({"problem": "bp1", "ni": 65, "nk": 33}, "square"),
           ({"problem": "bp1", "ni": 65, "dh": 30.0}, "dh"),
           ({"problem": "bp1", "tmax": 1.0, "t_year": 1}, "either"),
           ({"problem": "bp1", "ni": 17, "n_levels": 5}, "levels"),
           ({"problem": "bp1", "mindt": 0.0}, "mindt"),
           ({"problem": "reference", "sigma_n": 40.0}, "fixed")]
    for d, msg in bad:
        with pytest.raises(config.ParamsError, match=msg):
            load(d)
    (tmp_path / "x.json").write_text("[1, 2]")
    with pytest.raises(config.ParamsError):
        config.load_params_file(str(tmp_path / "x.json"))
    with pytest.raises(KeyError):
        config.workload("nope")
