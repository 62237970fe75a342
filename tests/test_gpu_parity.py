"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Integer-free FP64 path: the reference V-cycle, boundary handling, stress
stencil, Newton slip-rate solve and state rate are compared bit for bit (the
Newton solve and the state rate use the engine's own exp/sinh/cosh/asinh,
of which the oracle carries identical copies).
"""
import numpy as np
import pytest

import oracle
from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu


def _problem(n, basin=True, levels=None, **kw):
    return config.make_bp1_problem(n, basin=basin, levels=levels, **kw)


def _random_medium(n, seed):
    rng = np.random.default_rng(seed)
    mu = np.zeros((n + 2, n + 2))
    mu[1:-1, 1:-1] = np.exp(rng.normal(np.log(25.0), 0.4, size=(n, n)))
    return mu


@pytest.mark.parametrize("n,kind", [(33, "basin"), (65, "random"), (17, "uniform")])
def test_mu_equi_device_bitwise(n, kind):
    if kind == "basin":
        mu = _problem(n).mu
    elif kind == "random":
        mu = _random_medium(n, 7)
    else:
        mu = _problem(n, basin=False).mu
    ref = oracle.mu_equi(mu)
    got = lib.build_mu_equi(mu)
    np.testing.assert_array_equal(got, ref)


def test_set_boundary_and_stress_bitwise():
    pb = _problem(33)
    rng = np.random.default_rng(3)
    nz, nx = pb.mu.shape
    U = rng.normal(size=(nz, nx))
    slip = rng.normal(size=nz)
    ss = rng.normal(size=nz)
    p = pb.params
    with lib.Engine(pb, oracle.mu_equi(pb.mu)) as eng:
        U_gpu = U.copy()
        eng.setBoundary(U_gpu, slip, ss, 12345.5, 0.83, 17.25)
        U_ref = U.copy()
        oracle.set_boundary(U_ref, p.ni, p.nk, slip, ss, p.Vini, 12345.5, 0.83, 17.25)
        np.testing.assert_array_equal(U_gpu, U_ref)
        Ts, Tn = eng.eval_stress_on_fault(U_ref, pb.mu, pb.tau_n0)
        Ts_r, Tn_r = oracle.eval_stress(U_ref, pb.mu, pb.tau_n0, p.ni, p.nk, p.dh)
        np.testing.assert_array_equal(Ts, Ts_r)
        np.testing.assert_array_equal(Tn, Tn_r)


def test_nr_and_dpsi_match_oracle():
    pb = _problem(129, basin=False)
    p = pb.params
    rng = np.random.default_rng(11)
    nz = p.nz
    Ts = np.zeros(nz)
    psi = np.zeros(nz)
    Ts[1:-1] = rng.uniform(5.0, 40.0, nz - 2)
    psi[1:-1] = rng.uniform(0.5, 0.9, nz - 2)
    oracle.set_libm(False)
    with lib.Engine(pb, oracle.mu_equi(pb.mu)) as eng:
        V = eng.NRsearch(Ts, pb.tau_n0, psi, pb.a, pb.eta)
        st, V_r, _ = oracle.nr_search(Ts, pb.tau_n0, psi, pb.a, pb.eta, p.V0)
        assert st == 0
        np.testing.assert_array_equal(V[1:-1], V_r[1:-1])
        d = eng.cal_dpsi(psi, V_r, pb.b)
        d_r = oracle.cal_dpsi(psi, V_r, pb.b, p.f0, p.V0, p.L)
        np.testing.assert_array_equal(d[1:-1], d_r[1:-1])
        # guard branch: b <= 1e-3 gives 0 (odeSolver.cpp:1200)
        b0 = pb.b.copy()
        b0[5:9] = 1e-3
        d0 = eng.cal_dpsi(psi, V_r, b0)
        assert np.all(d0[5:9] == 0.0)


def _perturbed_u(pb, seed):
    rng = np.random.default_rng(seed)
    U = pb.U0.copy()
    U[1:-1, 2:-2] += rng.normal(scale=0.5, size=(pb.params.nk, pb.params.ni - 2))
    return U


@pytest.mark.parametrize("n,levels", [(33, 3), (65, 4)])
@pytest.mark.parametrize("ncap", [1, 2, 3])
def test_reference_vcycles_bitwise(n, levels, ncap):
    """n V-cycles of the reference cycle give the oracle's iterate exactly."""
    pb = _problem(n, levels=levels)
    p = pb.params
    me = oracle.mu_equi(pb.mu)
    U = _perturbed_u(pb, 5)
    mg = oracle.MG(p.ni, p.nk, p.dh, levels, me)
    U_ref = U.copy()
    st, vc, _, _ = mg.solve(U_ref, oracle.ORC_TOL_ABS, 1e-300, ncap)
    assert vc == ncap
    with lib.Engine(pb, me, tol=1e-300, max_vcycles=ncap) as eng:
        U_gpu = U.copy()
        with pytest.raises(lib.QdeqError) as ei:
            eng.ode_solve(U_gpu)
        assert ei.value.status == lib.QDEQ_ERR_MG_NOT_CONVERGED
    np.testing.assert_array_equal(U_gpu, U_ref)


@pytest.mark.parametrize("n,levels", [(65, 4), (129, 4)])
def test_reference_solve_abs_tol(n, levels):
    """Full solve with the reference stop rule: same V-cycle count, same U."""
    pb = _problem(n, levels=levels)
    p = pb.params
    me = oracle.mu_equi(pb.mu)
    U = _perturbed_u(pb, 9)
    mg = oracle.MG(p.ni, p.nk, p.dh, levels, me)
    U_ref = U.copy()
    st, vc_ref, _, res_ref = mg.solve(U_ref, oracle.ORC_TOL_ABS, 1e-9, 500)
    assert st == 0
    with lib.Engine(pb, me) as eng:
        U_gpu = U.copy()
        vc, _, res = eng.ode_solve(U_gpu)
    assert vc == vc_ref
    np.testing.assert_array_equal(U_gpu, U_ref)
    assert res <= 1e-9
    np.testing.assert_allclose(res, res_ref, rtol=1e-12)


def _rk_compare(pb, steps, **overrides):
    """Step-by-step comparison of the device-resident RK43 driver with the
    oracle: every scalar and every state array bit for bit."""
    me = oracle.mu_equi(pb.mu)
    oracle.set_libm(False)
    sim = oracle.Sim(pb, me)
    sim.init()
    with lib.Engine(pb, me, **overrides) as eng:
        eng.rk43_setup()
        for s in range(steps):
            sim.step()
            info = eng.rk43_step(1)
            ref = sim.scalars()
            for key in ("it", "attempts", "t", "dt", "totErr", "V_max", "Ts_max"):
                assert info[key] == ref[key], (s, key, info[key], ref[key])
            assert info["vcycles"] == ref["vcycles"], s
        st = eng.state()
        ar = sim.arrays()
        for key in ("slip", "psi", "V", "dpsi", "Ts", "U"):
            np.testing.assert_array_equal(st[key], ar[key], err_msg=key)
        return st, ar


def test_rk43_bp1_129_first_steps():
    """First 60 accepted RK43 steps of the 129^2 workload vs the oracle."""
    _rk_compare(_problem(129, basin=False), 60)


def test_rk43_basin_65_steps():
    """Heterogeneous medium, 65^2, 40 steps."""
    _rk_compare(_problem(65, basin=True, levels=3), 40)


def test_rk43_bp1_129_golden_2000_steps():
    """2000 accepted steps (five seismic events) against the committed oracle
    fixture: the whole step sequence and the fault state every 100 steps."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "bp1_129_oracle.npz"))
    names = [str(x) for x in g["names"]]
    series = g["series"]
    pb = _problem(129, basin=False)
    col = {n: i for i, n in enumerate(names)}
    with lib.Engine(pb, oracle.mu_equi(pb.mu)) as eng:
        eng.rk43_setup()
        eng.events_enable(1e-3)
        snap = 0
        for s in range(series.shape[0]):
            info = eng.rk43_step(1)
            for key in ("t", "dt", "V_max", "Ts_max", "attempts", "totErr",
                        "vcycles", "solves"):
                assert info[key] == series[s, col[key]], (s, key, info[key],
                                                           series[s, col[key]])
            if snap < len(g["snap_steps"]) and s + 1 == g["snap_steps"][snap]:
                st = eng.state()
                for key in ("slip", "psi", "V"):
                    np.testing.assert_array_equal(st[key], g[key][snap], err_msg=key)
                snap += 1
        events = eng.events()
    onsets = np.flatnonzero((series[1:, col["V_max"]] > 1e-3) &
                            (series[:-1, col["V_max"]] <= 1e-3))
    assert len(onsets) >= 2
    # the device event catalogue against the fixture's V_max series: step s
    # (0-based) of the series is accepted step it = s + 1
    vmax, t = series[:, col["V_max"]], series[:, col["t"]]
    ref = []
    s = 0
    while s < len(vmax):
        if vmax[s] > 1e-3:
            e = s
            while e < len(vmax) and vmax[e] > 1e-3:
                e += 1
            pk = s + int(np.argmax(vmax[s:e]))
            ref.append(dict(it_onset=s + 1, t_onset=t[s], v_peak=vmax[pk],
                            t_peak=t[pk], it_end=e + 1 if e < len(vmax) else -1,
                            t_end=t[e] if e < len(vmax) else -1.0))
            s = e
        else:
            s += 1
    assert len(events) == len(ref) == len(onsets) + (1 if vmax[0] > 1e-3 else 0)
    for got, exp in zip(events, ref):
        for key, val in exp.items():
            assert got[key] == val, (key, got[key], val)
        assert 1 <= got["node_onset"] <= 129 and got["slip_max"] > 0.0
