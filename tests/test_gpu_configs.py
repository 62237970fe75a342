"""GPU tests at the workloads' own sizes.

* basin_1025 (the headline workload): 1025^2 basin medium, 7 levels, the
  production V(1,3) cycle with relaxation 1.15 at a 1e-10 relative residual,
  autotuned tiles: the RK43 run equals the oracle bit for bit from t = 0, and
  again after a restart from the engine's state 600 steps in (larger steps,
  two or more V-cycles per solve).  One V-cycle of the reference's own cycle
  (32 sweeps per level) at the same size equals the oracle's.
* ensemble: members of the seeded 64-member draw at 257^2, 513^2 and
  1025^2, FP32-IR and FP64, stepped together; each equals its solo run bit
  for bit and the FP64 member equals the oracle.
* learned initial guess at 1025^2: trained on pairs harvested from a
  perturbed run; guesses are accepted, the slip at a common end time is
  within 1e-6 of the plain run's, and a non-finite prediction (NaN or inf)
  falls back to the plain run bit for bit.  (A finite but absurd one, x1e6,
  is not always refused here: the candidate also carries the exact response
  to the remote-edge change, so when the fault data barely change it can
  still beat the previous solution; tests/test_gpu_guess.py checks x1e6 at
  129^2.)
* the reference's own problem (801^2 elliptical basin, 6 levels, reference
  cycle, absolute 1e-9 stop rule, SetParams.cpp / odeSolver.cpp:272-319 /
  fault.cpp:73-135) with its initial displacement read from a file in the
  reference's format (odeSolver.cpp:96-128): equal to the oracle bit for bit.

The oracle's 1025^2 and 801^2 runs use the device-built mu_equi (checked bit
for bit against the oracle's at smaller sizes in test_gpu_parity.py): the
oracle's own build takes minutes at these sizes.
"""
import numpy as np
import pytest

import oracle
from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu

PROD = dict(cycle=config.CYCLE_FAST, n_pre=1, n_post=3, n_coarse=0, relax=1.15)


def _basin_1025():
    return config.make_bp1_problem(1025, basin=True, tol_mode=config.TOL_REL, tol=1e-10)


def _oracle_sim(pb, me, **cyc):
    oracle.set_libm(False)
    args = dict(cycle=1, n_pre=cyc["n_pre"], n_post=cyc["n_post"],
                n_coarse=cyc["n_coarse"], relax=cyc["relax"]) if cyc else {}
    return oracle.Sim(pb, me, **args)


def _compare_steps(sim, eng, steps, tag):
    for s in range(steps):
        sim.step()
        info = eng.rk43_step(1)
        ref = sim.scalars()
        for key in ("it", "t", "dt", "totErr", "V_max", "Ts_max", "attempts"):
            assert info[key] == ref[key], (tag, s, key, info[key], ref[key])
    st, ar = eng.state(), sim.arrays()
    for key in ("slip", "psi", "V", "dpsi", "Ts", "U"):
        np.testing.assert_array_equal(st[key], ar[key], err_msg=f"{tag} {key}")
    return st


def test_basin_1025_production_cycle_matches_oracle():
    pb = _basin_1025()
    me = lib.build_mu_equi(pb.mu)
    sim = _oracle_sim(pb, me, **PROD)
    sim.init()
    with lib.Engine(pb, me, **PROD) as eng:
        eng.set_tiles(True, 64, 16)
        eng.autotune_tiles(3)
        assert eng.uses_tiled_cycle
        eng.rk43_setup()
        v0 = eng.state()["info"]["vcycles"]
        _compare_steps(sim, eng, 10, "from t=0")
        # later in the interseismic period: restart the oracle from the
        # engine's state and step both
        eng.rk43_step(590)
        st = eng.state()
        i = st["info"]
        sim2 = _oracle_sim(pb, me, **PROD)
        sim2.set_state(i["t"], i["dt"], i["it"], i["errA0"], i["errA1"], st["U"],
                       st["slip"], st["psi"])
        vc_before = i["vcycles"]
        _compare_steps(sim2, eng, 8, "after 600 steps")
        i2 = eng.state()["info"]
        vps = (i2["vcycles"] - vc_before) / max(i2["solves"] - i["solves"], 1)
        print(f"1025^2: t {i['t']:.4g} -> {i2['t']:.4g} s, {vps:.2f} V-cycles/solve "
              f"in the compared steps; {v0} V-cycles in the t = 0 solve")
        assert vps >= 1.0


def test_basin_1025_reference_cycle_vcycle_matches_oracle():
    pb = _basin_1025()
    p = pb.params
    me = lib.build_mu_equi(pb.mu)
    rng = np.random.default_rng(21)
    U = pb.U0.copy()
    U[1:-1, 2:-2] += rng.normal(scale=0.05, size=(p.nk, p.ni - 2))
    mg = oracle.MG(p.ni, p.nk, p.dh, p.n_levels, me)
    ref = U.copy()
    st, vc, _, _ = mg.solve(ref, oracle.ORC_TOL_ABS, 1e-300, 1)
    assert vc == 1
    with lib.Engine(pb, me, tol_mode=config.TOL_ABS, tol=1e-300, max_vcycles=1) as eng:
        V = U.copy()
        with pytest.raises(lib.QdeqError) as ei:
            eng.ode_solve(V)
        assert ei.value.status == lib.QDEQ_ERR_MG_NOT_CONVERGED
    np.testing.assert_array_equal(V, ref)


ENS_LEVELS = {257: 5, 513: 6, 1025: 7}
ENS_MODES = {"fp64": dict(precision=config.PREC_FP64, tiles=(64, 16)),
             "ir32": dict(precision=config.PREC_FP32_IR, tiles=(64, 32))}


def _ensemble_cases():
    """The first member of each grid size in the seeded 64-member draw, in
    FP32-IR, and the first 257^2 member again in FP64 (oracle-checked)."""
    draws = config.ensemble_members(64, seed=0)
    first = {}
    for d in draws:
        first.setdefault(d["n"], d)
    cases = [(first[n], "ir32") for n in (257, 513, 1025)]
    cases.append((first[257], "fp64"))
    return cases


def _ens_problem(d):
    return config.make_bp1_problem(d["n"], basin=True, levels=ENS_LEVELS[d["n"]],
                                   Dc=d["Dc"], sigma_n=d["sigma_n"], b_seis=d["b_seis"],
                                   tol_mode=config.TOL_REL, tol=1e-10)


def test_ensemble_uneven_grids_equal_solo_runs_and_oracle():
    cases = _ensemble_cases()
    steps = 60
    media, shapes = {}, {}
    pbs = [_ens_problem(d) for d, _ in cases]

    def make(i):
        d, mode = cases[i]
        pb = pbs[i]
        if d["n"] not in media:
            media[d["n"]] = lib.build_mu_equi(pb.mu)
        m = ENS_MODES[mode]
        eng = lib.Engine(pb, media[d["n"]], max_vcycles=200, precision=m["precision"],
                         **PROD)
        eng.set_tiles(True, *m["tiles"])
        key = (d["n"], mode)
        if key not in shapes:
            shapes[key] = eng.autotune_tiles(3)
        else:
            eng.set_tile_shapes(shapes[key])
        eng.rk43_setup()
        return eng

    solo = []
    for i in range(len(cases)):
        with make(i) as e:
            e.rk43_step(steps)
            solo.append(e.state())
    engines = [make(i) for i in range(len(cases))]
    status = lib.ensemble_step(engines, steps)
    assert status == [lib.QDEQ_OK] * len(cases)
    for i, e in enumerate(engines):
        st = e.state()
        assert st["info"]["t"] == solo[i]["info"]["t"], i
        assert st["info"]["vcycles"] == solo[i]["info"]["vcycles"], i
        for k in ("slip", "psi", "V", "U"):
            np.testing.assert_array_equal(st[k], solo[i][k], err_msg=f"member {i} {k}")
    # the FP64 member against the oracle
    i64 = [i for i, (_, m) in enumerate(cases) if m == "fp64"][0]
    pb = pbs[i64]
    sim = _oracle_sim(pb, media[pb.params.ni], **PROD)
    sim.init()
    for _ in range(steps):
        sim.step()
    ar, st = sim.arrays(), engines[i64].state()
    assert st["info"]["t"] == sim.scalars()["t"]
    for k in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(st[k], ar[k], err_msg=k)
    for e in engines:
        e.close()


FAST_LEARNED = dict(PROD)


def _run_1025(pb, me, steps=None, tmax=None, guess=None, mode=lib.GUESS_OFF,
              inject=None):
    kw = dict(FAST_LEARNED)
    if tmax is not None:
        kw["tmax"] = tmax
    with lib.Engine(pb, me, **kw) as eng:
        eng.set_tiles(True, 64, 16)
        eng.rk43_setup()
        if guess is not None:
            eng.attach_guess(guess, mode)
            if inject is not None:
                eng.guess_inject(inject)
        if tmax is None:
            eng.rk43_step(steps)
        else:
            eng.rk43_step(10 ** 6)
        st = eng.state()
        stats = eng.guess_stats() if guess is not None else None
    return st, stats


def test_learned_guess_at_1025():
    pb = _basin_1025()
    me = lib.build_mu_equi(pb.mu)
    g = lib.Guess(1025, capacity=2500, stride=2)
    # harvest from a run with other frictional parameters than the test run
    pb_h = config.make_bp1_problem(1025, basin=True, tol_mode=config.TOL_REL, tol=1e-10,
                                   Dc=0.0065, sigma_n=55.0)
    _run_1025(pb_h, me, steps=400, guess=g, mode=lib.GUESS_HARVEST)
    assert g.pairs > 800
    loss = g.train(epochs=30, batch=256)
    assert np.isfinite(loss)
    steps = 300
    s0, _ = _run_1025(pb, me, steps=steps)
    T = s0["info"]["t"]
    s1, st = _run_1025(pb, me, tmax=T, guess=g, mode=lib.GUESS_PREDICT)
    assert s1["info"]["t"] == T
    dslip = np.max(np.abs(s1["slip"] - s0["slip"])) / np.max(np.abs(s0["slip"]))
    i0, i1 = s0["info"], s1["info"]
    print(f"1025^2 learned: V-cycles/solve {i1['vcycles'] / i1['solves']:.3f} "
          f"(plain {i0['vcycles'] / i0['solves']:.3f}), accepted "
          f"{st['accepted']}/{st['predicted']}, slip rel diff {dslip:.2e}")
    assert st["predicted"] > 1000 and st["accepted"] > 0.3 * st["predicted"]
    assert dslip <= 1e-6
    # a non-finite prediction is never used: bit for bit the plain run
    for factor in (float("nan"), float("inf")):
        s2, st2 = _run_1025(pb, me, steps=40, guess=g, mode=lib.GUESS_PREDICT,
                            inject=factor)
        s3, _ = _run_1025(pb, me, steps=40)
        assert st2["predicted"] > 100 and st2["accepted"] == 0
        assert s2["info"]["t"] == s3["info"]["t"]
        for k in ("slip", "psi", "V", "U"):
            np.testing.assert_array_equal(s2[k], s3[k], err_msg=f"{factor} {k}")
    g.close()


def test_reference_problem_801_with_u0_file_matches_oracle(tmp_path):
    base = config.make_reference_problem()
    p = base.params
    # an initial displacement in the reference's file format: ni*nk FP64
    # values, x fastest, bottom row first; here the uniform-prestress field
    # plus a smooth bump, so that it differs from the built-in default
    X, Z = config.coords(p.ni, p.nk, p.dh)
    u0 = base.U0[1:-1, 1:-1] + 0.02 * np.exp(-((X[1:-1, 1:-1] - 6.0) ** 2 +
                                               (Z[1:-1, 1:-1] + 9.0) ** 2) / 8.0)
    path = tmp_path / "U0.bin"
    np.ascontiguousarray(u0, dtype=np.float64).tofile(path)
    pb = config.make_reference_problem(U0=config.read_reference_u0(str(path), p.ni, p.nk))
    np.testing.assert_array_equal(pb.U0[1:-1, 1:-1], u0)
    me = lib.build_mu_equi(pb.mu)
    sim = _oracle_sim(pb, me)
    sim.init()
    with lib.Engine(pb, me) as eng:
        eng.rk43_setup()
        _compare_steps(sim, eng, 25, "ref_basin_801")
        info = eng.state()["info"]
        print(f"801^2 reference problem: 25 steps to t={info['t']:.4g} s, "
              f"{info['vcycles']} V-cycles, {info['solves']} solves")
