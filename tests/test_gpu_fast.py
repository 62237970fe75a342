"""GPU checks of the fast cycle (the engine's production solver).

* FP64 fast cycle on the tiled LDS kernels == the same cycle launched one
  half-sweep at a time == the oracle's fast cycle, bit for bit.
* FP32 V-cycles in FP64 iterative refinement reach the 1e-10 relative
  residual and the FP64 oracle's solution to the accuracy that residual
  implies; tiled and untiled IR are bit-identical.
* The RK43 driver in the fast FP64 mode matches the oracle step for step.
* The same for the production cycle of the bench, V(1,3) with relaxation
  1.15.
"""
import numpy as np
import pytest

import oracle
from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu

FAST = dict(cycle=config.CYCLE_FAST, n_pre=2, n_post=2, n_coarse=0)


def _perturbed(pb, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    U = pb.U0.copy()
    U[1:-1, 2:-2] += rng.normal(scale=scale, size=(pb.params.nk, pb.params.ni - 2))
    return U


def _fast_oracle_solve(pb, me, U, ncap, tol=1e-300, mode=oracle.ORC_TOL_ABS,
                      n_coarse=None):
    p = pb.params
    mg = oracle.MG(p.ni, p.nk, p.dh, p.n_levels, me)
    mg.set_cycle(FAST["n_pre"], FAST["n_post"],
                 FAST["n_coarse"] if n_coarse is None else n_coarse, True)
    V = U.copy()
    st, vc, r0, r = mg.solve(V, mode, tol, ncap)
    return V, vc, r0, r


def _gpu_solve(pb, me, U, tiled, tile=(64, 16), **kw):
    with lib.Engine(pb, me, **{**FAST, **kw}) as eng:
        eng.set_tiles(tiled, *tile)
        assert eng.uses_tiled_cycle == tiled
        V = U.copy()
        try:
            out = eng.ode_solve(V)
        except lib.QdeqError as e:
            if e.status != lib.QDEQ_ERR_MG_NOT_CONVERGED:
                raise
            out = None
    return V, out


@pytest.mark.parametrize("n,levels", [(65, 3), (129, 4), (257, 5), (129, 3)])
@pytest.mark.parametrize("ncap", [1, 3])
@pytest.mark.parametrize("n_coarse", [0, 32])
def test_fast_fp64_tiled_equals_untiled_equals_oracle(n, levels, ncap, n_coarse):
    """n_coarse = 0: exact coarsest solve; 32: coarsest sweeps."""
    pb = config.make_bp1_problem(n, basin=True, levels=levels)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 3)
    ref, vc, _, _ = _fast_oracle_solve(pb, me, U, ncap, n_coarse=n_coarse)
    assert vc == ncap
    untiled, _ = _gpu_solve(pb, me, U, False, tol=1e-300, max_vcycles=ncap,
                            n_coarse=n_coarse)
    np.testing.assert_array_equal(untiled, ref)
    for tile in ((64, 16), (32, 8), (128, 4)):
        tiled, _ = _gpu_solve(pb, me, U, True, tile=tile, tol=1e-300, max_vcycles=ncap,
                              n_coarse=n_coarse)
        np.testing.assert_array_equal(tiled, ref, err_msg=str(tile))


def test_fast_fp64_relative_tolerance_solve():
    pb = config.make_bp1_problem(257, basin=True, tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 4)
    ref, vc_ref, r0_ref, r_ref = _fast_oracle_solve(pb, me, U, 200, 1e-10,
                                                    oracle.ORC_TOL_REL)
    got, (vc, r0, r) = _gpu_solve(pb, me, U, True)
    assert vc == vc_ref
    assert r <= 1e-10 * r0
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_allclose(r0, r0_ref, rtol=1e-13)
    # the final norm from the level-0 up kernel's tile partials (fixed
    # wavefront and tile order) against the oracle's sequential sum over the
    # same, bit-identical residual: both lie within (n + 64) u of the exact
    # norm of n points (u = 2^-53)
    n_pts = pb.params.ni * pb.params.nk
    np.testing.assert_allclose(r, r_ref, rtol=2 * (n_pts + 64) * 2.0 ** -53)


@pytest.mark.parametrize("tiled", [False, True])
def test_ir_fp32_reaches_fp64_tolerance(tiled):
    """FP32 correction cycles in FP64 refinement: relative residual <= 1e-10
    and the solution within the error that residual allows of the FP64
    oracle's converged solution."""
    pb = config.make_bp1_problem(257, basin=True, tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 5)
    ref, _, _, _ = _fast_oracle_solve(pb, me, U, 400, 1e-13, oracle.ORC_TOL_REL)
    fp64, _ = _gpu_solve(pb, me, U, True)          # FP64 fast cycle, same tolerance
    got, (vc, r0, r) = _gpu_solve(pb, me, U, tiled, tile=(64, 32),
                                  precision=config.PREC_FP32_IR)
    assert r <= 1e-10 * r0
    assert vc < 40
    scale = np.max(np.abs(ref[1:-1, 1:-1]))
    err_ir = np.max(np.abs(got[1:-1, 1:-1] - ref[1:-1, 1:-1]))
    err_64 = np.max(np.abs(fp64[1:-1, 1:-1] - ref[1:-1, 1:-1]))
    # a 1e-10 relative residual bounds the error only through the operator's
    # condition number: both solves sit at that level, far below 1e-6
    assert err_ir <= 1e-6 * scale
    assert err_ir <= 10.0 * max(err_64, 1e-12 * scale)


def test_ir_tiled_equals_untiled():
    pb = config.make_bp1_problem(129, basin=True, tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 6)
    a, oa = _gpu_solve(pb, me, U, False, precision=config.PREC_FP32_IR)
    b, ob = _gpu_solve(pb, me, U, True, tile=(64, 32), precision=config.PREC_FP32_IR)
    assert oa[0] == ob[0]
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n,basin,levels,steps", [(129, False, 4, 60), (65, True, 3, 40)])
def test_rk43_fast_fp64_matches_oracle(n, basin, levels, steps):
    pb = config.make_bp1_problem(n, basin=basin, levels=levels,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    oracle.set_libm(False)
    sim = oracle.Sim(pb, me, cycle=1, n_pre=FAST["n_pre"], n_post=FAST["n_post"],
                     n_coarse=FAST["n_coarse"])
    sim.init()
    with lib.Engine(pb, me, **FAST) as eng:
        assert eng.uses_tiled_cycle
        eng.rk43_setup()
        for s in range(steps):
            sim.step()
            info = eng.rk43_step(1)
            ref = sim.scalars()
            for key in ("t", "dt", "totErr", "V_max", "attempts", "vcycles"):
                assert info[key] == ref[key], (s, key, info[key], ref[key])
        st = eng.state()
    ar = sim.arrays()
    for key in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(st[key], ar[key], err_msg=key)


@pytest.mark.parametrize("precision", [config.PREC_FP64, config.PREC_FP32_IR])
def test_autotuned_tiles_give_identical_solutions(precision):
    """The tile-shape autotuner only changes the launch shapes: the solve and
    the RK43 steps after tuning are bit-identical to the default shapes'."""
    pb = config.make_bp1_problem(257, basin=True, levels=5,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 9)
    out = []
    for tune in (False, True):
        with lib.Engine(pb, me, **{**FAST, "precision": precision}) as eng:
            eng.set_tiles(True, 64, 16)
            if tune:
                shapes = eng.autotune_tiles(3)
                assert len(shapes) == 3          # 257, 129, 65 run as tiles
                for down, up in shapes:
                    for tx, ty, rows in (down, up):
                        assert tx % 2 == 0 and ty % 2 == 0 and rows in (4, 8, 16)
            V = U.copy()
            vc = eng.ode_solve(V)[0]
            eng.rk43_setup()
            info = eng.rk43_step(25)
            st = eng.state()
            out.append((V, vc, info, st))
    (V0, vc0, i0, s0), (V1, vc1, i1, s1) = out
    np.testing.assert_array_equal(V0, V1)
    assert vc0 == vc1
    assert i0["t"] == i1["t"] and i0["vcycles"] == i1["vcycles"]
    for k in ("slip", "psi", "U"):
        np.testing.assert_array_equal(s0[k], s1[k])


def test_pinned_tile_shapes_round_trip_and_validation():
    pb = config.make_bp1_problem(257, basin=True, levels=5,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    with lib.Engine(pb, me, **FAST) as eng:
        eng.set_tiles(True, 64, 16)
        shapes = [((32, 8, 8), (64, 4, 4)), ((96, 16, 16), (32, 2, 8)),
                  ((64, 8, 16), (64, 8, 16))]
        eng.set_tile_shapes(shapes)
        assert eng.tile_shapes() == shapes
        with pytest.raises(lib.QdeqError):
            eng.set_tile_shapes(shapes[:2])                        # level count
        with pytest.raises(lib.QdeqError):
            eng.set_tile_shapes([((33, 8, 8), (64, 4, 4))] + shapes[1:])  # odd
        with pytest.raises(lib.QdeqError):
            eng.set_tile_shapes([((512, 64, 16), (64, 4, 4))] + shapes[1:])  # LDS
        assert eng.tile_shapes() == shapes
        U = _perturbed(pb, 4)
        V = U.copy()
        eng.ode_solve(V)
    V2, _ = _gpu_solve(pb, me, U, True)
    np.testing.assert_array_equal(V, V2)


def test_grid_fitted_tile_shapes_give_identical_solutions():
    """Tile sizes that split a 2^k + 1 grid into whole tiles (66 = 264/4 at
    257^2, one tile of 66 x 66 per CU) are among the autotuner's candidates;
    with them pinned, the solve is bit-identical to the default shapes'."""
    pb = config.make_bp1_problem(257, basin=True, levels=5,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 6)
    shapes = [((66, 18, 16), (66, 66, 16)), ((34, 10, 8), (44, 22, 16)),
              ((22, 12, 8), (18, 10, 4))]
    with lib.Engine(pb, me, **FAST) as eng:
        eng.set_tiles(True, 64, 16)
        eng.set_tile_shapes(shapes)
        assert eng.tile_shapes() == shapes
        V = U.copy()
        vc = eng.ode_solve(V)[0]
    V2, out2 = _gpu_solve(pb, me, U, True)
    np.testing.assert_array_equal(V, V2)
    assert vc == out2[0]


def test_level0_residual_in_registers_and_in_scratch_agree():
    """Level-0 k_down keeps the residuals of its window in registers (up to 8
    per thread) or, for a larger window per thread, in a scratch array:
    96x24 tiles on 4 thread rows (98 x 26 > 8 x 256 points) take the scratch
    path, 96x24 on 16 rows the register path; both solves are bit-identical
    to the default shapes'."""
    pb = config.make_bp1_problem(257, basin=True, levels=5,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 8)
    ref, out_ref = _gpu_solve(pb, me, U, True)
    rest = [((64, 8, 16), (64, 8, 16)), ((32, 4, 8), (32, 4, 8))]
    for rows in (4, 16):
        shapes = [((96, 24, rows), (64, 16, 16))] + rest
        with lib.Engine(pb, me, **FAST) as eng:
            eng.set_tiles(True, 64, 16)
            eng.set_tile_shapes(shapes)
            V = U.copy()
            vc = eng.ode_solve(V)[0]
        np.testing.assert_array_equal(V, ref, err_msg=str(rows))
        assert vc == out_ref[0]


def test_kernel_timing_does_not_change_results():
    """Timed V-cycles are launched directly instead of from the graph: the
    run must be bit-identical, and the timed kernel must have been timed."""
    pb = config.make_bp1_problem(257, basin=True, levels=5,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    out = []
    for which in (lib.TIME_OFF, lib.TIME_UP0, lib.TIME_TAIL):
        with lib.Engine(pb, me, **FAST) as eng:
            eng.set_tiles(True, 64, 16)
            eng.rk43_setup()
            eng.set_kernel_timing(which)
            info = eng.rk43_step(30)
            ms, calls = eng.kernel_timing()
            if which != lib.TIME_OFF:
                assert calls > 0 and ms > 0
            out.append((info, eng.state()))
    for info, st in out[1:]:
        assert info["t"] == out[0][0]["t"] and info["vcycles"] == out[0][0]["vcycles"]
        for k in ("slip", "psi", "U"):
            np.testing.assert_array_equal(st[k], out[0][1][k])


@pytest.mark.parametrize("precision", [config.PREC_FP64, config.PREC_FP32_IR])
def test_graph_and_direct_launches_are_bit_identical(precision, monkeypatch):
    """One simulation launches its V-cycles kernel by kernel by default; with
    QDEQ_GRAPH_LAUNCH=1 every V-cycle is a replay of the captured graph (many
    replays per run, with the host's own launches between them).  Both runs
    and the default must agree bit for bit."""
    pb = config.make_bp1_problem(257, basin=True, levels=5,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    out = []
    for env in (None, "QDEQ_GRAPH_LAUNCH", "QDEQ_DIRECT_LAUNCH"):
        monkeypatch.delenv("QDEQ_GRAPH_LAUNCH", raising=False)
        monkeypatch.delenv("QDEQ_DIRECT_LAUNCH", raising=False)
        if env:
            monkeypatch.setenv(env, "1")
        with lib.Engine(pb, me, **{**FAST, "precision": precision}) as eng:
            eng.set_tiles(True, 64, 16)
            eng.rk43_setup()
            info = eng.rk43_step(40)
            out.append((info, eng.state()))
    assert out[0][0]["vcycles"] > 40
    for info, st in out[1:]:
        assert info["t"] == out[0][0]["t"] and info["vcycles"] == out[0][0]["vcycles"]
        for k in ("slip", "psi", "U"):
            np.testing.assert_array_equal(st[k], out[0][1][k])


def test_ir_and_fp64_at_2049():
    """The mixed-precision workload's grid: 2049^2, 8 levels.  FP32 V-cycles
    in FP64 refinement and the FP64 cycle (bit-checked against the oracle at
    smaller sizes) both reach the 1e-10 relative residual on a perturbed
    state, with autotuned tiles, and agree to what that residual implies."""
    pb = config.make_bp1_problem(2049, basin=True, levels=8,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    U = _perturbed(pb, 11, scale=0.05)
    out = {}
    for prec, tile in ((config.PREC_FP64, (64, 16)), (config.PREC_FP32_IR, (64, 32))):
        with lib.Engine(pb, **{**FAST, "precision": prec, "max_vcycles": 60}) as eng:
            eng.set_tiles(True, *tile)
            eng.autotune_tiles(2)
            V = U.copy()
            vc, r0, r = eng.ode_solve(V)
            assert r <= 1e-10 * r0 and 1 <= vc <= 60
            out[prec] = V
    scale = np.max(np.abs(out[config.PREC_FP64]))
    assert np.max(np.abs(out[config.PREC_FP32_IR] - out[config.PREC_FP64])) <= 1e-6 * scale


# the bench's production cycle: V(1,3), over-relaxed red/black sweeps
PROD = dict(cycle=config.CYCLE_FAST, n_pre=1, n_post=3, n_coarse=0, relax=1.15)


@pytest.mark.parametrize("n,levels", [(129, 4), (257, 5)])
def test_production_cycle_equals_oracle(n, levels):
    """V(1,3) with relaxation 1.15 on the tiled kernels (several tile
    shapes) equals the oracle's fast cycle with the same shape and factor."""
    pb = config.make_bp1_problem(n, basin=True, levels=levels)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 5)
    p = pb.params
    mg = oracle.MG(p.ni, p.nk, p.dh, p.n_levels, me, relax=PROD["relax"])
    mg.set_cycle(PROD["n_pre"], PROD["n_post"], PROD["n_coarse"], True)
    ref = U.copy()
    st, vc, _, _ = mg.solve(ref, oracle.ORC_TOL_ABS, 1e-300, 3)
    assert vc == 3
    for tile in ((64, 16), (32, 8), (128, 4)):
        with lib.Engine(pb, me, tol=1e-300, max_vcycles=3, **PROD) as eng:
            eng.set_tiles(True, *tile)
            V = U.copy()
            try:
                eng.ode_solve(V)
            except lib.QdeqError as e:
                assert e.status == lib.QDEQ_ERR_MG_NOT_CONVERGED
        np.testing.assert_array_equal(V, ref, err_msg=str(tile))


def test_rk43_production_cycle_matches_oracle():
    pb = config.make_bp1_problem(65, basin=True, levels=3, tol_mode=config.TOL_REL,
                                 tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    oracle.set_libm(False)
    sim = oracle.Sim(pb, me, cycle=1, n_pre=PROD["n_pre"], n_post=PROD["n_post"],
                     n_coarse=PROD["n_coarse"], relax=PROD["relax"])
    sim.init()
    with lib.Engine(pb, me, **PROD) as eng:
        assert eng.uses_tiled_cycle
        eng.rk43_setup()
        for s in range(40):
            sim.step()
            info = eng.rk43_step(1)
            ref = sim.scalars()
            for key in ("t", "dt", "totErr", "V_max", "attempts", "vcycles"):
                assert info[key] == ref[key], (s, key, info[key], ref[key])
        st = eng.state()
    ar = sim.arrays()
    for key in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(st[key], ar[key], err_msg=key)


@pytest.mark.parametrize("n,levels", [(129, 4), (257, 5)])
def test_ir_device_start_equals_host_start(n, levels, monkeypatch):
    """FP32-IR solves decide their start on the device (the host path's
    residual and norm kernels, the sums in its order, its test) and launch
    their first V-cycles without waiting; QDEQ_HOST_START=1 keeps the host
    decision.  The runs are bit for bit equal, V-cycle counts included."""
    pb = config.make_bp1_problem(n, basin=True, levels=levels,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    runs = []
    for host in (False, True):
        if host:
            monkeypatch.setenv("QDEQ_HOST_START", "1")
        else:
            monkeypatch.delenv("QDEQ_HOST_START", raising=False)
        with lib.Engine(pb, me, **{**PROD, "precision": config.PREC_FP32_IR}) as eng:
            eng.set_tiles(True, 64, 32)
            eng.rk43_setup()
            eng.rk43_step(60)
            runs.append(eng.state())
    a, b = runs
    for key in ("t", "vcycles", "solves", "ir_sweeps"):
        assert a["info"][key] == b["info"][key], key
    for k in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.parametrize("n,levels", [(257, 5), (65, 3)])
@pytest.mark.parametrize("timed", [False, True])
def test_stages_launched_ahead_equal_waiting_runs(n, levels, timed, monkeypatch):
    """One simulation queues each stage's fault update, start and first
    V-cycles behind the previous stage's solve; they run only if that solve
    converged with the V-cycles already queued, else they abandon themselves
    and the solve continues.  The runs are bit for bit those of
    QDEQ_SPECULATE=0 (every stage waits), V-cycle counts included, with the
    folded start (257^2) and the separate start kernels (65^2), with and
    without kernel timing; stages are taken over and some abandoned."""
    pb = config.make_bp1_problem(n, basin=True, levels=levels,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    runs = []
    for spec in (True, False):
        if spec:
            monkeypatch.delenv("QDEQ_SPECULATE", raising=False)
        else:
            monkeypatch.setenv("QDEQ_SPECULATE", "0")
        with lib.Engine(pb, me, **PROD) as eng:
            eng.set_tiles(True, 64, 16)
            eng.rk43_setup()
            if timed:
                eng.set_kernel_timing(lib.TIME_UP0)
            eng.rk43_step(120)
            runs.append((eng.state(), eng.launch_ahead_stats()))
    (a, sa), (b, sb) = runs
    for key in ("t", "dt", "vcycles", "solves", "attempts"):
        assert a["info"][key] == b["info"][key], key
    for k in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(a[k], b[k])
    assert sb == {"taken": 0, "abandoned": 0, "next_step": 0}
    assert sa["taken"] > 0.5 * a["info"]["solves"]
    assert sa["abandoned"] > 0
    # the next step's stage 2 behind most accepted steps' final stage 1
    assert sa["next_step"] > 0.3 * 120


def test_next_step_launched_ahead_ends_at_tmax_and_stays_off_between_records(
        monkeypatch):
    """Stage 2 of the next step is queued behind an accepted step's final
    stage 1 only when that step runs within the same call and nothing is
    recorded between the steps.  Bit for bit against QDEQ_SPECULATE=0: a run
    whose last step is clipped to tmax inside one call (the queued stage used
    the clipped dt), a run one step per call and a run with events on (both
    without any stage queued across a step)."""
    n, levels = 129, 4
    pb = config.make_bp1_problem(n, basin=True, levels=levels,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    monkeypatch.setenv("QDEQ_SPECULATE", "0")
    with lib.Engine(pb, me, **PROD) as eng:
        eng.set_tiles(True, 64, 16)
        eng.rk43_setup()
        t30 = eng.rk43_step(30)["t"]
        dt30 = eng.rk43_step(0)["dt"]
    tmax = t30 + 0.37 * dt30  # step 31 clipped to end at tmax

    def run(spec, how):
        if spec:
            monkeypatch.delenv("QDEQ_SPECULATE", raising=False)
        else:
            monkeypatch.setenv("QDEQ_SPECULATE", "0")
        kw = dict(PROD, tmax=tmax) if how == "tmax" else dict(PROD)
        with lib.Engine(pb, me, **kw) as eng:
            eng.set_tiles(True, 64, 16)
            eng.rk43_setup()
            if how == "tmax":
                eng.rk43_step(50)
            elif how == "single":
                for _ in range(25):
                    eng.rk43_step(1)
            else:
                eng.events_enable(1e-3)
                eng.rk43_step(25)
            return eng.state(), eng.launch_ahead_stats()

    for how in ("tmax", "single", "events"):
        (a, sa), (b, sb) = run(True, how), run(False, how)
        for key in ("t", "dt", "vcycles", "solves", "attempts"):
            assert a["info"][key] == b["info"][key], (how, key)
        for k in ("slip", "psi", "V", "U"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=how)
        assert sb["taken"] == 0
        if how == "tmax":
            assert a["info"]["t"] >= tmax
            assert sa["next_step"] > 10
        else:
            assert sa["next_step"] == 0 and sa["taken"] > 0


SHAPES_1025 = [((130, 34, 16), (66, 66, 16)), ((34, 34, 16), (34, 34, 16)),
               ((32, 10, 8), (34, 10, 8)), ((12, 12, 8), (12, 8, 8)),
               ((10, 6, 8), (18, 2, 8))]


@pytest.mark.parametrize("n,levels,steps", [(1025, 7, 8), (257, 5, 60)])
def test_stop_tests_folded_equal_separate_runs(n, levels, steps, monkeypatch):
    """Within a batch launched kernel by kernel, the stop test of every
    V-cycle but the batch's last runs in the next V-cycle's level-0 down
    kernel (every workgroup sums the norm partials in the stop-test kernel's
    order).  Bit for bit the runs of QDEQ_FOLD_STOP=0, which launch a
    stop-test kernel after every V-cycle, V-cycle counts included; tests were
    folded (at 1025^2 with the 66x66 level-0 tiles: 256 partials)."""
    pb = config.make_bp1_problem(n, basin=True, levels=levels,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = lib.build_mu_equi(pb.mu)
    runs = []
    for fold in (True, False):
        if fold:
            monkeypatch.delenv("QDEQ_FOLD_STOP", raising=False)
        else:
            monkeypatch.setenv("QDEQ_FOLD_STOP", "0")
        with lib.Engine(pb, me, **PROD) as eng:
            eng.set_tiles(True, 64, 16)
            if n == 1025:
                eng.set_tile_shapes(SHAPES_1025)
            eng.rk43_setup()
            eng.rk43_step(steps)
            runs.append((eng.state(), eng.folded_stop_tests()))
    (a, fa), (b, fb) = runs
    for key in ("t", "dt", "vcycles", "solves", "attempts"):
        assert a["info"][key] == b["info"][key], key
    for k in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(a[k], b[k])
    assert fb == 0
    # batches of two or more V-cycles (few in the early steps, where most
    # solves need one)
    assert fa > 0


@pytest.mark.parametrize("cap", [2, 3, 5])
def test_folded_stop_tests_stop_at_the_cap(cap, monkeypatch):
    """A solve that cannot meet its tolerance stops at the V-cycle cap with
    the same result whether its stop tests are folded or not: no test that
    could reach the cap is folded."""
    pb = config.make_bp1_problem(129, basin=True, levels=4)
    me = oracle.mu_equi(pb.mu)
    U = _perturbed(pb, 11)
    out = []
    for fold in (True, False):
        if fold:
            monkeypatch.delenv("QDEQ_FOLD_STOP", raising=False)
        else:
            monkeypatch.setenv("QDEQ_FOLD_STOP", "0")
        with lib.Engine(pb, me, **{**PROD, "tol": 1e-300, "max_vcycles": cap}) as eng:
            eng.set_tiles(True, 64, 16)
            V = U.copy()
            with pytest.raises(lib.QdeqError) as ei:
                eng.ode_solve(V)
            assert ei.value.status == lib.QDEQ_ERR_MG_NOT_CONVERGED
            out.append((V, eng.folded_stop_tests()))
    (va, fa), (vb, fb) = out
    np.testing.assert_array_equal(va, vb)
    assert fb == 0
    assert fa >= 1


def test_accepted_case_launched_ahead_equal_waiting_runs_with_rejections(monkeypatch):
    """Behind each attempt's combination the commit, the final stage 1 and
    that solve's start and first batch are queued, gated on the device's own
    accept test; a rejected attempt finds them skipped.  Started with a first
    step far too long for the error tolerance (a series of rejected attempts
    until the controller has shrunk it), the run is bit for bit that of
    QDEQ_SPECULATE=0, V-cycle counts included."""
    pb = config.make_bp1_problem(129, basin=True, levels=4,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    steps = 80
    runs = []
    for spec in (True, False):
        if spec:
            monkeypatch.delenv("QDEQ_SPECULATE", raising=False)
        else:
            monkeypatch.setenv("QDEQ_SPECULATE", "0")
        with lib.Engine(pb, me, **{**PROD, "ini_dt": 1e7}) as eng:
            eng.set_tiles(True, 64, 16)
            eng.rk43_setup()
            eng.rk43_step(steps)
            runs.append((eng.state(), eng.launch_ahead_stats()))
    (a, sa), (b, sb) = runs
    for key in ("t", "dt", "vcycles", "solves", "attempts"):
        assert a["info"][key] == b["info"][key], key
    for k in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(a[k], b[k])
    attempts = (a["info"]["solves"] - steps - 1) / 5
    assert attempts >= steps + 3         # rejected attempts happened
    assert sb["taken"] == 0
    assert sa["abandoned"] > 0 and sa["taken"] > 0
