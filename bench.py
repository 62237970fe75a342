#!/usr/bin/env python3
"""Accepted RK43 steps per second at 1025^2 to a 1e-10 relative elliptic
residual, on one MI355X.

Workload (basin_1025, qdeqcycle_fdm_amd/params/basin_1025.json): 48 km x
48 km antiplane fault model on a 1025 x 1025 grid, low-rigidity basin (20 GPa
in the top 2 km within 10 km of the fault, 500 m tanh edges, 32 GPa
elsewhere), rate-and-state friction with radiation damping, 7-level
multigrid, 1e-10 relative residual per stage solve.  The run warms up W steps
from t = 0 and times the next K accepted steps (`value` = K / seconds, inputs
in HBM).  Inputs are synthetic (generated, no data files).

Modes: fast64 (the headline: FP64 V(1,3) fast cycle on the tiled LDS
kernels, direct coarsest solve), ir32 (FP32 correction cycles in FP64
iterative refinement) and ref64 (the reference's own V-cycle: 32 sweeps per
level, FP64).  One JSON line goes to stdout; progress goes to stderr.

Side measurements, each of a fixed size that does not depend on K:
* comparison window: the same warm-up, then --window-steps plain steps,
  which end in the quiet phase after the first event (the headline run
  itself when K reaches the window, else a separate run);
* learned vs plain: pairs are harvested from --harvest-runs runs of
  --harvest-steps steps with seeded Dc / sigma_n perturbations, the network
  is trained on the GPU (both untimed, their times reported), and the run
  with the learned initial guess goes from t = 0 to the window's end;
  V-cycles per solve, steps per second and the slip difference at the
  window's end are reported for both;
* projection start over the same window, and the plain start at 1e-12;
* ref64 over the headline's first min(K, --ref-steps) timed steps, against
  fast64's rate over the same steps (speedup_vs_ref64), and ir32;
* an ensemble leg: --ensemble-leg members of the seeded 64-member draw,
  --ensemble-leg-steps steps each, alone and stepped together;
* the CPU baseline: the project's FP64 port of the solver (oracle/) over the
  workload's first 20 accepted steps from t = 0 (fewer if one host core
  passes twice --cpu-seconds), on one core, then the same steps on every
  core; the GPU then runs the same steps from t = 0 and the three final
  states are compared bit for bit;
* with --c1, config 1 in full as well (129^2, the reference's own cycle,
  2000 steps): GPU, one core and every core, bit for bit.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("RK43 steps/s at 1025² to 1e-10 residual; V-cycles/solve, "
          "learned vs plain")

# The fast cycle: one pre- and three post-smoothing red/black sweeps,
# over-relaxed (1.15), and the direct 17^2 coarsest solve.  Swept over
# V(1,1)..V(4,4) and relaxations 1.0..1.4 at 257^2..2049^2; this shape and
# factor were best at every grid size in FP64 and FP32-IR.
MODES = {
    "fast64": dict(cycle=1, precision=0, n_pre=1, n_post=3, n_coarse=0,
                   relax=1.15, tiles=(64, 16)),
    "ir32": dict(cycle=1, precision=1, n_pre=1, n_post=3, n_coarse=0,
                 relax=1.15, tiles=(64, 32)),
    "ref64": dict(cycle=0, precision=0, n_pre=32, n_post=32, n_coarse=32,
                  relax=1.0, tiles=None),
}

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sync():
    """Device-wide barrier through the engine's HIP runtime.  torch is not
    imported here: its bundled HIP runtime shares the SONAME of /opt/rocm's,
    would be the one the library binds to, and does not match rocprofv3."""
    from qdeqcycle_fdm_amd import lib
    lib.device_synchronize(0)


TUNED = {}  # mode -> per-level tile shapes found by the autotuner


def make_engine(pb, me, mode, **extra):
    """Engine of a mode; tiled modes use the autotuned per-level tile shapes
    (tuned once per mode, on the first engine, before any step)."""
    from qdeqcycle_fdm_amd import lib
    kw = dict(MODES[mode])
    tiles = kw.pop("tiles")
    kw.update(extra)
    eng = lib.Engine(pb, me, **kw)
    if tiles:
        eng.set_tiles(True, *tiles)
        if AUTOTUNE:
            t0 = time.perf_counter()
            shapes = eng.autotune_tiles(5)
            if mode not in TUNED:
                TUNED[mode] = dict(shapes=shapes,
                                   seconds=round(time.perf_counter() - t0, 3))
    return eng


AUTOTUNE = True


def kernel_bytes(mode, n0, nu):
    """Algorithmic bytes of one launch of the level-0 up kernel (bilinear
    prolongation + correction, nu red/black sweeps, and in FP64 the residual
    norm): per point a prolongation + correction moves 2.25 words, a sweep 5
    (u read+write, f, two face coefficients), a residual 5."""
    s = 4 if mode == "ir32" else 8
    per = 2.25 + nu * 5 + (5 if mode == "fast64" else 0)
    return per * s * n0


def run_mode(pb, me, mode, warmup, steps, timed_kernel=None, marks=(), keep_state_at=None):
    """Warm-up, then `steps` timed accepted steps.  `marks`: step counts
    inside the timed region at which the cumulative rate is recorded (the
    timed region is split there, with a device barrier between segments);
    with keep_state_at (one of the marks) the state there is copied out
    between segments (the copy is not timed)."""
    from qdeqcycle_fdm_amd import lib
    eng = make_engine(pb, me, mode)
    eng.rk43_setup()
    if warmup:
        eng.rk43_step(warmup)
    sync()
    if timed_kernel is not None:
        eng.set_kernel_timing(timed_kernel)
    i0 = eng.state()["info"]
    cuts = sorted({m for m in marks if 0 < m < steps} | {steps})
    seconds, done, at, kept = 0.0, 0, {}, None
    for cut in cuts:
        t0 = time.perf_counter()
        info = eng.rk43_step(cut - done)
        sync()
        seconds += time.perf_counter() - t0
        done = cut
        at[cut] = dict(steps=cut, seconds=seconds, steps_per_s=cut / seconds,
                       t_start=i0["t"], t_end=info["t"],
                       vcycles_per_solve=(info["vcycles"] - i0["vcycles"]) /
                       max(info["solves"] - i0["solves"], 1),
                       solves_per_step=(info["solves"] - i0["solves"]) / cut,
                       sim_s_per_s=(info["t"] - i0["t"]) / seconds)
        if cut == keep_state_at and cut < steps:
            kept = eng.state()
    out = dict(seconds=seconds, steps=steps, steps_per_s=steps / seconds,
               vcycles=info["vcycles"] - i0["vcycles"],
               solves=info["solves"] - i0["solves"], t_start=i0["t"],
               t_end=info["t"], V_max_end=info["V_max"], marks=at)
    out["vcycles_per_solve"] = out["vcycles"] / max(out["solves"], 1)
    out["solves_per_step"] = out["solves"] / steps
    if keep_state_at == steps:
        kept = eng.state()
    out["kept_state"] = kept
    if timed_kernel is not None:
        ms, calls = eng.kernel_timing()
        out["kernel_ms_total"], out["kernel_calls"] = ms, calls
        eng.set_kernel_timing(lib.TIME_OFF)
    return eng, out


def train_guess(me, steps, runs, seed, epochs):
    """Learned initial guess: harvests (boundary change, solved displacement
    change) pairs from `runs` runs of the workload whose Dc and sigma_n are
    perturbed with a seeded generator (the timed run is not among them), one
    pair every second solve, then trains the network on the GPU."""
    import numpy as np
    from qdeqcycle_fdm_amd import config, lib
    rng = np.random.default_rng(seed)
    g = lib.Guess(1025, capacity=20000, stride=2)
    t0 = time.perf_counter()
    params = []
    for _ in range(runs):
        Dc = 0.008 * float(np.exp(rng.uniform(-0.35, 0.35)))
        sn = 50.0 * (1.0 + float(rng.uniform(-0.15, 0.15)))
        pb = config.make_bp1_problem(1025, basin=True, tol_mode=config.TOL_REL,
                                     tol=1e-10, Dc=Dc, sigma_n=sn)
        eng = make_engine(pb, me, "fast64")
        eng.rk43_setup()
        eng.attach_guess(g, lib.GUESS_HARVEST)
        eng.rk43_step(steps)
        eng.close()
        params.append({"Dc": round(Dc, 6), "sigma_n": round(sn, 3)})
        log(f"harvest run Dc={Dc:.4f} sigma_n={sn:.2f}: {g.pairs} pairs")
    sync()
    harvest_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    loss = g.train(epochs=epochs, batch=256)
    sync()
    train_s = time.perf_counter() - t1
    log(f"trained on {g.pairs} pairs in {train_s:.1f} s, loss {loss:.3e}")
    return g, dict(pairs=g.pairs, runs=params, harvest_s=round(harvest_s, 2),
                   train_s=round(train_s, 2), epochs=epochs, loss=loss, seed=seed)


def run_learned(pb, me, g, warmup, tmax):
    """The fast FP64 run with the learned guess from t = 0: warm-up, then
    timed until t = tmax (the end of the plain run's comparison window; the
    last step is shortened to end there)."""
    from qdeqcycle_fdm_amd import lib
    eng = make_engine(pb, me, "fast64", tmax=tmax)
    eng.rk43_setup()
    eng.attach_guess(g, lib.GUESS_PREDICT)
    if warmup:
        eng.rk43_step(warmup)
    sync()
    i0 = eng.state()["info"]
    s0 = eng.guess_stats()
    t0 = time.perf_counter()
    info = eng.rk43_step(10 ** 7)
    sync()
    t1 = time.perf_counter()
    s1 = eng.guess_stats()
    steps = info["it"] - i0["it"]
    out = dict(seconds=t1 - t0, steps=steps, steps_per_s=steps / (t1 - t0),
               vcycles=info["vcycles"] - i0["vcycles"],
               solves=info["solves"] - i0["solves"], t_start=i0["t"],
               t_end=info["t"], sim_s_per_s=(info["t"] - i0["t"]) / (t1 - t0),
               predicted=s1["predicted"] - s0["predicted"],
               accepted=s1["accepted"] - s0["accepted"])
    out["vcycles_per_solve"] = out["vcycles"] / max(out["solves"], 1)
    out["solves_per_step"] = out["solves"] / max(steps, 1)
    return eng, out


def run_projection(me, warmup, tmax, slots, tol):
    """The fast FP64 run from t = 0 with the projection start (`slots`
    stored solutions, 0 = the plain start) at relative residual `tol`:
    warm-up, then timed until t = tmax (the end of the plain run's
    comparison window).  Returns the rates and the final fault state."""
    from qdeqcycle_fdm_amd import config
    pb = config.make_bp1_problem(1025, basin=True, tol_mode=config.TOL_REL, tol=tol)
    eng = make_engine(pb, me, "fast64", tmax=tmax, max_vcycles=200)
    if slots:
        eng.projection_enable(slots)
    eng.rk43_setup()
    if warmup:
        eng.rk43_step(warmup)
    sync()
    i0 = eng.state()["info"]
    s0 = eng.projection_stats()
    t0 = time.perf_counter()
    info = eng.rk43_step(10 ** 7)
    sync()
    t1 = time.perf_counter()
    s1 = eng.projection_stats()
    st = eng.state()
    eng.close()
    steps = info["it"] - i0["it"]
    solves = info["solves"] - i0["solves"]
    out = dict(slots=slots, tol=tol, steps=steps, seconds=round(t1 - t0, 3),
               steps_per_s=steps / (t1 - t0),
               vcycles_per_solve=(info["vcycles"] - i0["vcycles"]) / max(solves, 1),
               solves_per_step=solves / max(steps, 1), t_start=i0["t"], t_end=info["t"])
    if slots:
        out["offered"] = s1["offered"] - s0["offered"]
        out["used"] = s1["used"] - s0["used"]
    return out, st


def fault_diff(a, b):
    """Largest relative slip difference and largest |log10 V| difference on
    the fault nodes between two states."""
    import numpy as np
    sa, sb = a["slip"][1:-1], b["slip"][1:-1]
    va = np.log10(np.maximum(a["V"][1:-1], 1e-300))
    vb = np.log10(np.maximum(b["V"][1:-1], 1e-300))
    return (float(np.max(np.abs(sa - sb)) / np.max(np.abs(sb))),
            float(np.max(np.abs(va - vb))))


def run_ensemble(members, steps, precision_mode, proj_slots=0):
    """Ensemble workload: `members` simulations (seeded draws of grid
    257/513/1025, Dc, sigma_n, b; basin medium), `steps` accepted steps each,
    first one after the other (each alone), then all together with
    qdeq_ensemble_step, and with proj_slots > 0 all together again with the
    projection start.  Returns the rates, V-cycles per solve and whether
    every member's final state is bit-identical between the first two runs."""
    import numpy as np
    from qdeqcycle_fdm_amd import config, lib
    levels = {257: 5, 513: 6, 1025: 7}
    draws = config.ensemble_members(members, seed=0)
    media, shapes, pbs = {}, {}, []
    for d in draws:
        pbs.append(config.make_bp1_problem(d["n"], basin=True, levels=levels[d["n"]],
                                           Dc=d["Dc"], sigma_n=d["sigma_n"],
                                           b_seis=d["b_seis"], tol_mode=config.TOL_REL,
                                           tol=1e-10))
        if d["n"] not in media:
            media[d["n"]] = lib.build_mu_equi(pbs[-1].mu)

    def make(i, slots=0):
        pb = pbs[i]
        n = draws[i]["n"]
        kw = dict(MODES[precision_mode])
        tiles = kw.pop("tiles")
        eng = lib.Engine(pb, media[n], max_vcycles=200, **kw)
        eng.set_tiles(True, *tiles)
        if n not in shapes:
            shapes[n] = eng.autotune_tiles(5)
        else:
            eng.set_tile_shapes(shapes[n])
        if slots:
            eng.projection_enable(slots)
        eng.rk43_setup()
        return eng

    def vcycles_per_solve(engines):
        infos = [e.state()["info"] for e in engines]
        return sum(i["vcycles"] for i in infos) / max(sum(i["solves"] for i in infos), 1)

    alone, t_alone = [], 0.0
    for i in range(members):
        eng = make(i)
        sync()
        t0 = time.perf_counter()
        eng.rk43_step(steps)
        sync()
        t_alone += time.perf_counter() - t0
        alone.append(eng.state()["slip"])
        eng.close()
        if (i + 1) % 8 == 0:
            log(f"ensemble members alone: {i + 1}/{members} ({t_alone:.1f} s)")
    engines = [make(i) for i in range(members)]
    sync()
    t0 = time.perf_counter()
    lib.ensemble_step(engines, steps)
    sync()
    t_ens = time.perf_counter() - t0
    same = all(np.array_equal(e.state()["slip"], a) for e, a in zip(engines, alone))
    vc_plain = vcycles_per_solve(engines)
    for e in engines:
        e.close()
    total = members * steps
    out = dict(members=members, steps_each=steps, mode=precision_mode,
               grids={str(n): sum(1 for d in draws if d["n"] == n) for n in levels},
               steps_per_s_alone=total / t_alone, steps_per_s_ensemble=total / t_ens,
               seconds_alone=t_alone, seconds_ensemble=t_ens,
               vcycles_per_solve=vc_plain, bitwise_equal_to_alone=bool(same))
    if proj_slots:
        engines = [make(i, proj_slots) for i in range(members)]
        sync()
        t0 = time.perf_counter()
        lib.ensemble_step(engines, steps)
        sync()
        t_proj = time.perf_counter() - t0
        out["projection"] = dict(slots=proj_slots, steps_per_s=total / t_proj,
                                 seconds=t_proj,
                                 vcycles_per_solve=vcycles_per_solve(engines),
                                 speedup=t_ens / t_proj)
        for e in engines:
            e.close()
        log(f"ensemble with the projection start ({proj_slots} slots): "
            f"{total / t_proj:.1f} steps/s ({t_ens / t_proj:.2f}x)")
    return out


def run_mixed_2049(n_events=2, tail_steps=200):
    """Mixed-precision workload: the basin model on 2049^2 (8 levels) with
    FP32 V-cycles in FP64 refinement, against the FP64 cycle on the same
    grid (the precision error) and the 1025^2 FP64 run (the resolution
    difference).  Every run starts at t = 0.  The 1025^2 FP64 run, the
    2049^2 FP32-IR run and the 2049^2 FP64 run with the projection start
    each step until n_events events (max V > 1e-3 m/s) have ended and then
    tail_steps more, into the quiet phase; the plain 2049^2 FP64 run stops
    exactly at the IR run's final time T.  Reported: rates, the slip and
    log10 V differences between the IR and plain FP64 2049^2 runs on the
    fault at T, and event onsets and recurrence intervals of all runs (the
    two grids resolve the nucleation differently, and the projection run's
    last event may fall on either side of T, so those are compared through
    the event catalogue, not through states at one time)."""
    import numpy as np
    from qdeqcycle_fdm_amd import config, lib

    def run(n, mode, tmax=None, proj=0):
        levels = {1025: 7, 2049: 8}[n]
        pb = config.make_bp1_problem(n, basin=True, levels=levels,
                                     tol_mode=config.TOL_REL, tol=1e-10)
        me = lib.build_mu_equi(pb.mu)
        extra = {} if tmax is None else {"tmax": tmax}
        eng = make_engine(pb, me, mode, max_vcycles=200, **extra)
        if proj:
            eng.projection_enable(proj)
        eng.rk43_setup()
        eng.events_enable(1e-3)
        sync()
        t0 = time.perf_counter()
        if tmax is not None:
            info = eng.rk43_step(10 ** 7)
        else:
            # until n_events events have ended, then tail_steps more
            while True:
                info = eng.rk43_step(100)
                done = [e for e in eng.events() if e["t_end"] >= 0]
                if len(done) >= n_events or info["it"] > 20000:
                    break
                if info["it"] % 1000 == 0:
                    log(f"{n}^2 {mode}: step {info['it']} t={info['t']:.6g} "
                        f"events ended {len(done)}")
            info = eng.rk43_step(tail_steps)
        sync()
        sec = time.perf_counter() - t0
        st = eng.state()
        ev = eng.events()
        eng.close()
        tag = f"{mode} + projection ({proj} slots)" if proj else mode
        log(f"{n}^2 {tag}: {info['it']} steps to t={info['t']:.6g} in {sec:.1f} s")
        return dict(info=info, seconds=sec, slip=st["slip"][1:-1],
                    V=st["V"][1:-1], events=ev)

    ref = run(1025, "fast64")
    ir = run(2049, "ir32")
    T = ir["info"]["t"]
    f64 = run(2049, "fast64", tmax=T)
    f64p = run(2049, "fast64", proj=6)

    def catalogue(a, b):
        ea = [e["t_onset"] for e in a["events"]]
        eb = [e["t_onset"] for e in b["events"]]
        k = min(len(ea), len(eb))
        return dict(onsets=[ea, eb],
                    onset_rel_diff=[abs(ea[i] - eb[i]) / eb[i] for i in range(k)],
                    recurrence=[[ea[i + 1] - ea[i] for i in range(len(ea) - 1)],
                                [eb[i + 1] - eb[i] for i in range(len(eb) - 1)]],
                    slip_per_event=[[e["slip_max"] for e in a["events"]],
                                    [e["slip_max"] for e in b["events"]]])

    def cmp(a, b):
        va = np.log10(np.maximum(a["V"], 1e-300))
        vb = np.log10(np.maximum(b["V"], 1e-300))
        out = dict(slip_max_rel=float(np.max(np.abs(a["slip"] - b["slip"])) /
                                      np.max(np.abs(b["slip"]))),
                   log10V_max_abs=float(np.max(np.abs(va - vb))))
        out.update(catalogue(a, b))
        return out

    def rate(r):
        return dict(steps=r["info"]["it"], seconds=round(r["seconds"], 2),
                    steps_per_s=r["info"]["it"] / r["seconds"],
                    vcycles_per_solve=r["info"]["vcycles"] / max(r["info"]["solves"], 1))

    return dict(T=T, runs={"1025_fp64": rate(ref), "2049_ir32": rate(ir),
                           "2049_fp64": rate(f64),
                           "2049_fp64_projection_6_slots": rate(f64p)},
                ir32_vs_fp64_2049=cmp(ir, f64),
                fp64_projection_vs_fp64_2049=catalogue(f64p, f64),
                ir32_2049_vs_fp64_1025=catalogue(ir, ref))


def cpu_cores() -> int:
    """Host cores this process may use (the affinity mask, capped by
    OMP_NUM_THREADS when that is set)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(n, int(env)) if env.isdigit() and int(env) > 0 else n


def _attempts_per_step(solves, steps):
    """RK43 attempts per accepted step from the solve count: one solve at
    t = 0, five per attempt (stages 2..6) and one per accepted step (its
    final stage 1)."""
    return (solves - steps - 1) / 5.0 / steps if steps else None


def cpu_baseline(pb, me, eng, seconds, steps=20):
    """The oracle from t = 0 over the workload's first `steps` accepted steps
    (fewer if one thread passes 2 x `seconds`): on one thread, then the same
    steps on every core (OpenMP over the multigrid's point loops); the
    engine, set back to t = 0, then runs the same steps.  Returns both CPU
    rates, the GPU rate on the same steps, V-cycles per solve and attempts
    per step, and whether the three final states agree bit for bit."""
    import numpy as np
    import oracle
    p = MODES["fast64"]
    oracle.set_libm(False)

    def fresh():
        sim = oracle.Sim(pb, me, cycle=1, n_pre=p["n_pre"], n_post=p["n_post"],
                         n_coarse=p["n_coarse"], relax=p["relax"])
        sim.init()
        return sim

    oracle.set_threads(1)
    sim1 = fresh()
    n = 0
    t0 = time.perf_counter()
    while n < steps:
        sim1.step()
        n += 1
        if time.perf_counter() - t0 >= 2 * seconds:
            break
    cpu1_s = time.perf_counter() - t0
    cores = cpu_cores()
    used = oracle.set_threads(cores)
    simn = fresh()
    t0 = time.perf_counter()
    for _ in range(n):
        simn.step()
    cpun_s = time.perf_counter() - t0
    oracle.set_threads(1)
    eng.rk43_setup()
    sync()
    g0 = time.perf_counter()
    eng.rk43_step(n)
    sync()
    gpu_s = time.perf_counter() - g0
    st2 = eng.state()
    same = True
    for sim in (sim1, simn):
        ar = sim.arrays()
        same = same and all(np.array_equal(st2[k], ar[k]) for k in ("slip", "psi", "V", "U"))
        same = same and st2["info"]["t"] == sim.scalars()["t"]
    sc = sim1.scalars()
    return dict(steps=n, cpu_seconds=cpu1_s, cpu_steps_per_s=n / cpu1_s,
                cpu_all_seconds=cpun_s, cpu_all_steps_per_s=n / cpun_s, threads=used,
                gpu_steps_per_s_same_steps=n / gpu_s, bitwise_equal=bool(same),
                vcycles_per_solve=sc["vcycles"] / sc["solves"],
                attempts_per_step=_attempts_per_step(sc["solves"], n),
                t_end=st2["info"]["t"])


def c1_leg():
    """Config 1 in full (bp1_129: 129^2, L = 4, the reference's own V-cycle
    with 32 sweeps, 2000 accepted steps): the engine, then the oracle on one
    thread and on every core; rates, wall times, V-cycles per solve,
    attempts per step and bit-equality of the three final states."""
    import numpy as np
    import oracle
    from qdeqcycle_fdm_amd import config, lib
    steps = 2000
    pb = config.workload("bp1_129")
    me = lib.build_mu_equi(pb.mu)
    oracle.set_libm(False)
    with lib.Engine(pb, me) as e1:
        e1.rk43_setup()
        sync()
        g0 = time.perf_counter()
        e1.rk43_step(steps)
        sync()
        gpu_s = time.perf_counter() - g0
        gst = e1.state()
    out = {"workload": "bp1_129", "steps": steps, "gpu_seconds": round(gpu_s, 3),
           "gpu_steps_per_s": round(steps / gpu_s, 2)}
    same = True
    for label, threads in (("cpu_1_thread", 1), ("cpu_all_cores", cpu_cores())):
        used = oracle.set_threads(threads)
        sim = oracle.Sim(pb, me)
        sim.init()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.step()
        sec = time.perf_counter() - t0
        sc, ar = sim.scalars(), sim.arrays()
        same = same and all(np.array_equal(gst[k], ar[k]) for k in ("slip", "psi", "V", "U"))
        same = same and gst["info"]["t"] == sc["t"]
        out[label] = {"threads": used, "seconds": round(sec, 2),
                      "steps_per_s": round(steps / sec, 3),
                      "vcycles_per_solve": round(sc["vcycles"] / sc["solves"], 4),
                      "attempts_per_step": round(_attempts_per_step(sc["solves"], steps), 4)}
        log(f"config 1 {label}: {out[label]['steps_per_s']} steps/s "
            f"({out[label]['seconds']} s on {used} threads)")
    oracle.set_threads(1)
    out["gpu_speedup_vs_cpu_1_thread"] = round(out["gpu_steps_per_s"] /
                                               out["cpu_1_thread"]["steps_per_s"], 1)
    out["bitwise_equal"] = bool(same)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mode", default="fast64", choices=sorted(MODES))
    ap.add_argument("--compare-steps", type=int, default=300,
                    help="timed steps of the ir32 mode (ref64 runs a tenth)")
    ap.add_argument("--no-compare", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--c1", action="store_true",
                    help="also run config 1 (129^2, 2000 steps) in full on the GPU "
                         "and the CPU (several minutes of CPU time)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="bound of the one-core CPU baseline: it stops before its "
                         "20 steps once twice this has passed (the all-core run "
                         "repeats the same steps)")
    ap.add_argument("--ensemble", type=int, default=0, metavar="M",
                    help="only run the ensemble workload: M members, "
                         "--ensemble-steps steps each, alone and together")
    ap.add_argument("--ensemble-steps", type=int, default=1000)
    ap.add_argument("--mixed2049", action="store_true",
                    help="only run the 2049^2 mixed-precision workload and its "
                         "accuracy comparison")
    ap.add_argument("--ensemble-mode", default="ir32", choices=("fast64", "ir32"))
    ap.add_argument("--ensemble-proj-slots", type=int, default=6,
                    help="also run the ensemble with the projection start (0: no)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="default tile shapes instead of the autotuned ones")
    ap.add_argument("--no-projection", action="store_true",
                    help="skip the projection-start runs")
    ap.add_argument("--proj-slots", type=int, default=8)
    ap.add_argument("--proj-tol", type=float, default=1e-12)
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-guess harvest, training and run")
    ap.add_argument("--harvest-runs", type=int, default=2)
    ap.add_argument("--window-steps", type=int, default=2780,
                    help="plain steps of the learned-vs-plain window (after the "
                         "warm-up; independent of --steps)")
    ap.add_argument("--harvest-steps", type=int, default=3000,
                    help="steps of each harvest run of the learned guess")
    ap.add_argument("--ref-steps", type=int, default=300,
                    help="at most this many of the headline's timed steps are "
                         "also run with the reference cycle (speedup_vs_ref64)")
    ap.add_argument("--ensemble-leg", type=int, default=8, metavar="M",
                    help="members of the short ensemble leg of the default run "
                         "(0: none)")
    ap.add_argument("--ensemble-leg-steps", type=int, default=50)
    ap.add_argument("--train-epochs", type=int, default=30)
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="run the V-cycle as one captured graph without the "
                         "HIP-event bracket of the level-0 up kernel (used for "
                         "the rocprofv3 passes)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="N",
                    help="only time a plain device copy of N doubles (counter "
                         "calibration) and exit")
    args = ap.parse_args()
    global AUTOTUNE
    AUTOTUNE = not args.no_autotune

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU "
              "path (independent simulations shard without communication, one "
              "process per GPU). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from qdeqcycle_fdm_amd import config, lib
    lib.load()
    if args.mixed2049:
        res = run_mixed_2049()
        r = res["runs"]["2049_ir32"]
        print(json.dumps({"metric": "RK43 steps/s at 2049^2, FP32 V-cycles in FP64 "
                                    "refinement, 1e-10 relative residual",
                          "value": round(r["steps_per_s"], 3), "unit": "steps/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f32+f64",
                          "data": "synthetic", "mixed2049": res}), flush=True)
        return
    if args.ensemble:
        res = run_ensemble(args.ensemble, args.ensemble_steps, args.ensemble_mode,
                           args.ensemble_proj_slots)
        log(f"ensemble: {res['steps_per_s_ensemble']:.1f} steps/s together, "
            f"{res['steps_per_s_alone']:.1f} one after the other, bitwise equal: "
            f"{res['bitwise_equal_to_alone']}")
        print(json.dumps({"metric": "ensemble accepted RK43 steps/s (all members)",
                          "value": round(res["steps_per_s_ensemble"], 2),
                          "unit": "steps/s", "n_gpus": 1, "higher_is_better": True,
                          "data": "synthetic", "ensemble": res}), flush=True)
        return
    pb = config.workload("basin_1025")
    if args.copy_probe:
        with lib.Engine(pb) as eng:
            ms = eng.copy_probe(args.copy_probe, reps=10)
        print(json.dumps({"copy_probe_doubles": args.copy_probe, "avg_ms": ms,
                          "GB/s": 2 * 8 * args.copy_probe / ms / 1e9 * 1e3}))
        return
    t0 = time.perf_counter()
    me = lib.build_mu_equi(pb.mu)
    log(f"mu_equi on the GPU: {time.perf_counter() - t0:.2f} s")

    timed = (lib.TIME_UP0 if args.mode != "ref64" and not args.no_kernel_timing
             else None)
    # the reference cycle is timed over the headline's first ref_steps steps
    # (its rate is 1/50 of the fast cycle's), and the fast cycle's rate over
    # the same steps is recorded for the comparison
    ref_steps = min(args.steps, args.ref_steps)
    window = args.window_steps
    compare = args.mode == "fast64" and not (args.no_learned and args.no_projection)
    marks = [ref_steps] + ([window] if compare and window < args.steps else [])
    eng, main_run = run_mode(pb, me, args.mode, args.warmup, args.steps, timed,
                             marks=marks,
                             keep_state_at=window if compare and window <= args.steps
                             else None)
    log(f"{args.mode}: {main_run['steps_per_s']:.2f} steps/s, "
        f"{main_run['vcycles_per_solve']:.2f} V-cycles/solve, "
        f"t {main_run['t_start']:.4g} -> {main_run['t_end']:.4g} s")

    # The learned-vs-plain comparison window: the same warm-up, then
    # window_steps plain steps, which end in the quiet phase after the first
    # event (the state there is copied out untimed).  It does not depend on
    # --steps: shorter headline runs get a separate plain run of the window.
    win, win_state = None, None
    if compare:
        if window <= args.steps:
            win, win_state = main_run["marks"][window], main_run["kept_state"]
        else:
            e_w, wr = run_mode(pb, me, "fast64", args.warmup, window, None,
                               keep_state_at=window)
            e_w.close()
            win, win_state = wr["marks"][window], wr["kept_state"]
        log(f"comparison window: {win['steps']} plain steps, t {win['t_start']:.4g} -> "
            f"{win['t_end']:.4g} s, {win['steps_per_s']:.2f} steps/s, "
            f"{win['vcycles_per_solve']:.3f} V-cycles/solve")
    main_run.pop("kept_state", None)

    learned = None
    if compare and not args.no_learned:
        import numpy as np
        g, tinfo = train_guess(me, args.harvest_steps, args.harvest_runs, 2024,
                               args.train_epochs)
        e3, lr = run_learned(pb, me, g, args.warmup, win["t_end"])
        st3 = e3.state()
        e3.close()
        g.close()
        ps = win_state["slip"]
        dslip = float(np.max(np.abs(st3["slip"] - ps)) / np.max(np.abs(ps)))
        log(f"window t {win['t_start']:.4g} -> {win['t_end']:.4g} s: plain "
            f"{win['steps']} steps, {win['steps_per_s']:.2f} steps/s, "
            f"{win['vcycles_per_solve']:.3f} V-cycles/solve; learned "
            f"{lr['steps']} steps, {lr['steps_per_s']:.2f} steps/s, "
            f"{lr['vcycles_per_solve']:.3f} V-cycles/solve, accepted "
            f"{lr['accepted']}/{lr['predicted']}; slip rel diff at the window "
            f"end {dslip:.2e}")
        learned = {k: lr[k] for k in ("steps", "steps_per_s", "vcycles_per_solve",
                                      "solves_per_step", "t_start", "t_end",
                                      "sim_s_per_s", "predicted", "accepted")}
        learned["plain_same_window"] = {k: win[k] for k in (
            "steps", "steps_per_s", "vcycles_per_solve", "solves_per_step",
            "t_start", "t_end", "sim_s_per_s")}
        learned["ends_at_window_end"] = lr["t_end"] == win["t_end"]
        learned["slip_max_rel_diff_vs_plain_at_window_end"] = dslip
        learned["training"] = tinfo
    # Projection start against the plain start over the same window, and the
    # accuracy of both against the plain start at a 100x tighter residual
    projection = None
    if compare and not args.no_projection:
        t_end = win["t_end"]
        ref, st_ref = run_projection(me, args.warmup, t_end, 0, 1e-12)
        ds, dv = fault_diff(win_state, st_ref)
        log(f"plain 1e-12 over the window: {ref['steps_per_s']:.2f} steps/s, "
            f"{ref['vcycles_per_solve']:.3f} V-cycles/solve; plain 1e-10 vs it: "
            f"slip {ds:.2e}, log10 V {dv:.2e}")
        projection = {
            "window": [win["t_start"], t_end],
            "plain_1e-10": {"steps_per_s": round(win["steps_per_s"], 2),
                            "vcycles_per_solve": round(win["vcycles_per_solve"], 3),
                            "slip_max_rel_diff_vs_plain_1e-12": ds,
                            "log10V_max_abs_diff_vs_plain_1e-12": dv},
            "plain_1e-12": {"steps_per_s": round(ref["steps_per_s"], 2),
                            "vcycles_per_solve": round(ref["vcycles_per_solve"], 3)},
            "runs": []}
        for slots, tol in ((args.proj_slots, args.proj_tol), (args.proj_slots, 1e-11),
                           (6, 1e-10)):
            pr, st = run_projection(me, args.warmup, t_end, slots, tol)
            ds, dv = fault_diff(st, st_ref)
            rec = {k: (round(v, 4) if isinstance(v, float) and k not in ("tol", "t_start", "t_end") else v)
                   for k, v in pr.items()}
            rec["ends_at_window_end"] = pr["t_end"] == t_end
            rec["speedup_vs_plain_1e-10"] = round(pr["steps_per_s"] / win["steps_per_s"], 3)
            rec["slip_max_rel_diff_vs_plain_1e-12"] = ds
            rec["log10V_max_abs_diff_vs_plain_1e-12"] = dv
            projection["runs"].append(rec)
            log(f"projection {slots} slots, residual {tol:g}: {pr['steps_per_s']:.2f} "
                f"steps/s ({rec['speedup_vs_plain_1e-10']:.2f}x plain 1e-10), "
                f"{pr['vcycles_per_solve']:.3f} V-cycles/solve, used "
                f"{pr.get('used')}/{pr.get('offered')}; vs plain 1e-12: slip {ds:.2e}, "
                f"log10 V {dv:.2e}")

    # per-kernel durations of the tiled cycle (each kernel launched back to
    # back on the engine's buffers, its shape as tuned)
    breakdown = None
    if args.mode != "ref64":
        us = eng.kernel_breakdown(20)
        lt = (len(us) - 1) // 2
        names = ([f"down{l}" for l in range(lt)] + ["tail"] +
                 [f"up{l}" for l in range(lt - 1, -1, -1)])
        breakdown = {n: round(v, 2) for n, v in zip(names, us)}
        log("kernels (us, back to back): " +
            " ".join(f"{n} {v:.1f}" for n, v in breakdown.items()))

    n0 = pb.params.ni * pb.params.nk
    nu = MODES[args.mode]["n_post"]  # the up kernel's sweeps
    roofline = None
    if timed is not None and main_run.get("kernel_calls"):
        avg_s = main_run["kernel_ms_total"] / main_run["kernel_calls"] * 1e-3
        bytes_ = kernel_bytes(args.mode, n0, nu)
        achieved = bytes_ / avg_s / 1e9
        # HBM bytes per launch from the rocprofv3 counter passes on the same
        # kernel and tile shape (profiles/counters_basin_1025_fast64.md)
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic.json")
        shape_up0 = None
        if TUNED.get(args.mode):
            shape_up0 = list(TUNED[args.mode]["shapes"][0][1])
        if os.path.exists(tpath):
            with open(tpath) as fh:
                tr = json.load(fh)
            rec = tr.get(f"{args.mode}:up0")
            if rec and list(rec.get("shape", [])) == shape_up0:
                traffic = rec["hbm_bytes_per_launch"]
        roofline = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": f"k_up level 0 ({args.mode}), tile {shape_up0}",
            "avg_launch_us": round(avg_s * 1e6, 2),
            "launches": main_run["kernel_calls"],
            "algorithmic_bytes_per_launch": bytes_,
        }

    modes = {}
    if not args.no_compare:
        for m, k in (("ir32", args.compare_steps), ("ref64", ref_steps)):
            if m == args.mode:
                continue
            e2, r = run_mode(pb, me, m, args.warmup, k)
            e2.close()
            modes[m] = {key: r[key] for key in ("steps", "steps_per_s",
                                                "vcycles_per_solve", "solves_per_step",
                                                "t_start", "t_end")}
            log(f"{m}: {r['steps_per_s']:.2f} steps/s, {r['vcycles_per_solve']:.2f} "
                f"V-cycles/solve over {k} steps")

    ensemble = None
    if args.ensemble_leg and args.mode != "ref64":
        ensemble = run_ensemble(args.ensemble_leg, args.ensemble_leg_steps,
                                args.ensemble_mode, 0)
        log(f"ensemble leg: {args.ensemble_leg} members x {args.ensemble_leg_steps} steps: "
            f"{ensemble['steps_per_s_ensemble']:.1f} steps/s together, "
            f"{ensemble['steps_per_s_alone']:.1f} one after the other, bitwise equal: "
            f"{ensemble['bitwise_equal_to_alone']}")

    cpu = None
    if not args.no_cpu and args.mode == "fast64":
        cb = cpu_baseline(pb, me, eng, args.cpu_seconds)
        log(f"cpu baseline: {cb['cpu_steps_per_s']:.3f} steps/s on one core, "
            f"{cb['cpu_all_steps_per_s']:.3f} on {cb['threads']} threads, over "
            f"{cb['steps']} steps; GPU on the same steps "
            f"{cb['gpu_steps_per_s_same_steps']:.1f}; bitwise equal: {cb['bitwise_equal']}")
        cpu = {"value": round(cb["cpu_steps_per_s"], 4), "unit": "steps/s",
               "cores": 1, "kind": "port",
               "sample": (f"the first {cb['steps']} accepted steps of basin_1025 "
                          f"from t = 0 (to t = {cb['t_end']:.6g} s), same fast FP64 "
                          f"cycle, oracle/ C port on one core"),
               "vcycles_per_solve": round(cb["vcycles_per_solve"], 4),
               "attempts_per_step": round(cb["attempts_per_step"], 4),
               "all_cores": {"value": round(cb["cpu_all_steps_per_s"], 4),
                             "cores": cb["threads"],
                             "sample": "the same steps, OpenMP over the multigrid's "
                                       "point loops"},
               "gpu_steps_per_s_same_steps": round(cb["gpu_steps_per_s_same_steps"], 2),
               "bitwise_equal_to_gpu": cb["bitwise_equal"]}
        if args.c1:
            cpu["config_1"] = c1_leg()
    eng.close()

    head = main_run
    same = main_run["marks"][ref_steps]
    line = {
        "metric": METRIC, "value": round(head["steps_per_s"], 3),
        "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / head["steps_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.mode == "ir32" else "f64",
        "data": "synthetic",
        "config": {"workload": "basin_1025", "grid": "1025x1025", "levels": 7,
                   "residual": "1e-10 relative", "mode": args.mode,
                   "cycle": ("reference V(32,32)" if args.mode == "ref64" else
                             "V(1,3), red/black relaxation 1.15, direct 17x17 "
                             "coarsest solve (its inverse stored in FP32, "
                             "products and sums in FP64)"),
                   "initial_guess": "previous solution",
                   "tile_shapes": TUNED.get(args.mode),
                   "timed_window": {"t_start_s": head["t_start"], "t_end_s": head["t_end"],
                                    "after_warmup_steps": args.warmup,
                                    "vcycles_per_solve": round(head["vcycles_per_solve"], 3),
                                    "solves_per_step": round(head["solves_per_step"], 3)}},
        "vcycles_per_solve": round(head["vcycles_per_solve"], 3),
        "solves_per_step": round(head["solves_per_step"], 3),
        "vcycles_per_solve_learned_vs_plain": (
            None if learned is None else
            [round(learned["vcycles_per_solve"], 3),
             round(learned["plain_same_window"]["vcycles_per_solve"], 3)]),
        "steps_per_s_learned_vs_plain": (
            None if learned is None else
            [round(learned["steps_per_s"], 2),
             round(learned["plain_same_window"]["steps_per_s"], 2)]),
        "plain": {k: main_run[k] for k in ("steps", "steps_per_s",
                                            "vcycles_per_solve", "solves_per_step")},
        "comparison_window": win,
        "learned": learned,
        "vcycles_per_solve_projection": (
            None if projection is None else
            round(projection["runs"][0]["vcycles_per_solve"], 3)),
        "projection": projection,
        "roofline": roofline,
        "kernel_breakdown_us": breakdown,
        "cpu_baseline": cpu,
        "modes": modes,
        "ensemble": ensemble,
    }
    if "ref64" in modes:
        line["speedup_vs_ref64"] = {
            "steps": ref_steps, "fast64_steps_per_s": round(same["steps_per_s"], 2),
            "ref64_steps_per_s": round(modes["ref64"]["steps_per_s"], 2),
            "speedup": round(same["steps_per_s"] / modes["ref64"]["steps_per_s"], 2),
            "same_t_end": modes["ref64"]["t_end"] == same["t_end"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
