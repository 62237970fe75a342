#!/usr/bin/env python3
"""Accepted RK43 steps per second at 1025^2 to a 1e-10 relative elliptic
residual, on one MI355X.

Workload (basin_1025): 48 km x 48 km antiplane fault model on a 1025 x 1025
grid, low-rigidity basin (20 GPa in the top 2 km within 10 km of the fault,
500 m tanh edges, 32 GPa elsewhere), rate-and-state friction with radiation
damping, 7-level multigrid, 1e-10 relative residual per stage solve.  The
default run warms up W steps from t = 0 and times the next K accepted steps
(W + K = 3020: the workload's 3000-step run, which spans its first seismic
events).  Inputs are synthetic (generated, no data files).

Modes: fast64 (the headline: FP64 V(2,2) fast cycle on the tiled LDS
kernels, exact coarsest solve), ir32 (FP32 correction cycles in FP64
iterative refinement) and ref64 (the reference's own V-cycle: 32 sweeps per
level, FP64).  One JSON line goes to stdout; progress goes to stderr.

The CPU baseline is the project's FP64 port of the solver (oracle/) on one
host core, restarted from the GPU's state at the end of the timed region and
run for about --cpu-seconds; the GPU then runs the same steps and the two
final states are compared bit for bit.
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

MODES = {
    "fast64": dict(cycle=1, precision=0, n_pre=2, n_post=2, n_coarse=0,
                   tiles=(64, 16)),
    "ir32": dict(cycle=1, precision=1, n_pre=2, n_post=2, n_coarse=0,
                 tiles=(64, 32)),
    "ref64": dict(cycle=0, precision=0, n_pre=32, n_post=32, n_coarse=32,
                  tiles=None),
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


def make_engine(pb, me, mode):
    from qdeqcycle_fdm_amd import lib
    kw = dict(MODES[mode])
    tiles = kw.pop("tiles")
    eng = lib.Engine(pb, me, **kw)
    if tiles:
        eng.set_tiles(True, *tiles)
    return eng


def kernel_bytes(mode, n0, nu):
    """Algorithmic bytes of one launch of the level-0 up kernel (bilinear
    prolongation + correction, nu red/black sweeps, and in FP64 the residual
    norm): per point a prolongation + correction moves 2.25 words, a sweep 5
    (u read+write, f, two face coefficients), a residual 5."""
    s = 4 if mode == "ir32" else 8
    per = 2.25 + nu * 5 + (5 if mode == "fast64" else 0)
    return per * s * n0


def run_mode(pb, me, mode, warmup, steps, timed_kernel=None):
    from qdeqcycle_fdm_amd import lib
    eng = make_engine(pb, me, mode)
    eng.rk43_setup()
    if warmup:
        eng.rk43_step(warmup)
    sync()
    if timed_kernel is not None:
        eng.set_kernel_timing(timed_kernel)
    i0 = eng.state()["info"]
    t0 = time.perf_counter()
    info = eng.rk43_step(steps)
    sync()
    t1 = time.perf_counter()
    out = dict(seconds=t1 - t0, steps=steps, steps_per_s=steps / (t1 - t0),
               vcycles=info["vcycles"] - i0["vcycles"],
               solves=info["solves"] - i0["solves"], t_start=i0["t"],
               t_end=info["t"], V_max_end=info["V_max"])
    out["vcycles_per_solve"] = out["vcycles"] / max(out["solves"], 1)
    out["solves_per_step"] = out["solves"] / steps
    if timed_kernel is not None:
        ms, calls = eng.kernel_timing()
        out["kernel_ms_total"], out["kernel_calls"] = ms, calls
        eng.set_kernel_timing(lib.TIME_OFF)
    return eng, out


def cpu_baseline(pb, me, eng, seconds):
    """Oracle restarted from the engine's state; the engine then runs the
    same steps; returns the CPU rate, the GPU rate on the same steps and
    whether the final states agree bit for bit."""
    import numpy as np
    import oracle
    st = eng.state()
    info = st["info"]
    p = MODES["fast64"]
    sim = oracle.Sim(pb, me, cycle=1, n_pre=p["n_pre"], n_post=p["n_post"],
                     n_coarse=p["n_coarse"])
    oracle.set_libm(False)
    sim.set_state(info["t"], info["dt"], info["it"], info["errA0"], info["errA1"],
                  st["U"], st["slip"], st["psi"])
    n = 0
    t0 = time.perf_counter()
    while True:
        sim.step()
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    cpu_s = time.perf_counter() - t0
    sync()
    g0 = time.perf_counter()
    eng.rk43_step(n)
    sync()
    gpu_s = time.perf_counter() - g0
    st2 = eng.state()
    ar = sim.arrays()
    same = all(np.array_equal(st2[k], ar[k]) for k in ("slip", "psi", "V", "U"))
    same = same and st2["info"]["t"] == sim.scalars()["t"]
    return dict(steps=n, cpu_seconds=cpu_s, cpu_steps_per_s=n / cpu_s,
                gpu_steps_per_s_same_steps=n / gpu_s, bitwise_equal=bool(same),
                t_start=info["t"])


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
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="run the V-cycle as one captured graph without the "
                         "HIP-event bracket of the level-0 up kernel (used for "
                         "the rocprofv3 passes)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="N",
                    help="only time a plain device copy of N doubles (counter "
                         "calibration) and exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU "
              "path (independent simulations shard without communication; see "
              "DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from qdeqcycle_fdm_amd import config, lib
    lib.load()
    pb = config.make_bp1_problem(1025, basin=True, tol_mode=config.TOL_REL,
                                 tol=1e-10)
    if args.copy_probe:
        with lib.Engine(pb) as eng:
            ms = eng.copy_probe(args.copy_probe, reps=10)
        print(json.dumps({"copy_probe_doubles": args.copy_probe, "avg_ms": ms,
                          "GB/s": 2 * 8 * args.copy_probe / ms / 1e6}))
        return
    t0 = time.perf_counter()
    me = lib.build_mu_equi(pb.mu)
    log(f"mu_equi on the GPU: {time.perf_counter() - t0:.2f} s")

    timed = (lib.TIME_UP0 if args.mode != "ref64" and not args.no_kernel_timing
             else None)
    eng, main_run = run_mode(pb, me, args.mode, args.warmup, args.steps, timed)
    log(f"{args.mode}: {main_run['steps_per_s']:.2f} steps/s, "
        f"{main_run['vcycles_per_solve']:.2f} V-cycles/solve, "
        f"t {main_run['t_start']:.4g} -> {main_run['t_end']:.4g} s")

    n0 = pb.params.ni * pb.params.nk
    nu = MODES[args.mode]["n_pre"]
    roofline = None
    if timed is not None and main_run.get("kernel_calls"):
        avg_s = main_run["kernel_ms_total"] / main_run["kernel_calls"] * 1e-3
        bytes_ = kernel_bytes(args.mode, n0, nu)
        achieved = bytes_ / avg_s / 1e9
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tpath):
            with open(tpath) as fh:
                tr = json.load(fh)
            key = f"{args.mode}:up0"
            if key in tr:
                traffic = tr[key]["hbm_bytes_per_launch"]
        roofline = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": f"k_up level 0 ({args.mode})",
            "avg_launch_us": round(avg_s * 1e6, 2),
            "launches": main_run["kernel_calls"],
            "algorithmic_bytes_per_launch": bytes_,
        }

    modes = {}
    if not args.no_compare:
        for m, k in (("ir32", args.compare_steps), ("ref64", max(args.compare_steps // 10, 5))):
            if m == args.mode:
                continue
            e2, r = run_mode(pb, me, m, args.warmup, k)
            e2.close()
            modes[m] = {key: r[key] for key in ("steps", "steps_per_s",
                                                "vcycles_per_solve", "solves_per_step",
                                                "t_start", "t_end")}
            log(f"{m}: {r['steps_per_s']:.2f} steps/s, {r['vcycles_per_solve']:.2f} "
                f"V-cycles/solve over {k} steps")

    cpu = None
    if not args.no_cpu and args.mode == "fast64":
        cb = cpu_baseline(pb, me, eng, args.cpu_seconds)
        log(f"cpu baseline: {cb['cpu_steps_per_s']:.3f} steps/s over {cb['steps']} "
            f"steps; GPU on the same steps {cb['gpu_steps_per_s_same_steps']:.1f}; "
            f"bitwise equal: {cb['bitwise_equal']}")
        cpu = {"value": round(cb["cpu_steps_per_s"], 4), "unit": "steps/s",
               "cores": 1, "kind": "port",
               "sample": (f"{cb['steps']} accepted steps of basin_1025 restarted "
                          f"from the GPU state at t={cb['t_start']:.6g} s "
                          f"(after the timed region), same fast FP64 cycle, "
                          f"oracle/ C port on one core"),
               "gpu_steps_per_s_same_steps": round(cb["gpu_steps_per_s_same_steps"], 2),
               "bitwise_equal_to_gpu": cb["bitwise_equal"]}
    eng.close()

    line = {
        "metric": METRIC, "value": round(main_run["steps_per_s"], 3),
        "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / main_run["steps_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.mode == "ir32" else "f64",
        "data": "synthetic",
        "config": {"workload": "basin_1025", "grid": "1025x1025", "levels": 7,
                   "residual": "1e-10 relative", "mode": args.mode,
                   "cycle": ("reference V(32,32)" if args.mode == "ref64" else
                             "V(2,2), exact 17x17 coarsest solve"),
                   "sim_time_s": [main_run["t_start"], main_run["t_end"]]},
        "vcycles_per_solve": round(main_run["vcycles_per_solve"], 3),
        "solves_per_step": round(main_run["solves_per_step"], 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "modes": modes,
    }
    if "ref64" in modes:
        line["speedup_vs_ref64"] = round(main_run["steps_per_s"] /
                                         modes["ref64"]["steps_per_s"], 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
