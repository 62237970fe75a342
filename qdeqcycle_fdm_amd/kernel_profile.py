"""Kernel-level measurement driver for profilers (rocprofv3 kernel traces and
counter passes).

Builds the engine for a workload grid, tunes the tile shapes, and launches
each kernel of the tiled V-cycle back to back (qdeq_kernel_breakdown: direct
launches, no captured graphs, no time stepping), plus optionally a plain
device copy of known size for counter calibration.  Prints one JSON line
with the shapes and the HIP-event durations.  Does not import torch (its
bundled HIP runtime would replace the one the library was built against).

    python -m qdeqcycle_fdm_amd.kernel_profile --n 1025 --mode fast64 --reps 20

With --rk-steps K it instead runs K accepted RK43 steps (after --rk-warmup
untimed ones) of the basin workload at residual --tol, with the projection
start from --proj-slots stored solutions (0: the previous solution), and
prints the rate.  A single simulation launches its V-cycle kernels one by
one already; QDEQ_DIRECT_LAUNCH=1 under a kernel tracer makes sure no HIP
graph launch reaches it.
"""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# the fast cycle of the bench: V(1,3), relaxation 1.15
MODES = {
    "fast64": dict(cycle=1, precision=0, n_pre=1, n_post=3, n_coarse=0, relax=1.15,
                   tiles=(64, 16)),
    "ir32": dict(cycle=1, precision=1, n_pre=1, n_post=3, n_coarse=0, relax=1.15,
                 tiles=(64, 32)),
}
LEVELS = {257: 5, 513: 6, 1025: 7, 2049: 8}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--n", type=int, default=1025, choices=sorted(LEVELS))
    ap.add_argument("--mode", default="fast64", choices=sorted(MODES))
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--shapes", default=None,
                    help="JSON per-level shapes to pin instead of tuning "
                         "(as bench.py reports them)")
    ap.add_argument("--copy-doubles", type=int, default=0,
                    help="also time a plain copy of this many doubles")
    ap.add_argument("--rk-steps", type=int, default=0,
                    help="run this many accepted RK43 steps instead")
    ap.add_argument("--rk-warmup", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--proj-slots", type=int, default=0)
    ap.add_argument("--n-pre", type=int, default=None,
                    help="pre-smoothing sweeps (default: the mode's)")
    ap.add_argument("--n-post", type=int, default=None)
    ap.add_argument("--relax", type=float, default=None)
    args = ap.parse_args(argv)
    for key in ("n_pre", "n_post", "relax"):
        if getattr(args, key) is not None:
            MODES[args.mode][key] = getattr(args, key)

    from qdeqcycle_fdm_amd import config, lib
    pb = config.make_bp1_problem(args.n, basin=True, levels=LEVELS[args.n],
                                 tol_mode=config.TOL_REL, tol=args.tol)
    if args.rk_steps:
        import time
        me = lib.build_mu_equi(pb.mu)
        kw = dict(MODES[args.mode])
        tiles = kw.pop("tiles")
        with lib.Engine(pb, me, max_vcycles=200, **kw) as eng:
            eng.set_tiles(True, *tiles)
            shapes = eng.autotune_tiles(5) if not args.no_autotune else None
            if args.proj_slots:
                eng.projection_enable(args.proj_slots)
            eng.rk43_setup()
            eng.rk43_step(args.rk_warmup)
            lib.device_synchronize(0)
            i0 = eng.state()["info"]
            t0 = time.perf_counter()
            info = eng.rk43_step(args.rk_steps)
            lib.device_synchronize(0)
            sec = time.perf_counter() - t0
            solves = info["solves"] - i0["solves"]
            out = {"n": args.n, "mode": args.mode, "tol": args.tol,
                   "proj_slots": args.proj_slots, "tile_shapes": shapes,
                   "steps": args.rk_steps, "steps_per_s": args.rk_steps / sec,
                   "solves": solves,
                   "vcycles_per_solve": (info["vcycles"] - i0["vcycles"]) / max(solves, 1)}
        print(json.dumps(out), flush=True)
        return
    me = lib.build_mu_equi(pb.mu)
    kw = dict(MODES[args.mode])
    tiles = kw.pop("tiles")
    out = {"n": args.n, "mode": args.mode, "reps": args.reps}
    with lib.Engine(pb, me, **kw) as eng:
        eng.set_tiles(True, *tiles)
        if args.shapes:
            eng.set_tile_shapes(json.loads(args.shapes))
        elif not args.no_autotune:
            eng.autotune_tiles(5)
        out["tile_shapes"] = eng.tile_shapes()
        us = eng.kernel_breakdown(args.reps)
        lt = (len(us) - 1) // 2
        names = ([f"down{l}" for l in range(lt)] + ["tail"] +
                 [f"up{l}" for l in range(lt - 1, -1, -1)])
        out["kernel_us"] = {n: round(v, 3) for n, v in zip(names, us)}
        if args.copy_doubles:
            ms = eng.copy_probe(args.copy_doubles, reps=args.reps)
            out["copy"] = {"doubles": args.copy_doubles, "avg_ms": ms,
                           "GB/s": 16.0 * args.copy_doubles / ms / 1e6}
        lib.device_synchronize(0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
