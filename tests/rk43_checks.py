"""Mathematical checks of the oracle's RK43 step (orc_sim_step), the
composition of odeSolver.cpp:1462-1633:

* commit: the accepted step adds dt * sum_j b_j F_j to the state (the 4th
  order weights are propagated, odeSolver.cpp:1583-1592, 1608), and the
  embedded solution is y_n + dt * sum_j hb_j F_j (1569-1580);
* stage sums: stage state s is y_n + dt * sum_{j<s} a_sj F_j (1491-1563);
* boundary times: the solve of stage s sees the remote column at
  Vini (t + c_s dt) / 2 + SlipShift and the fault column at half the stage
  slip (setBoundary, 1150-1188), and the stage-1 solve after the step sees
  time t + dt;
* convergence order: from one fixed state, with the controller out of the
  way (every attempt accepted), the local error of y4 falls like dt^5 and
  that of y3 like dt^4, measured against 16 substeps of dt/16.

F_j are the stage rates (V, dpsi) the oracle computed, a, b, hb, c the
tableau, whose order conditions tests/test_oracle.py checks.  The state is
the 65^2 basin model after 100 accepted steps (interseismic creep), solved
to a 1e-13 relative residual so that the elliptic solves do not blur the
local errors.

Run as a script (python tests/rk43_checks.py) it prints the results as one
JSON line; the tests run it that way against oracle builds with planted
mistakes (QDEQ_ORACLE_LIB), which must fail at least one check.
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from qdeqcycle_fdm_amd import config  # noqa: E402

CYCLE = dict(cycle=1, n_pre=1, n_post=3, n_coarse=0, relax=1.15, max_vcycles=200)
# rounding allowance of a sum of six FP64 terms, relative to its magnitude
ROUND = 1e-14


def problem(n: int = 65, tol: float = 1e-13):
    pb = config.make_bp1_problem(n, basin=True, tol_mode=config.TOL_REL, tol=tol)
    # every attempt accepted: the controller never rejects (totErr < totTol)
    return dataclasses.replace(pb, params=dataclasses.replace(pb.params, totTol=1e300))


def base_state(pb, steps: int = 100):
    """(t, dt, U, slip, psi) after `steps` accepted steps of the normal run."""
    real = dataclasses.replace(pb, params=dataclasses.replace(pb.params, totTol=1e-7))
    sim = oracle.Sim(real, **CYCLE)
    sim.init()
    for _ in range(steps):
        sim.step()
    sc, ar = sim.scalars(), sim.arrays()
    return sc["t"], sc["dt"], ar["U"], ar["slip"], ar["psi"]


def step_from(pb, state, h: float, nsub: int = 1):
    """A simulation restarted at `state` that has taken nsub steps of h/nsub;
    also returns the stage-1 rates at the start."""
    t, _, U, slip, psi = state
    sim = oracle.Sim(pb, **CYCLE)
    sim.set_state(t, h / nsub, 10, 0.0, 0.0, U, slip, psi)
    a0 = sim.arrays()
    for _ in range(nsub):
        sim.set_dt(h / nsub)
        sim.step()
    return sim, a0


def _rel(diff, scale):
    return float(np.max(np.abs(diff)) / max(float(np.max(np.abs(scale))), 1e-300))


def check_commit(pb, state, h):
    T = oracle.tableau()
    sim, a0 = step_from(pb, state, h)
    la, a1 = sim.last_attempt(), sim.arrays()
    V = la["V"].copy()
    D = la["dpsi"].copy()
    V[1], D[1] = a0["V"], a0["dpsi"]
    b, hb = T["b"], T["hb"]
    iv = slice(1, -1)
    out = {}
    for name, y0, F, y1, y3 in (("slip", a0["slip"], V, a1["slip"], la["y3s"]),
                                ("psi", a0["psi"], D, a1["psi"], la["y3p"])):
        inc4 = h * sum(b[j - 1] * F[j] for j in range(1, 7))
        inc3 = h * sum(hb[j - 1] * F[j] for j in range(1, 7))
        scale = np.abs(y0[iv]) + h * sum(abs(b[j - 1]) * np.abs(F[j][iv]) for j in range(1, 7))
        out[f"{name}_y4"] = _rel((y1 - y0 - inc4)[iv], scale)
        out[f"{name}_y3"] = _rel((y3 - y0 - inc3)[iv], scale)
    return out


def check_stage_sums(pb, state, h):
    T = oracle.tableau()
    sim, a0 = step_from(pb, state, h)
    la = sim.last_attempt()
    V = la["V"].copy()
    D = la["dpsi"].copy()
    V[1], D[1] = a0["V"], a0["dpsi"]
    a = T["a"]
    iv = slice(1, -1)
    worst = 0.0
    for s in range(2, 7):
        for y0, F, ks in ((a0["slip"], V, la["slip"][s]), (a0["psi"], D, la["psi"][s])):
            inc = h * sum(a[s - 1, j - 1] * F[j] for j in range(1, s))
            scale = np.abs(y0[iv]) + h * sum(abs(a[s - 1, j - 1]) * np.abs(F[j][iv])
                                             for j in range(1, s))
            worst = max(worst, _rel((ks - y0 - inc)[iv], scale))
    return {"stage_sums": worst}


def check_boundary_times(pb, state, h):
    T = oracle.tableau()
    sim, a0 = step_from(pb, state, h)
    la = sim.last_attempt()
    log = sim.stage_log()  # stages 2..6 of the attempt, then stage 1 at t + h
    t0 = state[0]
    p = pb.params
    ss = pb.U0[1, p.ni]  # SlipShift of row 1 (fault_init, fault.cpp:129-133)
    remote, fault = 0.0, 0.0
    expect = [(s, t0, T["c"][s - 1]) for s in range(2, 7)] + [(1, t0 + h, 0.0)]
    for (s, t, c), rec in zip(expect, log[-6:]):
        want = 0.5 * p.Vini * (t + c * h) + ss
        remote = max(remote, abs(rec[3] - want) / abs(want))
        slip_s = la["slip"][s][1] if s > 1 else sim.arrays()["slip"][1]
        fault = max(fault, abs(rec[4] - 0.5 * slip_s) / max(abs(slip_s), 1e-300))
    return {"remote_column": remote, "fault_column": fault, "calls": int(len(log))}


def convergence_orders(pb, state, fracs=(4, 8, 16), nsub=16):
    """Observed orders log2(e(h)/e(h/2)) of y4 and y3 for h = dt_n / frac."""
    dt0 = state[1]
    e4, e3 = [], []
    for f in fracs:
        h = dt0 / f
        one, _ = step_from(pb, state, h)
        ref, _ = step_from(pb, state, h, nsub)
        la, r = one.last_attempt(), ref.arrays()
        iv = slice(1, -1)
        e4.append(float(np.max(np.abs(la["y4s"][iv] - r["slip"][iv])) +
                        np.max(np.abs(la["y4p"][iv] - r["psi"][iv]))))
        e3.append(float(np.max(np.abs(la["y3s"][iv] - r["slip"][iv])) +
                        np.max(np.abs(la["y3p"][iv] - r["psi"][iv]))))
    p4 = [math.log2(e4[i] / e4[i + 1]) for i in range(len(fracs) - 1)]
    p3 = [math.log2(e3[i] / e3[i + 1]) for i in range(len(fracs) - 1)]
    return {"e4": e4, "e3": e3, "order_y4": p4, "order_y3": p3}


# acceptance of each check (the tests use the same rules)
def verdicts(res) -> dict:
    c = res["commit"]
    v = {
        "commit_y4": max(c["slip_y4"], c["psi_y4"]) < ROUND,
        "embedded_y3": max(c["slip_y3"], c["psi_y3"]) < ROUND,
        "stage_sums": res["stage_sums"]["stage_sums"] < ROUND,
        "remote_column": res["boundary"]["remote_column"] < 1e-15,
        "fault_column": res["boundary"]["fault_column"] == 0.0,
        # asymptotic orders 5 and 4; the last ratio of each within 0.35
        "order_y4": abs(res["orders"]["order_y4"][-1] - 5.0) < 0.35,
        "order_y3": abs(res["orders"]["order_y3"][-1] - 4.0) < 0.35,
    }
    return {k: bool(x) for k, x in v.items()}


def run_all(n: int = 65, steps: int = 100) -> dict:
    oracle.set_libm(False)
    pb = problem(n)
    st = base_state(pb, steps)
    res = {"dt_n": st[1], "t_n": st[0],
           "commit": check_commit(pb, st, st[1]),
           "stage_sums": check_stage_sums(pb, st, st[1]),
           "boundary": check_boundary_times(pb, st, st[1]),
           "orders": convergence_orders(pb, st)}
    res["pass"] = verdicts(res)
    return res


if __name__ == "__main__":
    print(json.dumps(run_all()))
