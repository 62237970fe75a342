"""The oracle's RK43 step composition (odeSolver.cpp:1462-1633) pinned by
mathematics, not by comparison with another implementation: the committed
step and the embedded solution are the tableau's weighted sums of the stage
rates, every stage state is the tableau's row sum, every stage solve sees
the boundary data at its stage time, and the propagated / embedded
solutions converge with orders 5 / 4 in the step size (tests/rk43_checks.py).
The same checks, run against oracle builds with a planted mistake, fail.
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)

import rk43_checks  # noqa: E402


@pytest.fixture(scope="module")
def results():
    return rk43_checks.run_all()


def test_rk43_commits_fourth_order_weights(results):
    c = results["commit"]
    assert max(c["slip_y4"], c["psi_y4"]) < rk43_checks.ROUND, c
    assert max(c["slip_y3"], c["psi_y3"]) < rk43_checks.ROUND, c


def test_rk43_stage_states_follow_the_tableau(results):
    assert results["stage_sums"]["stage_sums"] < rk43_checks.ROUND, results["stage_sums"]


def test_rk43_stage_solves_see_their_stage_times(results):
    b = results["boundary"]
    assert b["calls"] >= 6
    assert b["remote_column"] < 1e-15, b
    assert b["fault_column"] == 0.0, b


def test_rk43_convergence_orders(results):
    o = results["orders"]
    # errors well above the solves' noise (1e-13 relative residual)
    assert min(o["e4"]) > 1e-13 and min(o["e3"]) > 1e-12, o
    for p4, p3 in zip(o["order_y4"], o["order_y3"]):
        assert 4.6 < p4 < 5.5, o
        assert 3.6 < p3 < 4.6, o
    assert abs(o["order_y4"][-1] - 5.0) < 0.35 and abs(o["order_y3"][-1] - 4.0) < 0.35, o
    assert all(results["pass"].values()), results["pass"]


# planted mistake -> checks that must fail (ORC_MUTANT in oracle/qdeq_oracle.c)
MUTANTS = {
    1: ("stage_sums", "order_y4"),      # a42 and a43 swapped in stage 4
    2: ("remote_column", "order_y4"),   # stage 4 solved at stage 3's time
    3: ("commit_y4",),                  # embedded y3 committed instead of y4
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_rk43_checks_detect_planted_mistakes(mutant, tmp_path):
    so = str(tmp_path / f"liboracle_m{mutant}.so")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mutant",
                    f"MUTANT={mutant}", f"OUT={so}"], check=True)
    env = dict(os.environ, QDEQ_ORACLE_LIB=so)
    out = subprocess.run([sys.executable, os.path.join(HERE, "rk43_checks.py")],
                         env=env, check=True, capture_output=True, text=True,
                         timeout=300)
    res = json.loads(out.stdout.strip().splitlines()[-1])
    failed = [k for k, ok in res["pass"].items() if not ok]
    for name in MUTANTS[mutant]:
        assert name in failed, (mutant, res["pass"])
