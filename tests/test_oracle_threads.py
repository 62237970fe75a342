"""The oracle's multigrid loops over independent points run under OpenMP
(levels above 16384 points); every point is computed by one thread with the
same expression, so RK43 runs on 1 and 4 threads are bit for bit equal."""
import numpy as np

import oracle
from qdeqcycle_fdm_amd import config


def _run(threads, steps=25):
    pb = config.make_bp1_problem(129, basin=True, levels=4,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    oracle.set_libm(False)
    used = oracle.set_threads(threads)
    try:
        sim = oracle.Sim(pb, me, cycle=1, n_pre=1, n_post=3, n_coarse=0, relax=1.15)
        sim.init()
        for _ in range(steps):
            sim.step()
        return used, sim.scalars(), sim.arrays()
    finally:
        oracle.set_threads(1)


def test_oracle_threads_do_not_change_results():
    u1, s1, a1 = _run(1)
    u4, s4, a4 = _run(4)
    assert u1 == 1 and u4 >= 2
    for key in ("t", "dt", "totErr", "V_max", "attempts", "vcycles"):
        assert s1[key] == s4[key], key
    for k in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(a1[k], a4[k], err_msg=k)
