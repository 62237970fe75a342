"""GPU checks of the reference's fault output records (qdeq_output_open):
record 0 and the periodic records hold, byte for byte in the reference's
layout, the state the CPU oracle computes for the same steps."""
import os

import numpy as np
import pytest

import oracle
from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu

FAST = dict(cycle=config.CYCLE_FAST, n_pre=2, n_post=2, n_coarse=0)


def test_output_records_match_oracle(tmp_path):
    pb = config.make_bp1_problem(65, basin=True, levels=3,
                                 tol_mode=config.TOL_REL, tol=1e-10)
    me = oracle.mu_equi(pb.mu)
    p = pb.params
    nz, nx = p.nz, p.nx
    oracle.set_libm(False)
    sim = oracle.Sim(pb, me, cycle=1, n_pre=2, n_post=2, n_coarse=0)
    sim.init()
    ref = [(sim.scalars(), sim.arrays())]
    steps, every = 24, 3
    for _ in range(steps):
        sim.step()
        ref.append((sim.scalars(), sim.arrays()))
    with lib.Engine(pb, me, **FAST) as eng:
        eng.rk43_setup()
        eng.output_open(str(tmp_path), every)
        eng.rk43_step(steps)
        eng.output_close()
    # record 0: t, V, Slip, Ts, psi, U, Tn, a, b, dt
    r0 = np.fromfile(tmp_path / "output_fault_0.bin", dtype=np.float64)
    assert r0.size == 1 + 4 * nz + nx * nz + 3 * nz + 1
    sc, ar = ref[0]
    o = 1
    assert r0[0] == sc["t"] == 0.0
    for key in ("V", "slip", "Ts", "psi"):
        np.testing.assert_array_equal(r0[o:o + nz], ar[key], err_msg=key)
        o += nz
    np.testing.assert_array_equal(r0[o:o + nx * nz].reshape(nz, nx), ar["U"])
    o += nx * nz
    for arr in (pb.tau_n0, pb.a, pb.b):
        np.testing.assert_array_equal(r0[o:o + nz], arr)
        o += nz
    assert r0[o] == sc["dt"]
    # periodic records every `every` steps: t, V, Slip, Ts, psi, dt
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("output_fault_"))
    assert len(files) == 1 + steps // every
    for k in range(1, steps // every + 1):
        r = np.fromfile(tmp_path / f"output_fault_{k}.bin", dtype=np.float64)
        assert r.size == 4 * nz + 2
        sc, ar = ref[k * every]
        assert r[0] == sc["t"] and r[-1] == sc["dt"]
        for j, key in enumerate(("V", "slip", "Ts", "psi")):
            np.testing.assert_array_equal(r[1 + j * nz:1 + (j + 1) * nz], ar[key],
                                          err_msg=f"record {k} {key}")
