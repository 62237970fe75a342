"""GPU checks of ensemble stepping (qdeq_ensemble_step): members with uneven
grids and seeded frictional parameters, stepped together from one host
thread on their own streams, are each bit-identical to the same simulation
run alone (FP64, FP32-IR, and both with the projection start); a member
that fails (forced from the host: a V-cycle cap of one at an unreachable
tolerance) reports its status without disturbing the others."""
import numpy as np
import pytest

from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu

LEVELS = {65: 3, 129: 4, 257: 5}


def _members(count, seed):
    out = []
    for m in config.ensemble_members(count, seed, sizes=(65, 129, 257)):
        pb = config.make_bp1_problem(m["n"], basin=True, levels=LEVELS[m["n"]],
                                     Dc=m["Dc"], sigma_n=m["sigma_n"],
                                     b_seis=m["b_seis"], tol_mode=config.TOL_REL,
                                     tol=1e-10)
        out.append(pb)
    return out


def _engine(pb, precision=0, slots=0, **kw):
    eng = lib.Engine(pb, cycle=config.CYCLE_FAST, n_pre=2, n_post=2, n_coarse=0,
                     precision=precision, **kw)
    eng.set_tiles(True, 64, 16 if precision == 0 else 32)
    if slots:
        eng.projection_enable(slots)
    eng.rk43_setup()
    return eng


@pytest.mark.parametrize("precision,slots", [(config.PREC_FP64, 0), (config.PREC_FP32_IR, 0),
                                             (config.PREC_FP64, 6), (config.PREC_FP32_IR, 6)])
def test_ensemble_members_equal_solo_runs(precision, slots):
    pbs = _members(6, seed=3)
    steps = 40
    solo = []
    for pb in pbs:
        with _engine(pb, precision, slots) as e:
            info = e.rk43_step(steps)
            solo.append((info, e.state()))
    engines = [_engine(pb, precision, slots) for pb in pbs]
    status = lib.ensemble_step(engines, steps)
    assert status == [lib.QDEQ_OK] * len(pbs)
    for (info, st), e in zip(solo, engines):
        s2 = e.state()
        assert s2["info"]["t"] == info["t"]
        assert s2["info"]["vcycles"] == info["vcycles"]
        for k in ("slip", "psi", "V", "U"):
            np.testing.assert_array_equal(s2[k], st[k])
    # a second call continues from where the first stopped
    status = lib.ensemble_step(engines, 5)
    with _engine(pbs[0], precision, slots) as e:
        e.rk43_step(steps + 5)
        ref = e.state()
    np.testing.assert_array_equal(engines[0].state()["slip"], ref["slip"])
    for e in engines:
        e.close()


def test_failing_member_is_isolated():
    pbs = _members(3, seed=5)
    engines = [_engine(pbs[0]), _engine(pbs[1], tol=1e-30, max_vcycles=1),
               _engine(pbs[2])]
    with pytest.raises(lib.QdeqError):
        lib.ensemble_step(engines, 10)
    # member 1 failed; the others completed their steps
    with _engine(pbs[2]) as e:
        e.rk43_step(10)
        ref = e.state()
    np.testing.assert_array_equal(engines[2].state()["slip"], ref["slip"])
    assert engines[0].state()["info"]["it"] == 10
    for e in engines:
        e.close()
