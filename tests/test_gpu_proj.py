"""GPU checks of the projection start (qdeq_projection_*, csrc/proj.hip).

* The least-squares fit kernel equals the CPU restatement oracle/proj.py bit
  for bit (same summation order), for 2..8 stored columns at 257 and 1025
  fault nodes, including a dependent column.
* Disabled again (0 slots), a context runs bit for bit like one that never
  had it.
* At a 1e-12 relative residual the projection-started run tracks the plain
  run at the same tolerance (slip within 3e-7 relative at a common end
  time) with fewer than half the V-cycles per solve; nearly every offered
  start is used.
* It cannot be combined with an attached learned guess.
* The start decided on the device equals the host decision bit for bit,
  and an absurd or non-finite start (fault injection) is never used: the
  run then equals the plain run bit for bit on both paths.
"""
import numpy as np
import pytest

from oracle import proj as oproj
from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu

LEVELS = {129: 4, 257: 5, 1025: 7}


def _engine(n, tol=1e-10, **kw):
    pb = config.make_bp1_problem(n, basin=True, levels=LEVELS[n],
                                 tol_mode=config.TOL_REL, tol=tol)
    eng = lib.Engine(pb, cycle=config.CYCLE_FAST, n_pre=2, n_post=2, n_coarse=0,
                     max_vcycles=200, **kw)
    eng.set_tiles(True, 64, 16)
    return eng


def _columns(rng, cnt, n):
    x = np.linspace(0.0, 1.0, n)
    base = 3.0 * np.exp(-3.0 * x)
    cols = [base + 1e-3 * sum(rng.standard_normal() * np.sin((k + 1) * np.pi * x)
                              for k in range(cnt + 1)) + 1e-7 * rng.standard_normal(n)
            for _ in range(cnt + 1)]
    return cols[0], np.array(cols[1:])


@pytest.mark.parametrize("n", [257, 1025])
def test_fit_kernel_equals_restatement(n):
    rng = np.random.default_rng(n)
    with _engine(n) as eng:
        for cnt in (2, 3, 5, 6, 8):
            g_new, g_old = _columns(rng, cnt, n)
            if cnt == 5:
                g_old[2] = g_old[4]  # dependent column
            c, ratio, kept = eng.projection_fit(g_new, g_old)
            c_ref, ratio_ref, kept_ref = oproj.fit(g_new, g_old)
            print(f"n={n} cnt={cnt} kept={kept} ratio={ratio:.3e}", flush=True)
            np.testing.assert_array_equal(c, c_ref)
            assert ratio == ratio_ref and kept == kept_ref


def _run(n, tol, slots, steps=None, tmax=None):
    kw = {} if tmax is None else {"tmax": tmax}
    with _engine(n, tol=tol, **kw) as eng:
        if slots is not None:
            eng.projection_enable(slots)
        eng.rk43_setup()
        info = eng.rk43_step(steps if steps else 10 ** 7)
        st = eng.state()
        stats = eng.projection_stats()
    return info, st, stats


def test_disabled_projection_is_bitwise_plain():
    with _engine(129) as a, _engine(129) as b:
        b.projection_enable(6)
        b.projection_enable(0)
        for e in (a, b):
            e.rk43_setup()
            e.rk43_step(60)
        sa, sb = a.state(), b.state()
        assert sa["info"]["vcycles"] == sb["info"]["vcycles"]
        for k in ("slip", "psi", "V", "U"):
            np.testing.assert_array_equal(sa[k], sb[k])


def test_projection_tracks_plain_at_tight_tolerance():
    info0, st0, _ = _run(257, 1e-12, None, steps=250)
    T = info0["t"]
    info1, st1, stats = _run(257, 1e-12, 6, tmax=T)
    assert info1["t"] == T
    ds = np.max(np.abs(st1["slip"] - st0["slip"])) / np.max(np.abs(st0["slip"]))
    vc0 = info0["vcycles"] / info0["solves"]
    vc1 = info1["vcycles"] / info1["solves"]
    print(f"slip rel {ds:.3e}; V-cycles/solve plain {vc0:.3f} projection {vc1:.3f}; {stats}",
          flush=True)
    assert ds <= 3e-7
    assert vc1 < 0.5 * vc0
    assert stats["offered"] >= info1["solves"] - 1
    assert stats["used"] >= 0.9 * stats["offered"]


def test_projection_and_learned_guess_are_exclusive():
    g = lib.Guess(129, hidden=128, capacity=1)
    try:
        with _engine(129) as eng:
            eng.projection_enable(4)
            with pytest.raises(lib.QdeqError):
                eng.attach_guess(g, lib.GUESS_PREDICT)
            eng.projection_enable(0)
            eng.attach_guess(g, lib.GUESS_PREDICT)
            with pytest.raises(lib.QdeqError):
                eng.projection_enable(4)
            eng.attach_guess(None, lib.GUESS_OFF)
    finally:
        g.close()


@pytest.mark.parametrize("slots,precision", [(0, config.PREC_FP64), (6, config.PREC_FP64),
                                             (0, config.PREC_FP32_IR),
                                             (6, config.PREC_FP32_IR)])
def test_device_start_equals_host_start(slots, precision, monkeypatch):
    """Solves decide their start on the device (norms, candidate choice,
    stop test) and launch the first V-cycles without waiting; QDEQ_HOST_START=1
    restores the host decision.  Same norms, same choices: the runs are bit
    for bit equal.  FP32-IR solves with a projection start take the host
    path either way (the projection then puts its start where that path
    reads it); the starts are used."""
    runs = []
    for host in (False, True):
        if host:
            monkeypatch.setenv("QDEQ_HOST_START", "1")
        else:
            monkeypatch.delenv("QDEQ_HOST_START", raising=False)
        with _engine(129, tol=1e-10, precision=precision) as eng:
            if slots:
                eng.projection_enable(slots)
            eng.rk43_setup()
            eng.rk43_step(80)
            runs.append((eng.state(), eng.projection_stats()))
    (a, pa), (b, pb_) = runs
    assert a["info"]["vcycles"] == b["info"]["vcycles"]
    assert a["info"]["solves"] == b["info"]["solves"]
    assert pa == pb_
    if slots:
        assert pa["used"] > 0.5 * pa["offered"] > 0
    for k in ("slip", "psi", "V", "U"):
        assert np.all(np.isfinite(a[k])), k
        np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.parametrize("factor", [1e6, float("nan")])
@pytest.mark.parametrize("host", [False, True])
def test_bad_projection_start_falls_back_bit_for_bit(factor, host, monkeypatch):
    """A start made absurd or non-finite (fault injection) is never used:
    each solve falls back to the previous solution (restored over u's
    interior on the device path), and the run equals the plain run bit for
    bit."""
    if host:
        monkeypatch.setenv("QDEQ_HOST_START", "1")
    else:
        monkeypatch.delenv("QDEQ_HOST_START", raising=False)
    with _engine(129) as a, _engine(129) as b:
        b.projection_enable(6)
        b.projection_inject(factor)
        for e in (a, b):
            e.rk43_setup()
            e.rk43_step(40)
        sa, sb = a.state(), b.state()
        stats = b.projection_stats()
        assert stats["offered"] > 0 and stats["used"] == 0
        assert sa["info"]["vcycles"] == sb["info"]["vcycles"]
        for k in ("slip", "psi", "V", "U"):
            np.testing.assert_array_equal(sa[k], sb[k])
