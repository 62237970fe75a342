"""GPU checks of the learned initial guess (include/qdeq_guess.h).

* The BF16 MFMA forward pass, and the forward + backward pass (loss and every
  parameter gradient), agree with the CPU restatement oracle/nn.py, which
  rounds to bfloat16 at the same points (only the FP32 accumulation order
  differs), at a small and at the production shape (1025 fault nodes).
* The 16-row inference path of a context equals the training-path forward.
* A harvested pair is the linear response of the elliptic problem to the
  normalised fault-boundary change (checked with the CPU multigrid).
* A non-finite or absurd prediction is rejected: the run is then bit-for-bit
  the run without the network.
* Trained on pairs from perturbed runs, the network lowers the V-cycles per
  solve of an unseen run, and the slip at a common end time stays within
  1e-6 (relative) of the run without it.
"""
import numpy as np
import pytest

import oracle
from oracle import nn as onn
from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu

FAST = dict(cycle=config.CYCLE_FAST, n_pre=2, n_post=2, n_coarse=0,
            tol_mode=config.TOL_REL, tol=1e-10)


def _dims(g: lib.Guess) -> onn.Dims:
    c = g.cfg
    return onn.Dims(nk=g.nk, factor=c.factor, channels=c.channels, width=c.width,
                    hidden=c.hidden)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) /
                 max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("nk,hidden,batch", [(129, 128, 5), (1025, 512, 3)])
def test_forward_matches_oracle(nk, hidden, batch):
    g = lib.Guess(nk, hidden=hidden, capacity=1)
    d = _dims(g)
    flat = onn.init_params(d, seed=11)
    g.set_params(flat)
    np.testing.assert_array_equal(g.get_params(), flat)
    x = np.random.default_rng(12).standard_normal((batch, nk)).astype(np.float32)
    y = g.forward(x)
    ref = onn.forward(d, flat.astype(np.float64), x.astype(np.float64), bf16=True)
    assert y.shape == ref.shape
    assert _rel(y, ref) < 5e-3
    g.close()


@pytest.mark.parametrize("nk,hidden,batch", [(129, 128, 37), (257, 256, 130),
                                             (1025, 512, 64)])
def test_forward_backward_matches_oracle(nk, hidden, batch):
    g = lib.Guess(nk, hidden=hidden, capacity=1)
    d = _dims(g)
    rng = np.random.default_rng(nk)
    flat = onn.init_params(d, seed=nk)
    g.set_params(flat)
    x = rng.standard_normal((batch, nk)).astype(np.float32)
    t = (0.1 * rng.standard_normal((batch, d.nout))).astype(np.float32)
    loss, grad = g.forward_backward(x, t)
    rl, rg = onn.forward_backward(d, flat.astype(np.float64), x.astype(np.float64),
                                  t.astype(np.float64), bf16=True)
    assert abs(loss - rl) <= 1e-4 * rl
    o = 0
    for i, (r, c) in enumerate(d.shapes()):
        n = r * c
        a, b = grad[o:o + n].astype(np.float64), rg[o:o + n]
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
        assert _rel(a, b) < 2e-2 and cos > 0.9998, (i, _rel(a, b), cos)
        o += n
    g.close()


def _engine(n, levels, basin=True, **kw):
    pb = config.make_bp1_problem(n, basin=basin, levels=levels,
                                 tol_mode=config.TOL_REL, tol=1e-10, **kw)
    me = lib.build_mu_equi(pb.mu)
    return pb, me


def test_infer_path_equals_training_forward():
    pb, me = _engine(257, 5)
    g = lib.Guess(257, hidden=256, capacity=1)
    g.set_params(onn.init_params(_dims(g), seed=3))
    with lib.Engine(pb, me, **FAST) as eng:
        eng.set_tiles(True, 64, 16)
        eng.rk43_setup()
        eng.attach_guess(g, lib.GUESS_PREDICT)
        x = np.random.default_rng(4).standard_normal(257).astype(np.float32)
        y1 = eng.guess_infer(x)
        y2 = g.forward(x[None, :])[0]
        assert _rel(y1, y2.astype(np.float64)) < 2e-3
    g.close()


def test_harvested_pair_is_the_linear_response():
    pb, me = _engine(129, 4, basin=False)
    g = lib.Guess(129, hidden=128, capacity=4000)
    with lib.Engine(pb, me, **{**FAST, "tol": 1e-13}) as eng:
        eng.set_tiles(True, 64, 16)
        eng.rk43_setup()
        eng.attach_guess(g, lib.GUESS_HARVEST)
        eng.rk43_step(150)
        stats = eng.guess_stats()
    n = g.pairs
    assert n == stats["harvested"] > 100
    x, t, s = g.read_pairs(0, n)
    i = int(np.argmax(s))
    assert s[i] > 0
    # fault column x at every row, remote column 0, Neumann top/bottom: the
    # harvested target is the solution sampled at every 4th node
    p = pb.params
    U = np.zeros((p.nz, p.nx))
    U[1:-1, 1] = x[i].astype(np.float64)
    mg = oracle.MG(p.ni, p.nk, p.dh, p.n_levels, me)
    mg.set_cycle(2, 2, 0, True)
    st, vc, r0, r = mg.solve(U, oracle.ORC_TOL_REL, 1e-13, 200)
    assert st == 0
    ref = U[1:-1, 1:-1][::4, ::4]
    got = t[i].reshape(33, 33)
    np.testing.assert_array_equal(got[:, 0], x[i][::4])
    np.testing.assert_array_equal(got[:, -1], 0.0)
    assert np.max(np.abs(got - ref)) < 1e-5 * np.max(np.abs(ref))
    g.close()


@pytest.mark.parametrize("factor", [float("nan"), 1e6])
def test_bad_prediction_falls_back_bit_for_bit(factor):
    pb, me = _engine(129, 4)
    g = lib.Guess(129, hidden=128, capacity=1)
    g.set_params(onn.init_params(_dims(g), seed=5))
    runs = []
    for use in (False, True):
        with lib.Engine(pb, me, **FAST) as eng:
            eng.set_tiles(True, 64, 16)
            eng.rk43_setup()
            if use:
                eng.attach_guess(g, lib.GUESS_PREDICT)
                eng.guess_inject(factor)
            eng.rk43_step(40)
            st = eng.state()
            runs.append(st)
            if use:
                stats = eng.guess_stats()
    assert stats["predicted"] > 100 and stats["accepted"] == 0
    for k in ("slip", "psi", "V", "U"):
        np.testing.assert_array_equal(runs[0][k], runs[1][k])
    assert runs[0]["info"]["t"] == runs[1]["info"]["t"]
    assert runs[0]["info"]["vcycles"] == runs[1]["info"]["vcycles"]
    g.close()


def _run(pb, me, steps, guess=None, mode=lib.GUESS_OFF, tmax=None):
    """Runs `steps` accepted steps, or until t reaches tmax; returns the
    final state and the guess statistics."""
    kw = dict(FAST)
    if tmax is not None:
        kw["tmax"] = tmax
    with lib.Engine(pb, me, **kw) as eng:
        eng.set_tiles(True, 64, 16)
        eng.rk43_setup()
        if guess is not None:
            eng.attach_guess(guess, mode)
        if tmax is None:
            eng.rk43_step(steps)
        else:
            for _ in range(4 * steps):
                if eng.rk43_step(1)["t"] >= tmax:
                    break
        st = eng.state()
        stats = eng.guess_stats() if guess is not None else None
    return st, stats


def test_learned_guess_lowers_vcycles_per_solve():
    n, levels, steps = 257, 5, 300
    g = lib.Guess(n, hidden=512, capacity=20000)
    # harvest from runs with other frictional parameters than the test run
    for Dc, sn in ((0.006, 45.0), (0.011, 58.0)):
        pb_h, me_h = _engine(n, levels, Dc=Dc, sigma_n=sn)
        _run(pb_h, me_h, steps, g, lib.GUESS_HARVEST)
    assert g.pairs > 2000
    loss = g.train(epochs=40, batch=256, lr=1e-3, lr_final=2e-5)
    assert np.isfinite(loss)
    pb, me = _engine(n, levels)
    s0, _ = _run(pb, me, steps)
    T = s0["info"]["t"]
    s1, st = _run(pb, me, steps, g, lib.GUESS_PREDICT, tmax=T)
    i0, i1 = s0["info"], s1["info"]
    vps0 = i0["vcycles"] / i0["solves"]
    vps1 = i1["vcycles"] / i1["solves"]
    # both runs end exactly at T (the last step is clipped to tmax)
    assert i1["t"] == T
    dslip = np.max(np.abs(s1["slip"] - s0["slip"])) / np.max(np.abs(s0["slip"]))
    print(f"V-cycles/solve plain {vps0:.3f} learned {vps1:.3f}; accepted "
          f"{st['accepted']}/{st['predicted']}; loss {loss:.3e}; steps "
          f"{i0['it']}/{i1['it']}; slip rel diff at t={T:.6e}: {dslip:.2e}")
    assert vps1 < 0.9 * vps0
    assert st["accepted"] > 0.5 * st["predicted"]
    assert dslip <= 1e-6
    g.close()
