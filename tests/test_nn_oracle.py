"""Pins the CPU restatement of the learned-guess network (oracle/nn.py):
bfloat16 rounding against torch's conversion, the encoder's tap positions on
impulses, the gradient formulas against central finite differences, and
positive homogeneity of the bias-free ReLU network."""
import numpy as np
import pytest

from oracle import nn as onn


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([
        rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
        # exact ties between two bfloat16 neighbours (round half to even)
        (np.arange(1, 2000, dtype=np.float64) * 2.0 ** -7 + 2.0 ** -9),
        np.array([0.0, -0.0, 1.0, -1.0, 3.0e38, 1e-38]),
    ]).astype(np.float32)
    ours = onn.bf16_round(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(ours, ref)


def test_bf16_round_is_idempotent_and_monotone():
    rng = np.random.default_rng(1)
    a = np.sort(rng.standard_normal(5000).astype(np.float32))
    r = onn.bf16_round(a)
    np.testing.assert_array_equal(onn.bf16_round(r), r)
    assert np.all(np.diff(r) >= 0)
    assert np.max(np.abs(r - a) / np.abs(a)) <= 2.0 ** -8 * 1.0000001  # 8-bit significand


def test_conv_taps_follow_stride_and_padding():
    d = onn.Dims(nk=33, factor=4, channels=1, width=8, hidden=128)
    q = onn.conv_index(d)
    # position p reads nodes 4p-2 .. 4p+5, clipped to [0, nk)
    assert q.shape == (8, 8)
    np.testing.assert_array_equal(q[0], [-1, -1, 0, 1, 2, 3, 4, 5])
    np.testing.assert_array_equal(q[7], [26, 27, 28, 29, 30, 31, 32, -1])
    # an impulse at node n feeds output p with tap weight w[n - 4p + 2]
    flat = onn.pack([np.arange(1, 9, dtype=float)[None, :],
                     np.zeros((128, 8)), np.zeros((128, 128)),
                     np.zeros((128, 128)), np.zeros((d.nout, 128))])
    for n in (0, 5, 13, 32):
        x = np.zeros((1, 33))
        x[0, n] = 1.0
        _, c = onn.forward(d, flat, x, bf16=False, cache=True)
        h0 = c["H"][0][0]
        for p in range(8):
            t = n - 4 * p + 2
            expect = float(t + 1) if 0 <= t < 8 else 0.0
            assert h0[p] == expect


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    d = onn.Dims(nk=33, factor=4, channels=3, width=6, hidden=7)
    rng = np.random.default_rng(seed)
    flat = onn.init_params(d, seed=seed).astype(np.float64)
    x = rng.standard_normal((5, d.nk))
    t = rng.standard_normal((5, d.nout)) * 0.1
    loss, g = onn.forward_backward(d, flat, x, t, bf16=False)
    eps = 1e-6
    idx = rng.choice(flat.size, size=120, replace=False)
    # every tensor contributes checked entries
    starts = np.cumsum([0] + [r * c for r, c in d.shapes()])
    for a, b in zip(starts[:-1], starts[1:]):
        idx = np.concatenate([idx, rng.integers(a, b, size=6)])
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += eps
        fm[i] -= eps
        lp = onn.forward_backward(d, fp, x, t, bf16=False)[0]
        lm = onn.forward_backward(d, fm, x, t, bf16=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_network_is_positively_homogeneous():
    d = onn.Dims(nk=65, factor=4, channels=4, width=8, hidden=16)
    flat = onn.init_params(d, seed=3).astype(np.float64)
    x = np.random.default_rng(4).standard_normal((3, d.nk))
    y1 = onn.forward(d, flat, x, bf16=False)
    y2 = onn.forward(d, flat, 3.5 * x, bf16=False)
    np.testing.assert_allclose(y2, 3.5 * y1, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(onn.forward(d, flat, 0 * x, bf16=False), 0.0)


def test_bf16_forward_stays_close_to_exact():
    d = onn.Dims(nk=129, factor=4, channels=8, width=8, hidden=128)
    flat = onn.init_params(d, seed=5).astype(np.float64)
    x = np.random.default_rng(6).standard_normal((4, d.nk))
    y64 = onn.forward(d, flat, x, bf16=False)
    y16 = onn.forward(d, flat, x, bf16=True)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert rel < 2e-2
