"""Pins the CPU restatement of the projection start's least-squares fit
(oracle/proj.py) against numpy's SVD least squares, exact-span and
dependent-column cases, and the fixed summation order on sums whose exact
value is known; and the start field against the linearity it relies on."""
import numpy as np
import pytest

from oracle import proj


def _columns(rng, cnt, n, noise=1e-3):
    """Fault columns shaped like stage data: a smooth slip profile advanced
    by smooth rate profiles, plus a little rough noise."""
    x = np.linspace(0.0, 1.0, n)
    base = 3.0 * np.exp(-3.0 * x) + 0.5 * x
    rates = [np.sin((k + 1) * np.pi * x) * np.exp(-x) for k in range(cnt + 1)]
    cols = []
    for i in range(cnt + 1):
        w = rng.standard_normal(cnt + 1)
        cols.append(base + 1e-3 * sum(wk * r for wk, r in zip(w, rates))
                    + noise * 1e-3 * rng.standard_normal(n))
    return cols[0], np.array(cols[1:])


@pytest.mark.parametrize("cnt", [2, 3, 4, 6, 8])
@pytest.mark.parametrize("n", [257, 1025])
def test_fit_matches_svd_least_squares(cnt, n):
    rng = np.random.default_rng(100 * cnt + n)
    g_new, g_old = _columns(rng, cnt, n)
    c, ratio, kept = proj.fit(g_new, g_old)
    ref = proj.lstsq_reference(g_new, g_old)
    assert kept == cnt - 1
    np.testing.assert_allclose(c, ref, rtol=1e-7, atol=1e-9 * np.max(np.abs(ref)))
    d = g_new - g_old[0]
    mis = d - (g_old[1:] - g_old[0]).T @ ref[1:]
    assert ratio == pytest.approx(np.linalg.norm(mis) / np.linalg.norm(d), rel=1e-6)


def test_exact_span_is_reproduced():
    rng = np.random.default_rng(7)
    _, g_old = _columns(rng, 5, 1025, noise=1.0)
    w = np.array([0.0, 0.7, -0.2, 1.3, 0.05])
    g_new = g_old[0] + (g_old[1:] - g_old[0]).T @ w[1:]
    c, ratio, kept = proj.fit(g_new, g_old)
    assert kept == 4
    np.testing.assert_allclose(c, w, atol=1e-8)
    # the misfit norm comes from a Cholesky pivot: about sqrt(eps) of ||d||
    # when d lies in the span
    assert ratio < 1e-6


def test_dependent_column_is_left_out():
    rng = np.random.default_rng(8)
    g_new, g_old = _columns(rng, 4, 513)
    g_old[3] = g_old[1]  # the same stored data twice
    c, ratio, kept = proj.fit(g_new, g_old)
    assert kept == 2
    assert c[3] == 0.0
    ref = proj.lstsq_reference(g_new, np.delete(g_old, 3, axis=0))
    np.testing.assert_allclose(np.delete(c, 3), ref, rtol=1e-7,
                               atol=1e-9 * np.max(np.abs(ref)))


def test_no_change_gives_zero_start_change():
    rng = np.random.default_rng(9)
    _, g_old = _columns(rng, 3, 257)
    c, ratio, kept = proj.fit(g_old[0].copy(), g_old)
    np.testing.assert_array_equal(c, np.zeros(3))
    assert ratio == 0.0


def test_block_sum_order_on_exact_sums():
    # integers (exact in float64): every order gives the same value
    v = np.arange(1, 1026, dtype=np.float64)
    assert proj.block_sum(v) == 1025 * 1026 / 2
    # one 1.0 and 1023 terms of 1e-16: summed left to right every small term
    # is lost against 1.0 (the result is exactly 1.0); in the kernel's order
    # threads 1..255 first sum their four small terms, and the pairwise tree
    # adds those sums to each other before they meet the 1.0, so the result
    # is 1 + 1020e-16 to within a few units in the last place
    v = np.full(1024, 1e-16)
    v[0] = 1.0
    left_to_right = 0.0
    for x in v:
        left_to_right += x
    assert left_to_right == 1.0
    s = proj.block_sum(v)
    assert abs(s - (1.0 + 1020e-16)) <= 4 * 2.0 ** -52


def test_start_field_is_exact_for_linear_responses():
    # fields that depend linearly on their fault column and remote value
    # reproduce the new field exactly when the new column is in the stored
    # span, whatever the remote values
    rng = np.random.default_rng(11)
    N, K = 65, 33
    prof = np.exp(-np.linspace(0.0, 3.0, K))  # response profile across the grid
    wr = np.tile(np.linspace(0.0, 1.0, K), (N, 1))  # unit remote response
    _, g_old = _columns(rng, 4, N, noise=1.0)
    r_old = np.array([2.0, 1.9, 1.7, 1.6])
    w = np.array([0.0, 0.3, -0.1, 0.6])
    g_new = g_old[0] + (g_old[1:] - g_old[0]).T @ w[1:]
    ring = np.array([np.outer(g, prof) + r * wr for g, r in zip(g_old, r_old)])
    r_new = 2.5
    u = np.zeros((N, K))
    u[:, 0] = g_new
    u[:, K - 1] = r_new
    c, _, _ = proj.fit(g_new, g_old)
    start = proj.apply(ring, r_old, wr, u, c)
    exact = np.outer(g_new, prof) + r_new * wr
    exact[:, 0] = g_new
    exact[:, K - 1] = r_new
    np.testing.assert_allclose(start, exact, rtol=0, atol=1e-12)
