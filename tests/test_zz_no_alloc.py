"""The mu_equi build that allocates nothing (qdeq_build_mu_equi_noalloc):
through the C ABI, on device arrays the test allocates, it gives the same
array as qdeq_build_mu_equi, bit for bit, on two calls in a row, the second
with a smaller grid (the scratch tables are recomputed by every call).
Also: grids outside the supported range and a null pointer are refused
before anything runs."""
import numpy as np
import pytest

from qdeqcycle_fdm_amd import config, lib

pytestmark = pytest.mark.gpu


def _random_medium(n, seed):
    rng = np.random.default_rng(seed)
    mu = np.zeros((n + 2, n + 2))
    mu[1:-1, 1:-1] = np.exp(rng.normal(np.log(25.0), 0.4, size=(n, n)))
    # ghost entries are never read: garbage there must not matter
    mu[0, :] = mu[-1, :] = mu[:, 0] = mu[:, -1] = np.nan
    return mu


def _noalloc(mu):
    L = lib.load()
    nz, nx = mu.shape
    d_mu = lib.DeviceArray.from_host(mu)
    d_me = lib.DeviceArray((2 * nz - 1, 2 * nx - 1))
    d_me.copy_from(np.full(d_me.shape, 7.0))   # stale contents are overwritten
    st = L.qdeq_build_mu_equi_noalloc(nx - 2, nz - 2, d_mu.ptr, d_me.ptr, 0)
    assert st == lib.QDEQ_OK, L.qdeq_global_error().decode()
    out = d_me.to_host()
    d_mu.free()
    d_me.free()
    return out


def test_noalloc_mu_equi_equals_the_allocating_build():
    cases = [config.make_bp1_problem(257, basin=True).mu, _random_medium(65, 3)]
    for mu in cases:  # the second call has the smaller grid
        ref = lib.build_mu_equi(np.nan_to_num(mu))
        got = _noalloc(mu)
        np.testing.assert_array_equal(got, ref)


def test_noalloc_mu_equi_rejects_bad_sizes():
    L = lib.load()
    d = lib.DeviceArray((16,))
    for ni, nk in ((2, 65), (65, 2), (70000, 65)):
        assert L.qdeq_build_mu_equi_noalloc(ni, nk, d.ptr, d.ptr, 0) == lib.QDEQ_ERR_ARG
        assert "grid" in L.qdeq_global_error().decode()
    assert L.qdeq_build_mu_equi_noalloc(65, 65, None, d.ptr, 0) == lib.QDEQ_ERR_ARG
    d.free()
