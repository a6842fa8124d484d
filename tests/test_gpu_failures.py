"""Failure handling through the C ABI.  The reference's device layer calls exit() on any CUDA or
cuBLAS error (cudaKernel.cu:1492-1508); here every entry point returns a code and leaves the context
usable: non-SPD or non-finite parameters, priors and posteriors, bad relabel tables, point ranges
outside the data, and device allocations that fail (forced from the host with
dpmm_debug_fail_allocations, so no kernel ever runs on a missing buffer)."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture
from helpers import gaussian_params, have_gpu, say

pytestmark = pytest.mark.gpu
EINVAL, ENOMEM, ENOTSPD, ESTATE = -1, -3, -4, -5


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")
    yield
    dp.load_library().dpmm_debug_fail_allocations(-1)


def _code(fn, *a, **k):
    with pytest.raises(dp.DPMMError) as e:
        fn(*a, **k)
    return e.value.code


def _setup(D=4, K=3, N=5000, dtype=np.float32):
    X, gt, _ = gaussian_mixture(N, D, K, seed=D, dtype=dtype)
    rng = np.random.default_rng(1)
    mu, lam, ld = gaussian_params(rng, X, K)
    smu, slam, sld = gaussian_params(rng, X, K, sub=True)
    return X, (mu, lam, ld, np.full(K, 1.0 / K), smu, slam, sld, np.full(2 * K, 0.5))


def test_bad_priors():
    X, _ = _setup()
    D = X.shape[1]
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    say("non-SPD prior psi")
    psi = np.eye(D)
    psi[0, 0] = -1.0
    assert _code(dev.set_prior, dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, psi)) == ENOTSPD
    assert _code(dev.set_prior, dp.niw_hyperparams(1.0, np.zeros(D), D - 1.5, np.eye(D))) == EINVAL   # nu <= D - 1
    assert _code(dev.set_prior, dp.niw_hyperparams(0.0, np.zeros(D), D + 3.0, np.eye(D))) == EINVAL   # kappa 0
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, psi)
    assert _code(dp.Sampler, dev, hyper, 1.0, 10) == ENOTSPD
    dev.set_prior(dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D)))   # still usable
    dev.close()


def test_bad_parameters():
    X, params = _setup()
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    say("non-SPD / NaN / negative cluster parameters")
    mu, lam, ld, w, smu, slam, sld, lrw = params
    bad = lam.copy()
    bad[1] = -bad[1]
    assert _code(dev.set_gaussian_params, mu, bad, ld, w, smu, slam, sld, lrw) == ENOTSPD
    nan_mu = mu.copy()
    nan_mu[0, 0] = np.nan
    assert _code(dev.set_gaussian_params, nan_mu, lam, ld, w, smu, slam, sld, lrw) == EINVAL
    nan_lam = lam.copy()
    nan_lam[2, 0, 0] = np.nan
    assert _code(dev.set_gaussian_params, mu, nan_lam, ld, w, smu, slam, sld, lrw) == ENOTSPD
    for wbad in (np.array([0.5, -0.1, 0.6]), np.array([0.5, np.nan, 0.5]), np.array([0.5, np.inf, 0.5])):
        assert _code(dev.set_gaussian_params, mu, lam, ld, wbad, smu, slam, sld, lrw) == EINVAL
    assert _code(dev.sample_cluster_labels, 1) == ESTATE   # no parameters were accepted
    dev.set_gaussian_params(*params)   # the context is unchanged and usable
    dev.sample_cluster_labels(1)
    lab, _ = dev.get_labels()
    assert lab.min() >= 0 and lab.max() < 3
    say("bad relabel tables and point ranges")
    assert _code(dev.relabel, 1, np.zeros(2 * 3000, np.int32), np.zeros(3000, np.int32)) == EINVAL   # K above 2048
    assert _code(dev.log_likelihoods, 3, 0, X.shape[0] - 5, 10) == EINVAL
    assert _code(dev.log_likelihoods, 4, 0, 0, 10) == EINVAL   # K differs from the parameters'
    dev.close()


# D = 4 and 64 take the one-row register factorisations, D = 128 the two-row ones; "mid" puts the one
# negative pivot in the middle (a later panel, another wave's), also for the reversed factorisation
@pytest.mark.parametrize("D,where", [(4, "all"), (4, "mid"), (64, "mid"), (128, "all"), (128, "mid")])
def test_non_spd_posterior_gives_nan_and_the_sampler_reports_it(D, where):
    X, _ = _setup(D=D)
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    dev.set_prior(dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D)))
    say(f"statistics whose posterior is not SPD, D={D} {where}")
    Nsub = np.array([10, 10], np.int64)
    sums = np.zeros((2, D))
    bad = -1e6 * np.eye(D) if where == "all" else np.diag(np.where(np.arange(D) == D // 2, -1e6, 1.0))
    S = np.stack([bad, np.eye(D)])
    dev.set_statistics(Nsub, sums, S)
    L, ldet = dev.log_marginal_batch(np.array([[0, -1, -1, -1], [1, -1, -1, -1], [0, 1, -1, -1]]))
    assert np.isnan(L[0]) and np.isfinite(L[1]) and np.isnan(L[2])
    assert _code(dev.sample_params, 1, np.array([[0, 1, -1, -1], [0, -1, -1, -1], [1, -1, -1, -1]]), [1.0],
                 [0.5, 0.5]) == ENOTSPD
    dev.close()


def test_allocation_failures_are_reported_and_recovered():
    L = dp.load_library()
    X, params = _setup(N=20000)
    D = X.shape[1]
    say("allocation failure at context creation")
    L.dpmm_debug_fail_allocations(0)
    assert _code(dp.DeviceSweep, X, dp.prior_type.Gaussian, 1) == ENOMEM
    L.dpmm_debug_fail_allocations(2)   # X and the labels succeed, the sub-labels do not
    assert _code(dp.DeviceSweep, X, dp.prior_type.Gaussian, 1) == ENOMEM
    L.dpmm_debug_fail_allocations(-1)
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    for after in range(0, 12):   # a failure at every allocation the sampler's set-up makes
        say(f"allocation failure after {after} allocations in the sampler set-up")
        L.dpmm_debug_fail_allocations(after)
        try:
            s = dp.Sampler(dev, hyper, 10.0, 20, seed=3, precision=dp.Precision.MIXED)
            s.step(1)
            s.close()
        except dp.DPMMError as e:
            assert e.code in (ENOMEM, -2), e
        L.dpmm_debug_fail_allocations(-1)
    say("the context after the failures")
    s = dp.Sampler(dev, hyper, 10.0, 30, seed=3, precision=dp.Precision.MIXED)
    for t in range(1, 31):
        s.step(t)
    assert s.num_clusters() >= 2
    s.close()
    dev.close()
