"""Verbose diagnostics of dp_parallel (dp_parallel_sampling.cpp:160-192, 218-236) on the device path:
the label x class contingency table and its NMI against the oracle's NMI of the same labels, and the
per-iteration log posterior against one recomputed from the oracle's statistics and log marginal
likelihoods of the labels the sampler holds."""
import math

import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture, multinomial_mixture
import oracle_ctypes as orc
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def test_contingency_and_nmi_match_oracle():
    rng = np.random.default_rng(4)
    N, K, C = 200_003, 7, 5
    gt = rng.integers(0, C, N).astype(np.int32)
    lab = ((gt * 3 + rng.integers(0, 3, N)) % K).astype(np.int32)   # labels that follow the classes in part
    X = np.zeros((N, 2), np.float32)
    say("contingency and NMI")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    dev.set_labels(lab, np.zeros(N, np.int32))
    dev.set_ground_truth(gt)
    t = dev.label_contingency(K, C)
    want = np.zeros((K, C), np.int64)
    np.add.at(want, (lab, gt), 1)
    np.testing.assert_array_equal(t, want)
    v = dev.label_nmi(K)
    o = orc.nmi(gt, lab)
    print(f"NMI device {v:.15f} oracle {o:.15f}")
    assert abs(v - o) <= 1e-12
    assert abs(v - dp._nmi(gt, lab)) <= 1e-12
    dev.set_labels(gt % K, None)   # labels equal to the classes: NMI 1
    assert abs(dev.label_nmi(K) - 1.0) <= 1e-12
    with pytest.raises(dp.DPMMError):
        dev.set_ground_truth(np.full(N, -1, np.int32))
    dev.close()


def _expected_log_posterior(alpha, N, Ns, Ls):
    v = math.lgamma(alpha) - math.lgamma(N + alpha)
    for n, L in zip(Ns, Ls):
        if n > 0:
            v += L + math.log(alpha) + math.lgamma(n)
    return v


def test_log_posterior_gaussian_matches_oracle():
    N, D, K = 40_000, 8, 6
    X, gt, _ = gaussian_mixture(N, D, K, seed=12, dtype=np.float64)
    say("log posterior gaussian")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=2)
    alpha, k0, v0 = 10.0, 1.0, D + 3.0
    smp = dp.Sampler(dev, dp.niw_hyperparams(k0, np.zeros(D), v0, np.eye(D)), alpha, 40, seed=2,
                     precision=dp.Precision.FP64)
    for t in range(1, 31):
        smp.step(t)
    lp = smp.log_posterior()
    Kc = smp.num_clusters()
    lab, sub = dev.get_labels()
    _, _, _, Nc, sc, Sc = orc.suff_stats(0, X, lab, sub, Kc)
    Ls = []
    for k in range(Kc):
        kn, mn, vn, psin = orc.niw_posterior(k0, np.zeros(D), v0, np.eye(D), Nc[k], sc[k], Sc[k])
        Ls.append(orc.niw_log_marginal(k0, v0, np.eye(D), kn, vn, psin, Nc[k]))
    want = _expected_log_posterior(alpha, N, Nc, Ls)
    print(f"K={Kc} log posterior {lp:.10f} oracle {want:.10f}")
    assert abs(lp - want) <= 1e-9 * abs(want)


def test_log_posterior_multinomial_matches_oracle():
    N, D, K = 20_000, 60, 5
    X, gt, _ = multinomial_mixture(N, D, K, seed=6)
    say("log posterior multinomial")
    dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=3)
    alpha = 10.0
    a0 = np.ones(D)
    smp = dp.Sampler(dev, dp.multinomial_hyper(a0), alpha, 40, seed=3, precision=dp.Precision.FP32)
    for t in range(1, 31):
        smp.step(t)
    lp = smp.log_posterior()
    Kc = smp.num_clusters()
    lab, sub = dev.get_labels()
    _, _, _, Nc, sc, _ = orc.suff_stats(1, X, lab, sub, Kc)
    Ls = [orc.mult_log_marginal(a0, a0 + sc[k]) for k in range(Kc)]
    want = _expected_log_posterior(alpha, N, Nc, Ls)
    print(f"K={Kc} log posterior {lp:.10f} oracle {want:.10f}")
    assert abs(lp - want) <= 1e-9 * abs(want)


def test_dp_parallel_verbose_histories():
    """dp_parallel in verbose mode with ground truth fills the log posterior and the device NMI
    every iteration (the reference's ModelInfo fields)."""
    X, gt, _ = gaussian_mixture(30_000, 2, 3, seed=0, dtype=np.float64)
    say("dp_parallel verbose histories")
    model = dp.dp_parallel_sampling_class(X.shape[0], X.T, 2, dp.prior_type.Gaussian)
    mi = model.dp_parallel(dp.niw_hyperparams(1.0, np.zeros(2), 5.0, np.eye(2)), 1.0, iters=40, verbose=True, gt=gt)
    assert len(mi.likelihood_history) == 40 and len(mi.nmi_score_history) == 40
    assert all(np.isfinite(mi.likelihood_history)) and all(v != 1.0 for v in mi.likelihood_history)
    assert abs(mi.nmi_score_history[-1] - dp._nmi(gt, mi.labels)) <= 1e-12


def test_hard_clustering_takes_the_argmax():
    """global_params::hard_clustering (local_clusters_actions.cpp:466): every label sweep takes the
    argmax.  After a step that changed no cluster, the labels are the argmax of the noise-free
    scores of the parameters that step drew (the soft sampler's are not, on overlapping clusters)."""
    N, D, K = 60_000, 8, 5
    X, _, _ = gaussian_mixture(N, D, K, seed=31, a=2.0, ridge=0.3, dtype=np.float32)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    say("hard clustering")
    mism = {}
    for hard in (True, False):
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=9)
        smp = dp.Sampler(dev, hyper, 10.0, 200, seed=9, precision=dp.Precision.FP32, hard_clustering=hard)
        kprev = 0
        for t in range(1, 61):
            smp.step(t)
            k = smp.num_clusters()
            if t >= 30 and k == kprev:
                lab, _ = dev.get_labels()
                s = dev.log_likelihoods(k, 0, 0, N)
                mism[hard] = int(np.count_nonzero(lab != np.argmax(s, axis=1)))
                break
            kprev = k
        smp.close()
        dev.close()
    print("labels off the argmax: hard", mism.get(True), "soft", mism.get(False))
    assert mism[True] == 0
