"""Batched FP64 conjugate work on the device (posteriors, Cholesky, log marginal likelihoods,
parameter draws, parameter-slot packing) against the FP64 oracle, and the reference's golden
split test (test/local_clusters_actions_test.cpp:174-279) run through the device.

Bars: log marginal likelihoods and log det psi_n within 1e-9 relative of the oracle; draws (same
Philox stream) within 1e-8 relative; slots packed on the device give labels identical to slots
packed on the host from the same draws."""
import json
import math
import os

import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
import oracle_ctypes as orc
from helpers import have_gpu, random_spd, say

pytestmark = pytest.mark.gpu
G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _data(D, N, K, seed, empty=3, dtype=np.float64):
    rng = np.random.default_rng(seed)
    centres = rng.normal(size=(K, D)) * 4
    labels = rng.integers(0, K, N).astype(np.int32)
    if empty is not None and empty < K:
        labels[labels == empty] = 0   # cluster `empty` has no points
    subs = rng.integers(0, 2, N).astype(np.int32)
    X = (centres[labels] + rng.normal(size=(N, D)) * (1 + rng.random(D))).astype(dtype)
    return X, labels, subs


def _jobs(K, rng):
    jobs = []
    for k in range(K):
        jobs += [[2 * k, 2 * k + 1, -1, -1], [2 * k, -1, -1, -1], [2 * k + 1, -1, -1, -1]]
    for i in range(K):
        for j in range(i + 1, K):
            jobs.append([2 * i, 2 * i + 1, 2 * j, 2 * j + 1])
    jobs.append([1, 4, 5, -1])
    jobs.append([-1, -1, -1, -1])   # empty sum: the prior
    return np.array(jobs, np.int32)


def _sum(jobs, Nsub, s, S):
    for m in jobs:
        idx = [q for q in m if q >= 0]
        yield (int(Nsub[idx].sum()) if idx else 0, s[idx].sum(0) if idx else np.zeros(s.shape[1]),
               S[idx].sum(0) if (idx and S is not None) else (np.zeros((s.shape[1],) * 2) if S is not None else None))


@pytest.mark.parametrize("D,dtype", [(2, np.float64), (3, np.float64), (16, np.float64), (64, np.float64),
                                     (100, np.float32), (128, np.float32)])
def test_log_marginal_batch_vs_oracle(D, dtype):
    K, N = 5, 4000
    X, labels, subs = _data(D, N, K, seed=D, dtype=dtype)
    rng = np.random.default_rng(D + 1)
    k0, v0 = 0.7, D + 3.0
    m0 = rng.normal(size=D)
    psi0 = random_spd(rng, D)
    say(f"log marginal batch D={D}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    dev.set_labels(labels, subs)
    dev.sufficient_statistics(K, counts_only=True)
    dev.set_prior(dp.niw_hyperparams(k0, m0, v0, psi0))
    jobs = _jobs(K, rng)
    L, ld = dev.log_marginal_batch(jobs)
    Nsub, s, S, _, _, _ = orc.suff_stats(0, X.astype(np.float64), labels, subs, K)
    for j, (n, su, SS) in enumerate(_sum(jobs, Nsub, s, S)):
        if n == 0:
            assert L[j] == 0.0
            np.testing.assert_allclose(ld[j], np.linalg.slogdet(psi0)[1], rtol=1e-12)
            continue
        kn, mn, vn, psin = orc.niw_posterior(k0, m0, v0, psi0, n, su, SS)
        want = orc.niw_log_marginal(k0, v0, psi0, kn, vn, psin, n)
        np.testing.assert_allclose(L[j], want, rtol=1e-9, err_msg=f"job {j} {jobs[j]}")
        np.testing.assert_allclose(ld[j], np.linalg.slogdet(psin)[1], rtol=1e-9, atol=1e-9)


def test_golden_split_through_device():
    """The reference's split test (test/local_clusters_actions_test.cpp:174-279): left / right
    sub-cluster statistics of its 10 points, log H = 0.52088 > log 0.5, so the cluster splits."""
    sp = G["lca_should_split"]
    _, l, r = sp["parts"]
    dev = dp.DeviceSweep(np.zeros((4, 2)), dp.prior_type.Gaussian, seed=0)
    dev.set_statistics([l["N"], r["N"]], np.array([l["sum"], r["sum"]]),
                       np.array([l["S"], r["S"]]).reshape(2, 2, 2))
    dev.set_prior(dp.niw_hyperparams(sp["k0"], sp["m0"], sp["v0"], np.array(sp["psi0"]).reshape(2, 2)))
    L, ld = dev.log_marginal_batch([[0, 1, -1, -1], [0, -1, -1, -1], [1, -1, -1, -1]])
    for j, p in enumerate(sp["parts"]):   # log det of the reference's posterior psi
        np.testing.assert_allclose(ld[j], np.linalg.slogdet(np.reshape(p["psin"], (2, 2)))[1], atol=1e-5)
    N, Nl, Nr = l["N"] + r["N"], l["N"], r["N"]
    lhr = (math.log(sp["alpha"]) + math.lgamma(Nl) + L[1] + math.lgamma(Nr) + L[2] - math.lgamma(N) - L[0])
    assert abs(lhr - sp["expected_log_hr"]) < 1e-4
    assert (lhr > math.log(sp["uniform"])) == bool(sp["expected_split"])


def test_golden_update_stats_through_device():
    """test/local_clusters_actions_test.cpp:15-172: statistics of the reference's 10 points computed on
    the device, posteriors' log det psi_n against the reference's posterior psi of cluster, l, r."""
    e = G["lca_update_stats"]
    X = np.array(e["points_DxN"], np.float64).reshape(e["D"], e["N"]).T.copy()   # reference D x N fill
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=0)
    dev.set_labels(np.array(e["labels"], np.int32) - 1, np.array(e["sub_labels"], np.int32) - 1)
    Nsub, _, _ = dev.sufficient_statistics(1, counts_only=True)
    dev.set_prior(dp.niw_hyperparams(e["k0"], e["m0"], e["v0"], np.array(e["psi0"]).reshape(2, 2)))
    _, ld = dev.log_marginal_batch([[0, 1, -1, -1], [0, -1, -1, -1], [1, -1, -1, -1]])
    exp = np.array([x[0] for x in e["expected"]]).reshape(3, 15)   # N, sum(2), S(4), k, m(2), v, psi(4)
    assert list(Nsub) == [int(exp[1, 0]), int(exp[2, 0])]
    for j in range(3):
        np.testing.assert_allclose(ld[j], np.linalg.slogdet(exp[j, 11:15].reshape(2, 2))[1], atol=2e-4)


def test_log_marginal_batch_multinomial():
    rng = np.random.default_rng(4)
    N, D, K = 3000, 50, 4
    theta = rng.dirichlet(np.full(D, 0.3), K)
    labels = rng.integers(0, K, N).astype(np.int32)
    X = np.stack([rng.multinomial(40, theta[k]) for k in labels]).astype(np.int32)
    subs = rng.integers(0, 2, N).astype(np.int32)
    a0 = rng.uniform(0.2, 2.0, D)
    dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=3)
    dev.set_labels(labels, subs)
    dev.sufficient_statistics(K, counts_only=True)
    dev.set_prior(dp.multinomial_hyper(a0))
    jobs = _jobs(K, rng)
    say("log marginal batch multinomial D=50")
    L, _ = dev.log_marginal_batch(jobs)
    Nsub, s, _, _, _, _ = orc.suff_stats(1, X, labels, subs, K, want_S=False)
    for j, (n, su, _) in enumerate(_sum(jobs, Nsub, s, None)):
        want = orc.mult_log_marginal(a0, a0 + su)
        np.testing.assert_allclose(L[j], want, rtol=1e-10, atol=1e-9)


def test_member_outside_statistics_rejected():
    X, labels, subs = _data(4, 500, 3, seed=2)
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    dev.set_labels(labels, subs)
    dev.sufficient_statistics(3, counts_only=True)
    dev.set_prior(dp.niw_hyperparams(1.0, np.zeros(4), 7.0, np.eye(4)))
    with pytest.raises(dp.DPMMError):
        dev.log_marginal_batch([[6, -1, -1, -1]])


@pytest.mark.parametrize("D,full", [(2, 0), (16, 0), (64, 0), (64, 1), (128, 0)])
def test_draws_vs_oracle(D, full):
    K, N, it = 4, 3000, 7
    dtype = np.float64 if D <= 64 else np.float32
    X, labels, subs = _data(D, N, K, seed=10 + D, empty=2, dtype=dtype)
    rng = np.random.default_rng(D)
    k0, v0 = 1.0, D + 3.0
    m0 = rng.normal(size=D)
    psi0 = np.eye(D) * 0.5 + random_spd(rng, D, scale=0.1)
    say(f"draws D={D} full_mu_cov={full}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=11)
    dev.set_labels(labels, subs)
    dev.sufficient_statistics(K, counts_only=True)
    dev.set_prior(dp.niw_hyperparams(k0, m0, v0, psi0))
    mem = np.array([[2 * k, 2 * k + 1, -1, -1] if p == 0 else [2 * k + p - 1, -1, -1, -1]
                    for k in range(K) for p in range(3)], np.int32)
    w = rng.dirichlet(np.ones(K))
    lrw = rng.dirichlet(np.ones(2), K).ravel()
    dev.sample_params(it, mem, w, lrw, full_mu_cov=full)
    mu, R, ld = dev.drawn_params(K)
    Nsub, s, S, _, _, _ = orc.suff_stats(0, X.astype(np.float64), labels, subs, K)
    for p, (n, su, SS) in enumerate(_sum(mem, Nsub, s, S)):
        kn, mn, vn, psin = orc.niw_posterior(k0, m0, v0, psi0, n, su, SS) if n else (k0, m0, v0, psi0)
        omu, olam, old = orc.niw_draw(kn, mn, vn, psin, 11, p, it, full_mu_cov=bool(full))
        lam = R[p] @ R[p].T
        assert np.allclose(np.triu(R[p], 1), 0.0)
        np.testing.assert_allclose(lam, olam, rtol=1e-8, atol=1e-8 * np.abs(olam).max(), err_msg=f"part {p}")
        np.testing.assert_allclose(ld[p], old, rtol=1e-9, atol=1e-8)
        np.testing.assert_allclose(mu[p], omu, rtol=1e-8, atol=1e-8 * (1 + np.abs(omu).max()))


def test_multinomial_draws_vs_oracle_stream():
    rng = np.random.default_rng(8)
    N, D, K = 2000, 30, 3
    labels = rng.integers(0, K, N).astype(np.int32)
    X = rng.poisson(2.0, size=(N, D)).astype(np.int32)
    subs = rng.integers(0, 2, N).astype(np.int32)
    a0 = np.full(D, 0.5)
    dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=21)
    dev.set_labels(labels, subs)
    dev.sufficient_statistics(K, counts_only=True)
    dev.set_prior(dp.multinomial_hyper(a0))
    mem = np.array([[2 * k, 2 * k + 1, -1, -1] if p == 0 else [2 * k + p - 1, -1, -1, -1]
                    for k in range(K) for p in range(3)], np.int32)
    dev.sample_params(5, mem, np.full(K, 1.0 / K), np.full(2 * K, 0.5))
    lt, _, _ = dev.drawn_params(K)
    Nsub, s, _, _, _, _ = orc.suff_stats(1, X, labels, subs, K, want_S=False)
    for p, (n, su, _) in enumerate(_sum(mem, Nsub, s, None)):
        want = np.log(orc.stream("dirichlet_substreams", 21, p, 5, 6, D, alpha=a0 + su))
        np.testing.assert_allclose(lt[p], want, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("prec", ["FP32", "FP64", "MIXED"])
def test_device_packing_matches_host_packing(prec):
    """Slots written on the device from the draws sweep exactly like slots the host packs from
    the same draws (labels and sub-labels identical)."""
    D, K, N, it = 64, 6, 20000, 3
    X, labels, subs = _data(D, N, K, seed=5, empty=None, dtype=np.float64 if prec == "FP64" else np.float32)
    rng = np.random.default_rng(2)
    say(f"device packing {prec}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=4)
    dev.set_precision(getattr(dp.Precision, prec))
    dev.set_labels(labels, subs)
    dev.sufficient_statistics(K, counts_only=True)
    dev.set_prior(dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D)))
    mem = np.array([[2 * k, 2 * k + 1, -1, -1] if p == 0 else [2 * k + p - 1, -1, -1, -1]
                    for k in range(K) for p in range(3)], np.int32)
    w = rng.dirichlet(np.ones(K))
    lrw = rng.dirichlet(np.ones(2), K).ravel()
    dev.sample_params(it, mem, w, lrw)
    mu, R, ld = dev.drawn_params(K)
    out = []
    for host in (False, True):
        if host:
            c, sub = np.arange(K) * 3, np.sort(np.r_[np.arange(K) * 3 + 1, np.arange(K) * 3 + 2])
            dev.set_gaussian_params(mu[c], R[c], ld[c], w, mu[sub], R[sub], ld[sub], lrw, factors=True)
        dev.set_labels(labels, subs)
        dev.sample_cluster_labels(it, False)
        dev.sample_sub_labels(it)
        out.append(dev.get_labels())
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
