"""Parity of the HIP device layer (through the C ABI) with the FP64 oracle, on the same seeded
inputs and the same Philox noise.  Bar: labels and sub-labels identical wherever the oracle's
winning margin exceeds 1e-6 relative (FP32 sweep) or 1e-12 (FP64 sweep); counts exact; FP64
statistics within 1e-9 relative."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture, multinomial_mixture
import oracle_ctypes as orc
from helpers import gaussian_params, have_gpu, say

pytestmark = pytest.mark.gpu

FP32_MARGIN = 1e-6
FP64_MARGIN = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _check_labels(lab, olab, margin, tol, what):
    mism = np.nonzero(lab != olab)[0]
    rate = mism.size / lab.size
    print(f"{what}: {mism.size} mismatches of {lab.size} (rate {rate:.2e})")
    if mism.size:
        assert margin[mism].max() < tol, f"{what}: mismatch at margin {margin[mism].max():.3e}"
    return rate


@pytest.mark.parametrize("D,K,dtype", [(2, 3, np.float64), (3, 5, np.float64), (2, 4, np.float32),
                                       (5, 7, np.float32), (32, 13, np.float32), (50, 6, np.float32),
                                       (64, 9, np.float32), (100, 4, np.float32), (128, 3, np.float32)])
@pytest.mark.parametrize("final", [False, True])
def test_label_sweep_gaussian(D, K, dtype, final):
    rng = np.random.default_rng(D * 100 + K)
    N = 20011
    X, gt, _ = gaussian_mixture(N, D, max(2, K // 2), seed=D + K, dtype=dtype)
    mu, lam, ld = gaussian_params(rng, X, K)
    smu, slam, sld = gaussian_params(rng, X, K, sub=True)
    w = rng.dirichlet(np.ones(K))
    lrw = rng.dirichlet([1.0, 1.0], size=K).ravel()
    say(f"label sweep D={D} K={K} dtype={np.dtype(dtype).name} final={final}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=77)
    dev.set_gaussian_params(mu, lam, ld, w, smu, slam, sld, lrw)
    dev.sample_cluster_labels(5, final)
    lab, _ = dev.get_labels()
    olab, margin = orc.label_sweep(0, X, mu, lam, ld, weights=w, seed=77, it=5, final=final)
    tol = FP64_MARGIN if dtype == np.float64 else FP32_MARGIN
    _check_labels(lab, olab, margin, tol, f"labels D={D} K={K}")
    # sub-labels on the oracle's labels
    say(f"sub-label sweep D={D} K={K}")
    dev.set_labels(olab, np.zeros(N, np.int32))
    dev.sample_sub_labels(5)
    _, sub = dev.get_labels()
    osub, smargin = orc.sublabel_sweep(0, X, olab, smu, slam, sld, lr_weights=lrw, seed=77, it=5)
    _check_labels(sub, osub, smargin, tol, f"sub-labels D={D} K={K}")
    dev.close()


def test_label_sweep_realistic_margins_fp32():
    """Mixture-fitted parameters (posterior-like): the sweep must reproduce the oracle's labels."""
    rng = np.random.default_rng(3)
    D, K, N = 32, 10, 200_000
    X, gt, means = gaussian_mixture(N, D, K, seed=4, dtype=np.float32)
    mu = np.stack([X[gt == k].mean(0) for k in range(K)]).astype(np.float64)
    lam = np.stack([np.linalg.inv(np.cov(X[gt == k].T.astype(np.float64))) for k in range(K)])
    ld = np.array([-np.linalg.slogdet(l)[1] for l in lam])
    w = np.bincount(gt, minlength=K) / N
    sub_mu = np.repeat(mu, 2, axis=0) + rng.normal(size=(2 * K, D)) * 0.5
    sub_lam = np.repeat(lam, 2, axis=0)
    sub_ld = np.repeat(ld, 2)
    lrw = np.full(2 * K, 0.5)
    say("realistic label sweep D=32 K=10 N=2e5")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=5)
    dev.set_gaussian_params(mu, lam, ld, w, sub_mu, sub_lam, sub_ld, lrw)
    dev.sample_cluster_labels(11)
    dev.sample_sub_labels(11)
    lab, sub = dev.get_labels()
    olab, margin = orc.label_sweep(0, X, mu, lam, ld, weights=w, seed=5, it=11)
    _check_labels(lab, olab, margin, FP32_MARGIN, "labels")
    osub, smargin = orc.sublabel_sweep(0, X, olab, sub_mu, sub_lam, sub_ld, lr_weights=lrw, seed=5, it=11)
    same = lab == olab
    _check_labels(sub[same], osub[same], smargin[same], FP32_MARGIN, "sub-labels")


@pytest.mark.parametrize("D,K,dtype", [(2, 1, np.float64), (3, 4, np.float64), (7, 5, np.float32),
                                       (16, 4, np.float32), (17, 3, np.float32), (32, 9, np.float32),
                                       (40, 3, np.float64), (48, 2, np.float32), (50, 5, np.float32),
                                       (64, 3, np.float32), (64, 2, np.float64), (96, 2, np.float32),
                                       (128, 2, np.float32)])
def test_sufficient_statistics(D, K, dtype):
    """Padded widths 16 .. 64 take the wave-local MFMA kernel (vector row loads when D is a multiple
    of DP / 16, element loads otherwise), 96 and 128 the LDS-staged one."""
    rng = np.random.default_rng(D + 7 * K)
    N = 30007
    X = (rng.normal(size=(N, D)) * 20 + rng.uniform(-100, 100, size=D)).astype(dtype)
    labels = rng.integers(0, K, N).astype(np.int32)
    if K > 2:
        labels[labels == 1] = 0   # an empty cluster
    subs = rng.integers(0, 2, N).astype(np.int32)
    say(f"stats D={D} K={K}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    dev.set_labels(labels, subs)
    Nsub, s, S = dev.sufficient_statistics(K)
    oN, os_, oS, _, _, _ = orc.suff_stats(0, X, labels, subs, K)
    np.testing.assert_array_equal(Nsub, oN)
    scale_s = np.abs(X).astype(np.float64).sum() + 1
    np.testing.assert_allclose(s, os_, rtol=1e-9, atol=1e-12 * scale_s)
    scale_S = (X.astype(np.float64) ** 2).sum() + 1
    np.testing.assert_allclose(S, oS, rtol=1e-9, atol=1e-12 * scale_S)
    # deterministic: a second call is bitwise identical
    N2, s2, S2 = dev.sufficient_statistics(K)
    assert np.array_equal(s, s2) and np.array_equal(S, S2)


@pytest.mark.parametrize("prior", ["gauss", "mult"])
def test_statistics_after_resets_match_full_pass(prior):
    """A relabel that only redraws the sub-labels of some clusters (a reset) leaves the chunk
    partials of the other clusters valid: the next statistics pass recomputes the reset clusters
    only.  Two resets in a row, then a split, each against the oracle's statistics of the labels
    the device then holds (exact counts and integer sums, FP64 sums within 1e-9)."""
    rng = np.random.default_rng(17)
    K, N = 6, 40009
    if prior == "gauss":
        D = 24
        X = (rng.normal(size=(N, D)) * 5 + rng.uniform(-20, 20, size=D)).astype(np.float32)
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=5)
    else:
        D = 40
        X, _, _ = multinomial_mixture(N, D, K, seed=3)
        dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=5)
    labels = rng.integers(0, K, N).astype(np.int32)
    subs = rng.integers(0, 2, N).astype(np.int32)
    say(f"statistics after resets {prior}")
    dev.set_labels(labels, subs)
    dev.sufficient_statistics(K)   # full pass

    def check(K):
        lab, sub = dev.get_labels()
        Nsub, sm, S = dev.sufficient_statistics(K)
        oN, os_, oS, _, _, _ = orc.suff_stats(0 if prior == "gauss" else 1, X, lab, sub, K)
        np.testing.assert_array_equal(Nsub, oN)
        if prior == "gauss":
            np.testing.assert_allclose(sm, os_, rtol=1e-9, atol=1e-9 * np.abs(X).sum())
            np.testing.assert_allclose(S, oS, rtol=1e-9, atol=1e-9 * (X.astype(np.float64) ** 2).sum())
        else:
            np.testing.assert_array_equal(sm, os_)

    ident = np.arange(2 * K, dtype=np.int32) // 2
    dev.relabel(3, ident, np.array([0, 1, 0, 0, 1, 0], np.int32))
    check(K)
    dev.relabel(4, ident, np.array([1, 0, 0, 0, 0, 0], np.int32))
    dev.relabel(5, ident, np.array([0, 0, 1, 0, 0, 0], np.int32))   # two resets before one pass
    check(K)
    split = ident.copy()
    split[2 * 3 + 1] = K   # cluster 3's right half becomes cluster K
    dev.relabel(6, split, np.array([0, 0, 0, 2, 0, 0], np.int32))
    check(K + 1)


def test_relabel_table():
    rng = np.random.default_rng(9)
    N, K = 50003, 6
    X = rng.normal(size=(N, 2))
    labels = rng.integers(0, K, N).astype(np.int32)
    subs = rng.integers(0, 2, N).astype(np.int32)
    new_label = rng.integers(0, K + 2, 2 * K).astype(np.int32)
    act = np.array([0, 1, 2, 3, 4, 0], np.int32)
    say("relabel table")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=123)
    dev.set_labels(labels, subs)
    dev.relabel(8, new_label, act)
    lab, sub = dev.get_labels()
    exp_lab = new_label[2 * labels + subs]
    exp_sub = subs.copy()
    for i in np.nonzero(act[labels] == 1)[0][:200]:
        assert sub[i] == orc.philox([int(i), 0, 8, 4], [123, 0])[0] & 1
    for i in np.nonzero(act[labels] == 2)[0][:200]:
        assert sub[i] == orc.philox([int(i), 0, 8, 5], [123, 0])[0] & 1
    exp_sub[act[labels] == 3] = 0
    exp_sub[act[labels] == 4] = 1
    keep = np.isin(act[labels], [0, 3, 4])
    np.testing.assert_array_equal(lab, exp_lab)
    np.testing.assert_array_equal(sub[keep], exp_sub[keep])


def test_initial_labels_match_oracle_schedule():
    N = 10007
    X = np.zeros((N, 2))
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=42)
    say("initial labels")
    dev.sample_initial_labels(5)
    lab, sub = dev.get_labels()
    for i in [0, 1, 17, 5000, N - 1]:
        w = orc.philox([i, 0, 0, 1], [42, 0])
        assert lab[i] == min(int(((w[0] + 0.5) * 2.0 ** -32) * 5), 4)
        assert sub[i] == w[1] & 1


@pytest.mark.parametrize("K,D", [(1, 60), (6, 60), (37, 60), (6, 61), (5, 1030), (3, 2052)])
def test_multinomial_sweeps(K, D):
    """D = 61 and 1030 read rows one word at a time, the others four; D = 1030 and 2052 take
    several 256-lane column passes in the statistics kernel."""
    rng = np.random.default_rng(K)
    N = 20000
    X, gt, theta = multinomial_mixture(N, D, max(K, 2), seed=K)
    lt = np.log(rng.dirichlet(np.ones(D), size=K))
    slt = np.log(rng.dirichlet(np.ones(D), size=2 * K))
    w = rng.dirichlet(np.ones(K))
    lrw = rng.dirichlet([1.0, 1.0], size=K).ravel()
    say(f"multinomial sweeps K={K} D={D}")
    dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=3)
    dev.set_multinomial_params(lt, w, slt, lrw, use_weight=False)
    dev.sample_cluster_labels(2)
    lab, _ = dev.get_labels()
    olab, margin = orc.label_sweep(1, X, logtheta=lt, weights=w, seed=3, it=2)
    _check_labels(lab, olab, margin, FP32_MARGIN, f"multinomial labels K={K}")
    dev.set_labels(olab, None)
    dev.sample_sub_labels(2)
    _, sub = dev.get_labels()
    osub, sm = orc.sublabel_sweep(1, X, olab, logtheta=slt, lr_weights=lrw, seed=3, it=2)
    _check_labels(sub, osub, sm, FP32_MARGIN, f"multinomial sub-labels K={K}")
    Nsub, s, _ = dev.sufficient_statistics(K)
    oN, os_, _, _, _, _ = orc.suff_stats(1, X, olab, osub, K)
    np.testing.assert_array_equal(Nsub, oN)
    np.testing.assert_array_equal(s, os_)   # integer sums are exact


def test_end_to_end_readme_example_fp64():
    """The README basic example shape (module_tests.cpp:19-49): N=1e5, D=2, 2 clusters, FP64.

    The sampler does not always separate the two clusters within 100 iterations: the oracle on
    the same data stays at one cluster for 2 of the seeds 1 .. 20 (1 and 11), where the sub-cluster
    pair never lines up with the true split; seed 2 is a typical run, and the device and the
    oracle take the same number of clusters on it."""
    X, gt, _ = gaussian_mixture(100_000, 2, 2, seed=0, dtype=np.float64)
    say("end to end N=1e5 D=2 FP64")
    model = dp.dp_parallel_sampling_class(X.shape[0], X.T, 2, dp.prior_type.Gaussian)
    hyper = dp.niw_hyperparams(1.0, np.zeros(2), 5.0, np.eye(2))
    mi = model.dp_parallel(hyper, 1.0, iters=100, verbose=False, burnout=15)
    assert mi.cluster_count_history[-1] == 2
    assert dp._nmi(gt, mi.labels) > 0.99
    ref = orc.run(X, prior=0, alpha=1.0, iters=100, seed=2, k0=1.0, v0=5.0)
    assert ref["K"] == 2
    assert orc.nmi(ref["labels"], mi.labels - 1) > 0.99


def test_end_to_end_d32_fp32_matches_oracle_distribution():
    X, gt, _ = gaussian_mixture(60_000, 32, 8, seed=1, a=1.0, ridge=1.0, weight=1.0, dtype=np.float32)
    D = 32
    ks, nmis, oks, onmis = [], [], [], []
    for seed in range(2):
        say(f"end to end D=32 seed={seed}")
        model = dp.dp_parallel_sampling_class(X.shape[0], X.T, seed, dp.prior_type.Gaussian)
        hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
        mi = model.dp_parallel(hyper, 10.0, iters=100, verbose=False, burnout=15)
        ks.append(mi.cluster_count_history[-1])
        nmis.append(dp._nmi(gt, mi.labels))
        ref = orc.run(X, prior=0, alpha=10.0, iters=100, seed=seed, k0=1.0, v0=D + 3.0)
        oks.append(ref["K"])
        onmis.append(orc.nmi(gt, ref["labels"]))
    print("gpu K", ks, "nmi", nmis, "oracle K", oks, "nmi", onmis)
    assert abs(np.mean(nmis) - np.mean(onmis)) < 0.05
    assert abs(np.mean(ks) - np.mean(oks)) <= 2


def _splits_after_cap(khist, k_start, cap):
    """Iterations that began with at least `cap` clusters and ended with more than they began with."""
    starts = [k_start] + list(khist[:-1])
    return [t + 1 for t, (a, b) in enumerate(zip(starts, khist)) if a >= cap and b > a]


def test_max_clusters_stops_splits():
    """max_clusters (dp_parallel_sampling.cpp:147): an iteration that starts with at least
    max_clusters clusters runs with no_more_splits, so from then on the count only falls (merges,
    removals); splits accepted within the iteration that first reaches the cap may overshoot it.
    Checked on the device and on the oracle with the same data and seed."""
    D, cap = 8, 3
    X, gt, _ = gaussian_mixture(60_000, D, 8, seed=5, dtype=np.float64)
    say("max_clusters")
    model = dp.dp_parallel_sampling_class(X.shape[0], X.T, 4, dp.prior_type.Gaussian)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    mi = model.dp_parallel(hyper, 10.0, iters=60, verbose=False, max_clusters=cap)
    kd = list(mi.cluster_count_history)
    ref = orc.run(X, prior=0, alpha=10.0, iters=60, seed=4, k0=1.0, v0=D + 3.0, max_clusters=cap)
    ko = [int(k) for k in ref["k_hist"]]
    free = model.dp_parallel(hyper, 10.0, iters=60, verbose=False).cluster_count_history
    print("device", kd, "\noracle", ko, "\nuncapped device", list(free))
    for k in (kd, ko):
        assert max(k) >= cap and _splits_after_cap(k, 1, cap) == []
    assert kd[-1] < free[-1]
