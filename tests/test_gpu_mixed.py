"""The BF16-MFMA label sweep with FP32 recheck (DPMM_PREC_MIXED) against the FP32 sweep and the
FP64 oracle.  Bar: labels identical to the FP32 sweep (the recheck makes the screen exact when
its error bound holds), hence identical to the oracle wherever its margin exceeds 1e-6."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture
import oracle_ctypes as orc
from helpers import adversarial_mixture, gaussian_params, have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _fitted_params(X, gt, K, rng):
    D = X.shape[1]
    mu = np.stack([X[gt == k].mean(0) for k in range(K)]).astype(np.float64)
    lam = np.stack([np.linalg.inv(np.cov(X[gt == k].T.astype(np.float64)) + 1e-3 * np.eye(D)) for k in range(K)])
    ld = np.array([-np.linalg.slogdet(l)[1] for l in lam])
    w = np.bincount(gt, minlength=K) / len(gt)
    smu = np.repeat(mu, 2, axis=0) + rng.normal(size=(2 * K, D))
    return mu, lam, ld, w, smu, np.repeat(lam, 2, axis=0), np.repeat(ld, 2), np.full(2 * K, 0.5)


@pytest.mark.parametrize("D,K,close", [(32, 12, False), (64, 20, False), (64, 8, True), (100, 6, False),
                                       (128, 5, False), (20, 7, True)])
@pytest.mark.parametrize("final", [False, True])
def test_mixed_equals_fp32_and_oracle(D, K, close, final):
    rng = np.random.default_rng(D + K)
    N = 60_000
    spread = 8.0 if close else 100.0   # close: overlapping clusters, many near-ties
    X, gt, _ = gaussian_mixture(N, D, K, seed=D * 3 + K, spread=spread, dtype=np.float32)
    params = _fitted_params(X, gt, K, rng)
    say(f"mixed sweep D={D} K={K} close={close} final={final}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=9)
    dev.set_gaussian_params(*params)
    dev.sample_cluster_labels(3, final)
    ref, _ = dev.get_labels()
    dev.set_precision(dp.Precision.MIXED)
    dev.set_gaussian_params(*params)
    dev.set_labels(np.full(N, -7, np.int32), None)
    dev.sample_cluster_labels(3, final)
    mx, _ = dev.get_labels()
    nre = dev.last_recheck_count()
    print(f"D={D} K={K} close={close}: rechecked {nre} of {N} ({nre / N:.3%})")
    assert (mx >= 0).all(), "a point got no label"
    np.testing.assert_array_equal(mx, ref)
    olab, margin = orc.label_sweep(0, X, params[0], params[1], params[2], weights=params[3], seed=9, it=3, final=final)
    bad = (mx != olab) & (margin >= 1e-6)
    assert not bad.any()


@pytest.mark.parametrize("D,K", [(32, 16), (64, 20), (100, 8), (128, 6)])
@pytest.mark.parametrize("final", [False, True])
def test_pruned_sweep_equals_unpruned(D, K, final):
    """Exact pruning: labels identical with pruning on and off, from the mixture's own labels (points
    grouped by cluster, so most clusters are skipped) and from scrambled labels, some outside [0, K)."""
    rng = np.random.default_rng(D * K)
    N = 80_000
    X, gt, _ = gaussian_mixture(N, D, K, seed=D + 2 * K, dtype=np.float32)
    params = _fitted_params(X, gt, K, rng)
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=5)
    dev.set_precision(dp.Precision.MIXED)
    dev.set_gaussian_params(*params)
    for start in ("truth", "scrambled"):
        before = gt.astype(np.int32) if start == "truth" else rng.integers(-2, 2 * K, N).astype(np.int32)
        out, frac = {}, None
        for prune in (False, True):
            say(f"pruned sweep D={D} K={K} final={final} start={start} prune={prune}")
            dev.set_label_pruning(prune)
            dev.set_labels(before, None)
            dev.sample_cluster_labels(6, final)
            out[prune] = dev.get_labels()[0]
            if prune:
                frac = dev.last_label_evaluations()
            else:
                assert dev.last_label_evaluations() == -1.0
        print(f"D={D} K={K} final={final} {start}: evaluated fraction {frac:.3f}")
        assert (out[True] >= 0).all() and (out[True] < K).all()
        np.testing.assert_array_equal(out[True], out[False])
        if start == "truth":   # separated clusters: most of them are skipped
            assert frac < 0.75
    dev.close()


@pytest.mark.parametrize("D,K", [(64, 10), (128, 6)])
def test_pruned_sweep_with_sibling_clusters(D, K):
    """Clusters in pairs of near-identical siblings (an over-split start): the BF16 screen leaves
    most points to the FP32 recheck, which is pruned with the same bounds; labels stay identical to
    the unpruned sweep's and to the plain FP32 sweep's."""
    rng = np.random.default_rng(D + K)
    N = 60_000
    X, gt, _ = gaussian_mixture(N, D, K, seed=D * 7 + K, dtype=np.float32)
    mu, lam, ld, w, smu, slam, sld, lrw = _fitted_params(X, gt, K, rng)
    mu2 = np.repeat(mu, 2, axis=0) + rng.normal(scale=0.01, size=(2 * K, D))
    lam2, ld2 = np.repeat(lam, 2, axis=0), np.repeat(ld, 2)
    w2 = np.repeat(w, 2) / 2
    smu2 = np.repeat(smu, 2, axis=0)
    slam2, sld2, lrw2 = np.repeat(slam, 2, axis=0), np.repeat(sld, 2), np.repeat(lrw, 2)
    params = (mu2, lam2, ld2, w2, smu2, slam2, sld2, lrw2)
    before = (2 * gt + rng.integers(0, 2, N)).astype(np.int32)
    say(f"sibling clusters D={D} K={2 * K}")
    ref_dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=8)
    ref_dev.set_gaussian_params(*params)
    ref_dev.sample_cluster_labels(4, False)
    ref, _ = ref_dev.get_labels()
    ref_dev.close()
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=8)
    dev.set_precision(dp.Precision.MIXED)
    dev.set_gaussian_params(*params)
    out = {}
    for prune in (False, True):
        dev.set_label_pruning(prune)
        dev.set_labels(before, None)
        dev.sample_cluster_labels(4, False)
        out[prune] = dev.get_labels()[0]
        print(f"D={D} prune={prune}: rechecked {dev.last_recheck_count()} of {N}, "
              f"evaluated {dev.last_label_evaluations():.3f}")
    np.testing.assert_array_equal(out[True], out[False])
    np.testing.assert_array_equal(out[True], ref)
    dev.close()


@pytest.mark.parametrize("D", [64, 128])
@pytest.mark.parametrize("final", [False, True])
def test_mixed_equals_fp32_on_ill_conditioned_near_ties(D, final):
    """Adversarial data for the BF16 screen: covariances with ridge 1e-3 and a dominant correlated
    direction (condition number ~4e5) and means that differ only along high-variance directions, so
    clusters overlap and W (x - mu) cancels large terms (tests/helpers.adversarial_mixture).  The
    screen's bound scales with |W| |x - mu| and keeps only labels the FP32 sweep would draw: mixed
    labels, pruned or not, equal the FP32 sweep's; both agree with the FP64 oracle wherever its margin
    exceeds 1e-6.  The recheck fraction is reported."""
    K, N = 8, 40_000
    X, gt, mus, sig = adversarial_mixture(N, D, K, seed=D)
    rng = np.random.default_rng(D)
    lam = np.linalg.inv(sig)
    ld = np.linalg.slogdet(sig)[1]
    w = np.full(K, 1.0 / K)
    smu = np.repeat(mus, 2, axis=0) + rng.normal(scale=0.1, size=(2 * K, D))
    params = (mus, lam, ld, w, smu, np.repeat(lam, 2, axis=0), np.repeat(ld, 2), np.full(2 * K, 0.5))
    say(f"ill-conditioned near-ties D={D} K={K} final={final}: FP32 sweep")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=13)
    dev.set_gaussian_params(*params)
    dev.sample_cluster_labels(7, final)
    ref, _ = dev.get_labels()
    dev.set_precision(dp.Precision.MIXED)
    dev.set_gaussian_params(*params)
    for prune in (False, True):
        say(f"ill-conditioned near-ties D={D} K={K} final={final}: mixed, pruning={prune}")
        dev.set_label_pruning(prune)
        dev.set_labels(gt, None)
        dev.sample_cluster_labels(7, final)
        mx, _ = dev.get_labels()
        nre = dev.last_recheck_count()
        print(f"D={D} final={final} pruning={prune}: rechecked {nre} of {N} ({nre / N:.1%})")
        np.testing.assert_array_equal(mx, ref)
    olab, margin = orc.label_sweep(0, X, mus, lam, ld, weights=w, seed=13, it=7, final=final)
    flips = np.nonzero(ref != olab)[0]
    print(f"D={D}: {flips.size} labels differ from the FP64 oracle, largest margin "
          f"{margin[flips].max() if flips.size else 0:.2e}")
    assert not ((ref != olab) & (margin >= 1e-6)).any()
    dev.close()


def test_screen_alone_flip_rate():
    """A bound scale of 0 keeps every screened label: the raw BF16 flip rate against the oracle, reported."""
    rng = np.random.default_rng(1)
    D, K, N = 64, 20, 100_000
    X, gt, _ = gaussian_mixture(N, D, K, seed=5, dtype=np.float32)
    params = _fitted_params(X, gt, K, rng)
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=2)
    dev.set_precision(dp.Precision.MIXED)
    dev.set_mixed_tolerance(0.0)
    dev.set_gaussian_params(*params)
    say("mixed screen only D=64 K=20")
    dev.sample_cluster_labels(4)
    mx, _ = dev.get_labels()
    olab, margin = orc.label_sweep(0, X, params[0], params[1], params[2], weights=params[3], seed=2, it=4)
    rate = (mx != olab).mean()
    print(f"BF16 screen alone: flip rate {rate:.2e}, rechecked {dev.last_recheck_count()}")
    assert rate < 1e-2


def test_mixed_sampler_end_to_end():
    X, gt, _ = gaussian_mixture(80_000, 64, 10, seed=3, a=1.0, ridge=0.5, weight=0.5, dtype=np.float32)
    D = 64
    say("mixed end to end D=64")
    model = dp.dp_parallel_sampling_class(X.shape[0], X.T, 4, dp.prior_type.Gaussian)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    mi = model.dp_parallel(hyper, 10.0, iters=120, verbose=False, precision=dp.Precision.MIXED)
    ref = model.dp_parallel(hyper, 10.0, iters=120, verbose=False, precision=dp.Precision.FP32)
    nm, nr = dp._nmi(gt, mi.labels), dp._nmi(gt, ref.labels)
    print("mixed K", mi.cluster_count_history[-1], "nmi", nm, "fp32 K", ref.cluster_count_history[-1], "nmi", nr)
    assert abs(nm - nr) < 0.01
    np.testing.assert_array_equal(mi.labels, ref.labels)   # same draws, exact screen -> same trajectory


def _docs(N, D, K, seed, words=80):
    rng = np.random.default_rng(seed)
    theta = rng.dirichlet(np.full(D, 0.1), K)
    gt = rng.integers(0, K, N)
    cdf = np.cumsum(theta, 1)
    X = np.zeros((N, D), np.int32)
    for k in range(K):
        idx = np.nonzero(gt == k)[0]
        w = np.minimum(np.searchsorted(cdf[k], rng.random((idx.size, words)) * cdf[k, -1]), D - 1)
        np.add.at(X, (np.repeat(idx, words), w.ravel()), 1)
    return X, gt, theta


@pytest.mark.parametrize("D,K", [(300, 20), (1000, 100), (200, 200)])
@pytest.mark.parametrize("final", [False, True])
def test_mixed_multinomial_equals_fp32_and_oracle(D, K, final):
    """Split-BF16 MFMA screen + FP32 recheck: labels identical to the FP32 sweep."""
    rng = np.random.default_rng(D + K)
    X, gt, theta = _docs(20_000, D, K, seed=K)
    # parameters near the truth, plus a few near-duplicate topics to force close calls
    lt = np.log(np.stack([rng.dirichlet(theta[k] * 500 + 0.05) for k in range(K)]))
    lt[1::7] = lt[0::7][:len(lt[1::7])] + rng.normal(size=lt[1::7].shape) * 0.01
    w = rng.dirichlet(np.ones(K))
    slt, lrw = np.repeat(lt, 2, axis=0), np.full(2 * K, 0.5)
    say(f"mixed multinomial D={D} K={K} final={final}")
    dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=7)
    dev.set_multinomial_params(lt, w, slt, lrw)
    dev.sample_cluster_labels(5, final)
    ref, _ = dev.get_labels()
    dev.set_precision(dp.Precision.MIXED)
    dev.set_multinomial_params(lt, w, slt, lrw)
    dev.set_labels(np.full(len(X), -7, np.int32), None)
    dev.sample_cluster_labels(5, final)
    mx, _ = dev.get_labels()
    nre = dev.last_recheck_count()
    print(f"D={D} K={K} final={final}: rechecked {nre} of {len(X)}")
    assert (mx >= 0).all()
    np.testing.assert_array_equal(mx, ref)
    olab, margin = orc.label_sweep(1, X, logtheta=lt, weights=w, seed=7, it=5, final=final)
    assert not ((mx != olab) & (margin >= 1e-6)).any()


def test_mixed_multinomial_rejects_large_counts():
    X = np.zeros((100, 10), np.int32)
    X[3, 4] = 300
    dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=1)
    with pytest.raises(dp.DPMMError):
        dev.set_precision(dp.Precision.MIXED)


def test_mixed_multinomial_sampler_end_to_end():
    X, gt, _ = _docs(30_000, 200, 8, seed=3)
    say("mixed multinomial end to end")
    runs = {}
    for prec in (dp.Precision.FP32, dp.Precision.MIXED):
        dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=5)
        s = dp.Sampler(dev, dp.multinomial_hyper(np.ones(200)), 10.0, 60, seed=5, precision=prec)
        for t in range(1, 61):
            s.step(t)
        runs[prec] = dev.get_labels()[0]
        print(prec, "K", s.num_clusters(), "nmi", dp._nmi(gt, runs[prec]))
        s.close()
        dev.close()
    np.testing.assert_array_equal(runs[dp.Precision.MIXED], runs[dp.Precision.FP32])


@pytest.mark.parametrize("D,K,close", [(32, 12, False), (64, 20, False), (64, 8, True), (100, 6, False),
                                       (128, 5, False), (20, 7, True)])
def test_mixed_sublabels_equal_fp32_and_oracle(D, K, close):
    """BF16-MFMA sub-label screen + FP32 recheck: sub-labels identical to the FP32 sub-label sweep."""
    rng = np.random.default_rng(D * 7 + K)
    N = 60_000
    X, gt, _ = gaussian_mixture(N, D, K, seed=D * 5 + K, spread=8.0 if close else 100.0, dtype=np.float32)
    mu, lam, ld, w, smu, slam, sld, lrw = _fitted_params(X, gt, K, rng)
    if close:   # sub-clusters nearly equal: most draws are close calls
        smu = np.repeat(mu, 2, axis=0) + rng.normal(size=(2 * K, D)) * 0.01
    labels = gt.astype(np.int32)
    say(f"mixed sub-labels D={D} K={K} close={close}")
    out = {}
    for prec in (dp.Precision.FP32, dp.Precision.MIXED):
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=13)
        dev.set_precision(prec)
        dev.set_gaussian_params(mu, lam, ld, w, smu, slam, sld, lrw)
        dev.set_labels(labels, np.zeros(N, np.int32))
        dev.sample_sub_labels(6)
        out[prec] = dev.get_labels()[1]
        if prec == dp.Precision.MIXED:
            nre = dev.last_sub_recheck_count()
            print(f"D={D} K={K} close={close}: rechecked {nre} of {N}")
            assert nre < 0.05 * N   # the split screen decides the bulk of the points
        dev.close()
    np.testing.assert_array_equal(out[dp.Precision.MIXED], out[dp.Precision.FP32])
    osub, margin = orc.sublabel_sweep(0, X, labels, smu, slam, sld, lr_weights=lrw, seed=13, it=6)
    assert not ((out[dp.Precision.MIXED] != osub) & (margin >= 1e-6)).any()


@pytest.mark.parametrize("D,K", [(64, 6), (128, 5), (32, 8)])
def test_refined_sibling_ties_equal_fp32(D, K):
    """Sibling clusters (each cluster of the mixture present twice, with means 0.05 apart and the same
    covariance), the situation of a split or an over-split start: the BF16 screen leaves most points
    within its bound of a tie.  The pruned sweep's split-BF16 refinement must decide them as the FP32
    sweep does (labels identical, hence the oracle's wherever its margin exceeds 1e-6), and leave far
    fewer points to the FP32 recheck than with the refinement off."""
    rng = np.random.default_rng(D * 11 + K)
    N = 40_000
    X, gt, _ = gaussian_mixture(N, D, K, seed=D + K, dtype=np.float32)
    mu, lam, ld, w, smu, slam, sld, lrw = _fitted_params(X, gt, K, rng)
    mu2 = np.repeat(mu, 2, axis=0)
    mu2[1::2] += 0.05 * rng.normal(size=(K, D)) / np.sqrt(D)
    lam2, ld2, w2 = np.repeat(lam, 2, axis=0), np.repeat(ld, 2), np.repeat(w, 2) / 2
    K2 = 2 * K
    params = (mu2, lam2, ld2, w2, np.repeat(mu2, 2, axis=0), np.repeat(lam2, 2, axis=0), np.repeat(ld2, 2),
              np.full(2 * K2, 0.5))
    init = (2 * gt + rng.integers(0, 2, N)).astype(np.int32)
    say(f"refined sibling ties D={D} K={K2}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=12)
    dev.set_gaussian_params(*params)
    dev.set_labels(init, None)
    dev.sample_cluster_labels(5)
    ref, _ = dev.get_labels()
    counts = {}
    for refine in (True, False):
        dev.set_precision(dp.Precision.MIXED)
        dev.set_label_pruning(True)
        dev.set_label_refinement(refine)
        dev.set_gaussian_params(*params)
        dev.set_labels(init, None)
        dev.sample_cluster_labels(5)
        mx, _ = dev.get_labels()
        np.testing.assert_array_equal(mx, ref, err_msg=f"refinement {refine}")
        counts[refine] = dev.last_recheck_count()
        dev.set_precision(dp.Precision.FP32)
    print(f"D={D} K={K2}: FP32 rechecks {counts[True]} with the refinement, {counts[False]} without, of {N}")
    assert counts[False] > N // 10, "the siblings should defeat the BF16 screen"
    # at D = 128 the FP32 sums' share of the split bound, (5 DP + 24) 2^-24, exceeds the split's own
    # 2^-15, so fewer ties resolve there (about 10 % of the points stay undecided)
    assert counts[True] * (10 if D <= 64 else 5) < counts[False]
    olab, margin = orc.label_sweep(0, X, params[0], params[1], params[2], weights=params[3], seed=12, it=5)
    assert not ((ref != olab) & (margin >= 1e-6)).any()
    dev.close()
