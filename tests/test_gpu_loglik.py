"""Per-point log-likelihood scores (dpmm_log_likelihoods, the rows of the reference's d_r matrix)
against the FP64 oracle's scores of the same parameters, and the label-flip rate of a sweep at
parameters the sampler itself drew.

Tolerances, from the arithmetic rather than fitted.  A Gaussian score is c - q / 2 with
c = -logdet / 2 + log w (the common -D^2 log(2 pi) / 2 is added in FP64) and q the whitened squared distance.  The FP32
sweep rounds mu and the factor W to FP32 and sums D-term fma chains, so
|s - s_oracle| <= 2^-24 ((4 D + 16)(|c| + q / 2 + 1) + 4 sqrt(q) ||W||_2 max|mu| sqrt(D)); the FP64
sweep's error is below 1e-9 (|c| + q / 2 + 1).  Multinomial FP32 scores sum at most `words`
products of size <= max|log theta|: within 2^-20 words max|log theta|.  The noise-free sweep's
labels are the argmax of the same scores, bit for bit."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture, multinomial_mixture
import oracle_ctypes as orc
from helpers import gaussian_params, have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _gauss_bound(s_oracle, D, mu, lam, ld, logw, fp64):
    """Error bound of each score (rows: points, columns: clusters) from the oracle's score."""
    common = -0.5 * D * D * np.log(2 * np.pi)   # added in FP64 by the library, outside the sweep
    c = -0.5 * ld + logw
    q = np.maximum(0.0, 2.0 * (c[None, :] + common - s_oracle))
    scale = np.abs(c)[None, :] + 0.5 * q + 1.0
    if fp64:
        return 1e-9 * scale
    wn = np.sqrt(np.max([np.linalg.eigvalsh(l).max() for l in lam]))
    return 2.0 ** -24 * ((4 * D + 16) * scale + 4 * np.sqrt(q) * wn * np.abs(mu).max() * np.sqrt(D))


@pytest.mark.parametrize("D,K,dtype", [(2, 3, np.float64), (17, 5, np.float64), (64, 4, np.float64),
                                       (2, 4, np.float32), (32, 9, np.float32), (64, 7, np.float32),
                                       (100, 3, np.float32), (128, 5, np.float32)])
def test_gaussian_scores_match_oracle(D, K, dtype):
    rng = np.random.default_rng(D + 7 * K)
    N = 4099
    X, _, _ = gaussian_mixture(N, D, max(2, K // 2), seed=D * K, dtype=dtype)
    mu, lam, ld = gaussian_params(rng, X, K)
    smu, slam, sld = gaussian_params(rng, X, K, sub=True)
    w = rng.dirichlet(np.ones(K))
    lrw = rng.dirichlet([1.0, 1.0], size=K).ravel()
    say(f"scores gaussian D={D} K={K} dtype={np.dtype(dtype).name}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=9)
    dev.set_gaussian_params(mu, lam, ld, w, smu, slam, sld, lrw)
    first, n = 1000, 3000
    fp64 = dtype == np.float64
    s = dev.log_likelihoods(K, 0, first, n)
    olab, _, os_ = orc.label_sweep(0, X, mu, lam, ld, weights=w, seed=9, it=1, final=True, want_scores=True)
    want = os_[first:first + n]
    err = np.abs(s - want)
    bound = _gauss_bound(want, D, mu, lam, ld, np.log(w), fp64)
    print(f"D={D} K={K} {np.dtype(dtype).name}: cluster score error max {err.max():.2e}, "
          f"max error / bound {(err / bound).max():.3f}")
    assert np.all(err <= bound)
    # the noise-free sweep takes the argmax of exactly these scores
    dev.sample_cluster_labels(1, True)
    lab, _ = dev.get_labels()
    np.testing.assert_array_equal(lab[first:first + n], np.argmax(s, axis=1))
    # sub-cluster scores of each point's label
    dev.set_labels(olab, np.zeros(N, np.int32))
    ss = dev.log_likelihoods(K, 1, first, n)
    _, _, osub = orc.label_sweep(0, X, smu, slam, sld, weights=lrw, seed=9, it=1, final=True, want_scores=True)
    sbound_all = _gauss_bound(osub[first:first + n], D, smu, slam, sld, np.log(lrw), fp64)
    rows = np.arange(n)
    k = olab[first:first + n]
    want_s = np.stack([osub[first + rows, 2 * k], osub[first + rows, 2 * k + 1]], axis=1)
    sbound = np.stack([sbound_all[rows, 2 * k], sbound_all[rows, 2 * k + 1]], axis=1)
    err_s = np.abs(ss - want_s)
    print(f"D={D} K={K}: sub-cluster score error max {err_s.max():.2e}, max error / bound {(err_s / sbound).max():.3f}")
    assert np.all(err_s <= sbound)
    dev.close()


@pytest.mark.parametrize("K,D", [(1, 300), (6, 300), (70, 300), (6, 301)])
def test_multinomial_scores_match_oracle(K, D):
    rng = np.random.default_rng(K + 3)
    N = 3000
    X, _, _ = multinomial_mixture(N, D, max(K, 2), seed=K)
    lt = np.log(rng.dirichlet(np.ones(D), size=K))
    slt = np.log(rng.dirichlet(np.ones(D), size=2 * K))
    w = rng.dirichlet(np.ones(K))
    lrw = rng.dirichlet([1.0, 1.0], size=K).ravel()
    say(f"scores multinomial K={K} D={D}")
    dev = dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=4)
    dev.set_multinomial_params(lt, w, slt, lrw, use_weight=False)
    s = dev.log_likelihoods(K, 0, 0, N)
    olab, _, os_ = orc.label_sweep(1, X, logtheta=lt, weights=w, seed=4, it=1, final=True, want_scores=True)
    scale = X.sum(1).max() * max(np.abs(lt).max(), np.abs(slt).max())
    err = float(np.max(np.abs(s - os_)))
    print(f"multinomial K={K}: score error {err:.2e} (scale {scale:.1f})")
    assert err <= scale * 2.0 ** -20
    dev.sample_cluster_labels(1, True)
    lab, _ = dev.get_labels()
    np.testing.assert_array_equal(lab, np.argmax(s, axis=1))
    dev.set_labels(olab, None)
    ss = dev.log_likelihoods(K, 1, 0, N)
    _, _, osub = orc.label_sweep(1, X, logtheta=slt, weights=lrw, seed=4, it=1, final=True,
                                 mult_use_weight=True, want_scores=True)
    want = np.stack([osub[np.arange(N), 2 * olab], osub[np.arange(N), 2 * olab + 1]], axis=1)
    err_s = float(np.max(np.abs(ss - want)))
    print(f"multinomial K={K}: sub-cluster score error {err_s:.2e}")
    assert err_s <= scale * 2.0 ** -20
    dev.close()


def test_scores_reject_bad_arguments():
    X, _, _ = gaussian_mixture(100, 3, 2, seed=0, dtype=np.float32)
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=1)
    with pytest.raises(dp.DPMMError):
        dev.log_likelihoods(2, 0, 0, 10)          # no parameters set yet
    rng = np.random.default_rng(0)
    mu, lam, ld = gaussian_params(rng, X, 2)
    smu, slam, sld = gaussian_params(rng, X, 2, sub=True)
    dev.set_gaussian_params(mu, lam, ld, np.full(2, 0.5), smu, slam, sld, np.full(4, 0.5))
    with pytest.raises(dp.DPMMError):
        dev.log_likelihoods(3, 0, 0, 10)          # K differs from the parameters'
    with pytest.raises(dp.DPMMError):
        dev.log_likelihoods(2, 0, 95, 10)         # past the last point
    dev.set_labels(np.full(100, 5, np.int32), None)
    with pytest.raises(dp.DPMMError):
        dev.log_likelihoods(2, 1, 0, 10)          # labels outside the clusters set
    dev.close()


@pytest.mark.parametrize("prec", ["FP32", "MIXED"])
def test_flip_rate_at_sampler_drawn_parameters(prec):
    """Run the sampler, then sweep once more at the parameters it drew in its last iteration and
    compare with the FP64 oracle on the same noise: label-flip rate <= 1e-4, every flip at a
    relative margin below 1e-6, sub-label flips likewise."""
    D, Kt, N = 64, 12, 60_000
    X, gt, _ = gaussian_mixture(N, D, Kt, seed=21, a=1.0, ridge=0.5, weight=0.5, dtype=np.float32)
    say(f"flip rate at drawn parameters {prec}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=3)
    smp = dp.Sampler(dev, dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D)), 10.0, 60, seed=3,
                     precision=getattr(dp.Precision, prec))
    for t in range(1, 41):
        smp.step(t)
    K = smp.num_clusters()
    w = smp.weights()[0]
    lrw = np.random.default_rng(5).dirichlet([2.0, 2.0], size=K).ravel()
    dev.sufficient_statistics(K, counts_only=True)
    mem = np.array([[2 * k, 2 * k + 1, -1, -1] if p == 0 else [2 * k + p - 1, -1, -1, -1]
                    for k in range(K) for p in range(3)], np.int32)
    it = 41
    dev.sample_params(it, mem, w, lrw)          # draws packed into the sweep's slots on the device
    mu, R, ld = dev.drawn_params(K)
    c = np.arange(K) * 3
    sub = np.sort(np.r_[c + 1, c + 2])
    dev.sample_cluster_labels(it, False)
    lab, _ = dev.get_labels()
    lam = np.einsum("kij,klj->kil", R, R)
    olab, margin = orc.label_sweep(0, X, mu[c], lam[c], ld[c], weights=w, seed=3, it=it)
    flips = np.nonzero(lab != olab)[0]
    print(f"{prec}: K={K} label flips {flips.size}/{N}")
    assert flips.size / N <= 1e-4
    if flips.size:
        assert margin[flips].max() < 1e-6
    dev.set_labels(olab, np.zeros(N, np.int32))
    dev.sample_sub_labels(it)
    _, s = dev.get_labels()
    osub, smargin = orc.sublabel_sweep(0, X, olab, mu[sub], lam[sub], ld[sub], lr_weights=lrw, seed=3, it=it)
    sflips = np.nonzero(s != osub)[0]
    print(f"{prec}: sub-label flips {sflips.size}/{N}")
    assert sflips.size / N <= 1e-4
    if sflips.size:
        assert smargin[sflips].max() < 1e-6
    smp.close()
    dev.close()
