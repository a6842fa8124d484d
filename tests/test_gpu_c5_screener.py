"""The merge-screener workload at reduced size: D = 128, true clusters with A ~ N(0, 0.5^2), the
sampler started over-split into two fragments per cluster, mixed precision, exact merge tests only
for screened pairs with a full all-pairs scan every 10 iterations.  The screener is trained on merge
decisions the plain sampler logged on mixtures of the two small configurations' families (D = 2 and
D = 32), never on D = 128 data.  Bar: no accepted merge screened out, the cluster-count history and
final labels identical to the all-pairs run's, and the D = 128 log marginal likelihoods of the merge
pairs' unions within 1e-9 relative of the FP64 oracle's.
"""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd import screener as scr
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture
import oracle_ctypes as orc
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def test_c5_shape_screened_run_equals_all_pairs():
    runs = [("c1", 201), ("c1", 202), ("c2", 203), ("c2", 204), ("c2", 205), ("c2", 206)]
    say("screener training data from configs-1-2-style mixtures")
    X_tr, h, u = scr.collect_config_style_pairs(runs, n_points=60_000, pieces=3, iters=60,
                                                precision=dp.Precision.MIXED)
    y = scr.plausible(h)
    print(f"training pairs {len(y)}, plausible {int(y.sum())}, accepted {int((h > u).sum())}")
    assert y.sum() >= 20
    s, hist = scr.train(X_tr, y, epochs=30, batch=4096, lr=2e-3, seed=2)
    thr = scr.threshold_for_recall(s.logits(X_tr), y, recall=1.0, margin=8.0)

    N, D, K, iters = 200_000, 128, 40, 100
    X, gt, _ = gaussian_mixture(N, D, K, seed=31, a=0.5, ridge=1.0, dtype=np.float32)
    init = scr.over_split_labels(gt, 2, 31)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    out = {}
    for name, use in (("all_pairs", False), ("screened", True)):
        say(f"C5-shape run N={N} D={D} K*={K} over-split 2x, {name}")
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=32)
        smp = dp.Sampler(dev, hyper, 10.0, iters, seed=32, precision=dp.Precision.MIXED, init_labels=init)
        smp.set_screener(s if use else None, full_scan_every=10, threshold=thr)
        ks = []
        for t in range(1, iters + 1):
            smp.step(t)
            ks.append(smp.num_clusters())
        lab, _ = dev.get_labels()
        out[name] = dict(k=ks, labels=lab, stats=smp.screen_stats(), nmi=dp._nmi(gt, lab))
        if use:
            # log ML of every pair's union at the final statistics, device against the oracle
            Kf = smp.num_clusters()
            Nsub, sums, S = dev.sufficient_statistics(Kf)
            pairs = [(i, j) for i in range(Kf) for j in range(i + 1, Kf)][:300]
            mem = np.array([[2 * i, 2 * i + 1, 2 * j, 2 * j + 1] for i, j in pairs], np.int32)
            L, _ = dev.log_marginal_batch(mem)
            for (i, j), Lij in zip(pairs, L):
                q = [2 * i, 2 * i + 1, 2 * j, 2 * j + 1]
                n = int(Nsub[q].sum())
                kn, mn, vn, psin = orc.niw_posterior(1.0, np.zeros(D), D + 3.0, np.eye(D), n, sums[q].sum(0),
                                                     S[q].sum(0))
                want = orc.niw_log_marginal(1.0, D + 3.0, np.eye(D), kn, vn, psin, n)
                assert abs(Lij - want) <= 1e-9 * abs(want), f"pair ({i}, {j}): {Lij} vs {want}"
        smp.close()
        dev.close()
    a, b = out["all_pairs"], out["screened"]
    print("K history", b["k"][:12], "...", b["k"][-3:])
    print("exact tests all-pairs", a["stats"][0], "screened", b["stats"][0], "skipped", b["stats"][1],
          "missed", b["stats"][2], "full scans", b["stats"][3], "NMI", a["nmi"], b["nmi"], "threshold", thr)
    assert b["stats"][2] == 0, "an accepted merge was screened out"
    assert b["k"] == a["k"]
    np.testing.assert_array_equal(b["labels"], a["labels"])
    assert b["k"][-1] < 2 * K and b["nmi"] > 0.99
    assert b["stats"][0] < a["stats"][0]
