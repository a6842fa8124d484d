"""The device sampler followed iteration by iteration against the FP64 oracle's sampler.

Both run the same iteration (group_step, local_clusters_actions.cpp:460-490) from the same data
and seed: the device on its FP64 path through the C ABI, the oracle in plain FP64 C++ on the same
Philox schedule.  After every iteration the labels and sub-labels the two sweeps drew, the cluster
count, the per-cluster counts, the mixture weights and the final labels must be identical.  A
difference is allowed only where it starts: at a point whose winning margin in the oracle's sweep
is below 1e-12 relative (FP64 rounding can decide such a tie either way); the comparison stops
there.  This pins the order of group_step, reset_bad_clusters, the split and merge decisions, the
greedy merge resolution and the removal of empty clusters step by step.
"""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture
import oracle_ctypes as orc
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu

FP64_MARGIN = 1e-12


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


CASES = {
    # the README basic example (module_tests.cpp:19-49): N = 1e5, D = 2, two clusters, alpha 1, nu = 5
    "readme_d2": dict(N=100_000, D=2, K=2, data_seed=0, seed=2, alpha=1.0, v0=5.0),
    "d8_k8": dict(N=40_000, D=8, K=8, data_seed=3, seed=5, alpha=10.0, v0=11.0),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_lockstep_fp64_against_oracle(case):
    c = CASES[case]
    iters = 100
    X, gt, _ = gaussian_mixture(c["N"], c["D"], c["K"], seed=c["data_seed"], dtype=np.float64)
    D = c["D"]
    say(f"lockstep {case}: N={c['N']} D={D} FP64, {iters} iterations")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=c["seed"])
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), c["v0"], np.eye(D))
    smp = dp.Sampler(dev, hyper, c["alpha"], iters, seed=c["seed"], precision=dp.Precision.FP64)
    smp.trace_sweeps(True)
    om = orc.Model(X, prior=0, alpha=c["alpha"], iters=iters, seed=c["seed"], k0=1.0, v0=c["v0"], trace=True)
    khist = []
    for t in range(1, iters + 1):
        smp.step(t)
        om.step(t)
        glab, gsub = smp.sweep_trace()
        olab, osub, margin, smargin = om.sweep_trace()
        flips = np.nonzero(glab != olab)[0]
        if flips.size:
            print(f"{case}: iteration {t}: {flips.size} label flips, largest oracle margin {margin[flips].max():.3e}")
            assert margin[flips].max() < FP64_MARGIN
            return
        sflips = np.nonzero(gsub != osub)[0]
        if sflips.size:
            print(f"{case}: iteration {t}: {sflips.size} sub-label flips, largest margin {smargin[sflips].max():.3e}")
            assert smargin[sflips].max() < FP64_MARGIN
            return
        K = smp.num_clusters()
        assert K == om.num_clusters(), f"iteration {t}: {K} clusters on the device, {om.num_clusters()} in the oracle"
        w, cnt = smp.weights()
        lab, sub = dev.get_labels()
        olab2, osub2, ocnt, ow = om.state()
        np.testing.assert_array_equal(cnt, ocnt, err_msg=f"iteration {t}: cluster counts")
        np.testing.assert_array_equal(lab, olab2, err_msg=f"iteration {t}: labels after the step")
        np.testing.assert_array_equal(sub, osub2, err_msg=f"iteration {t}: sub-labels after the step")
        np.testing.assert_allclose(w, ow, rtol=1e-12, atol=0, err_msg=f"iteration {t}: mixture weights")
        khist.append(K)
    print(f"{case}: identical for all {iters} iterations; K history {khist}")
    assert dp._nmi(gt, dev.get_labels()[0]) > 0.95
    smp.close()
    dev.close()
    om.close()


def test_cluster_cap_stops_splitting_and_run_continues():
    """A context whose kernels take at most `cap` clusters (lowered with dpmm_limit_clusters): the
    sampler treats the cap as max_clusters (no splits from an iteration that starts at it) and
    accepts, within one iteration, only as many splits as fit, so K never exceeds the cap and the
    run finishes instead of failing in the next sweep."""
    D, cap = 8, 4
    X, gt, _ = gaussian_mixture(60_000, D, 8, seed=5, dtype=np.float32)
    say(f"cluster cap {cap}")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=4)
    dev.limit_clusters(cap)
    assert dev.max_clusters() == cap
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    smp = dp.Sampler(dev, hyper, 10.0, 60, seed=4, init_clusters=3)
    ks = []
    for t in range(1, 61):
        smp.step(t)
        ks.append(smp.num_clusters())
    print("K history", ks)
    assert max(ks) == cap and all(k <= cap for k in ks)
    dev.limit_clusters(0)
    assert dev.max_clusters() == 2048
    smp.close()
    dev.close()
