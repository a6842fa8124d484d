"""The device sampler followed iteration by iteration against the FP64 oracle's sampler.

Both run the same iteration (group_step, local_clusters_actions.cpp:460-490) from the same data
and seed: the device on its FP64 path through the C ABI, the oracle in plain FP64 C++ on the same
Philox schedule.  The device sampler runs in both of its modes: every iteration on the device with
the decisions in a device kernel, and the decisions on the host between device launches.  After every iteration the labels and sub-labels the two sweeps drew, the cluster
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


def _lockstep(name, X, gt, seed, alpha, v0, iters=100, init_clusters=1, outlier=None, min_nmi=0.95,
              host_decisions=False):
    """Device (FP64) and oracle samplers side by side; see the module docstring."""
    D = X.shape[1]
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=seed)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), v0, np.eye(D))
    kw, okw = {}, {}
    if outlier is not None:
        w, (ok0, om0, ov0, opsi0) = outlier
        kw = dict(outlier_weight=w, outlier_params=dp.niw_hyperparams(ok0, om0, ov0, opsi0))
        okw = dict(outlier_weight=w, outlier_params=(ok0, om0, ov0, opsi0))
    smp = dp.Sampler(dev, hyper, alpha, iters, seed=seed, precision=dp.Precision.FP64, init_clusters=init_clusters,
                     host_decisions=host_decisions, **kw)
    smp.trace_sweeps(True)
    om = orc.Model(X, prior=0, alpha=alpha, iters=iters, seed=seed, k0=1.0, v0=v0, init_clusters=init_clusters,
                   trace=True, **okw)
    khist = []
    for t in range(1, iters + 1):
        smp.step(t)
        om.step(t)
        glab, gsub = smp.sweep_trace()
        olab, osub, margin, smargin = om.sweep_trace()
        flips = np.nonzero(glab != olab)[0]
        if flips.size:
            print(f"{name}: iteration {t}: {flips.size} label flips, largest oracle margin {margin[flips].max():.3e}")
            assert margin[flips].max() < FP64_MARGIN
            return None
        sflips = np.nonzero(gsub != osub)[0]
        if sflips.size:
            print(f"{name}: iteration {t}: {sflips.size} sub-label flips, largest margin {smargin[sflips].max():.3e}")
            assert smargin[sflips].max() < FP64_MARGIN
            return None
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
    print(f"{name}: identical for all {iters} iterations; K history {khist}; events {smp.events()}")
    lab = dev.get_labels()[0]
    if min_nmi is not None:
        assert dp._nmi(gt, lab) > min_nmi
    smp.close()
    dev.close()
    om.close()
    return khist, lab, w


@pytest.mark.parametrize("mode", ["device", "host"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_lockstep_fp64_against_oracle(case, mode):
    c = CASES[case]
    X, gt, _ = gaussian_mixture(c["N"], c["D"], c["K"], seed=c["data_seed"], dtype=np.float64)
    say(f"lockstep {case} ({mode} decisions): N={c['N']} D={c['D']} FP64, 100 iterations")
    _lockstep(case, X, gt, c["seed"], c["alpha"], c["v0"], host_decisions=mode == "host")


@pytest.mark.parametrize("init_clusters", [1, 3])
def test_lockstep_outlier_mode_against_oracle(init_clusters):
    """The reference's outlier mode (dp_parallel's outlier_weight, dp_parallel_sampling.cpp:68-72): cluster
    0 is the outlier cluster with its own broad prior and weight 0.05; cluster index 1 is never redrawn and
    is left out of the mixture's Dirichlet draw (local_clusters_actions.cpp:348, 364-370), index 1 never
    splits (250), the outlier never merges (295), and the outlier (and cluster 1 while K = 2) is never
    removed (409-410).  The data carry 2 % of points spread uniformly over the whole box; with one initial
    cluster the outlier cluster first takes every point (cluster 1's weight is 0) and the true clusters are
    split off it, the spread points staying in it."""
    D = 4
    X, gt, _ = gaussian_mixture(20_000, D, 3, seed=1, dtype=np.float64)
    rng = np.random.default_rng(0)
    idx = rng.choice(len(X), 400, replace=False)
    X[idx] = rng.uniform(-300, 300, size=(400, D))
    gt = gt.copy()
    gt[idx] = 3
    say(f"lockstep outlier mode: N=20000 D={D} FP64, init_clusters={init_clusters}")
    out = _lockstep(f"outlier K0={init_clusters}", X, gt, 3, 10.0, D + 3.0, iters=80, init_clusters=init_clusters,
                    outlier=(0.05, (1.0, np.zeros(D), D + 3.0, 100.0 * np.eye(D))), min_nmi=None)
    if out is not None:
        khist, lab, w = out
        assert w[0] == 0.05   # the outlier cluster's weight is outlier_mod, fixed
        if init_clusters == 1:   # the spread points end in the outlier cluster, the true clusters split off it
            assert khist[-1] == 4 and dp._nmi(gt, lab) > 0.95


@pytest.mark.parametrize("mode", ["device", "host"])
def test_cluster_cap_stops_splitting_and_run_continues(mode):
    """A context whose kernels take at most `cap` clusters (lowered with dpmm_limit_clusters): the
    sampler treats the cap as max_clusters (no splits from an iteration that starts at it) and
    accepts, within one iteration, only as many splits as fit, so K never exceeds the cap and the
    run finishes instead of failing in the next sweep."""
    D, cap = 8, 4
    X, gt, _ = gaussian_mixture(60_000, D, 8, seed=5, dtype=np.float32)
    say(f"cluster cap {cap} ({mode} decisions)")
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=4)
    dev.limit_clusters(cap)
    assert dev.max_clusters() == cap
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    smp = dp.Sampler(dev, hyper, 10.0, 60, seed=4, init_clusters=3, host_decisions=mode == "host")
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
