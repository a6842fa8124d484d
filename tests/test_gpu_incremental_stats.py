"""Statistics updated from the moved points (dpmm_set_statistics_refresh) against statistics recomputed
from every point (create_sufficient_statistics, cudaKernel.cu:844-941, which the reference runs on every
point each iteration).

Behind a pruned mixed label sweep that moves no label, the partition of the labels stands and the
statistics of an iteration are the last ones plus the sums of the points whose sub-label moved.  Two
samplers with the same data and seed run side by side, one with updates (a full recomputation every 8
passes) and one recomputing every pass.  They take the same draws, so after every iteration their
sweeps' labels and sub-labels and their counts must be equal, and the updated statistics must equal
the recomputed ones within 1e-12 of the largest entry of each sub-cluster's sums and scatter matrix
(counts exactly).  A cluster that a moved label touched is recomputed in the same pass; most passes
must be updates."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _close(a, b, name):
    for q in range(a.shape[0]):
        scale = max(np.abs(b[q]).max(), 1.0)
        err = np.abs(a[q] - b[q]).max()
        assert err <= 1e-12 * scale, f"{name}, entry {q}: {err:.3e} against {scale:.3e}"


CASES = {
    # splits from one cluster (labels move on the iterations that split)
    "d32_splits": dict(N=200_000, D=32, K=10, iters=80, a=1.0, init=None, host=False),
    "d128_splits": dict(N=60_000, D=128, K=6, iters=70, a=0.5, init=None, host=False),
    # an over-split start plus a cluster of one point: merges, removals and resets (a reset redraws a
    # cluster's sub-labels, and the next pass updates it from its moved sub-labels), with the decisions
    # on the device and on the host
    "d64_oversplit": dict(N=100_000, D=64, K=10, iters=80, a=0.5, init="oversplit", host=False),
    "d64_oversplit_host": dict(N=100_000, D=64, K=10, iters=80, a=0.5, init="oversplit", host=True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_updated_statistics_equal_recomputed(case):
    c = CASES[case]
    N, D, K, iters = c["N"], c["D"], c["K"], c["iters"]
    X, gt, _ = gaussian_mixture(N, D, K, seed=21, a=c["a"], ridge=1.0, dtype=np.float32)
    kw = {}
    if c["init"] == "oversplit":
        rng = np.random.default_rng(3)
        init = (2 * gt + rng.integers(0, 2, gt.size)).astype(np.int32)
        init[0] = 2 * K   # a cluster of one point
        kw = dict(init_labels=init)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    runs = []
    for every in (8, 1):
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=4)
        dev.set_label_pruning(True)
        dev.set_statistics_refresh(every)
        smp = dp.Sampler(dev, hyper, 10.0, iters, seed=4, precision=dp.Precision.MIXED, host_decisions=c["host"], **kw)
        smp.trace_sweeps(True)
        runs.append((dev, smp))
    (du, su), (df, sf) = runs
    say(f"updated vs recomputed statistics: {case}")
    updates, khist = [], []
    for t in range(1, iters + 1):
        k_before, ev_before = su.num_clusters(), su.events()
        su.step(t)
        sf.step(t)
        upd = du.last_statistics_update()
        assert not df.last_statistics_update()
        lu, uu = su.sweep_trace()
        lf, uf = sf.sweep_trace()
        np.testing.assert_array_equal(lu, lf, err_msg=f"iteration {t}: labels")
        np.testing.assert_array_equal(uu, uf, err_msg=f"iteration {t}: sub-labels")
        K = su.num_clusters()
        assert K == sf.num_clusters()
        np.testing.assert_array_equal(su.weights()[1], sf.weights()[1], err_msg=f"iteration {t}: counts")
        if upd and su.events() == ev_before:
            # no decision relabelled the points: the pass's statistics are still the current ones
            Nu, su_, Su = du.sufficient_statistics(k_before)
            Nf, sf_, Sf = df.sufficient_statistics(k_before)
            np.testing.assert_array_equal(Nu, Nf, err_msg=f"iteration {t}: sub-cluster counts")
            _close(su_, sf_, f"iteration {t}: sums")
            _close(Su, Sf, f"iteration {t}: scatter")
        updates.append(upd)
        khist.append(K)
    n_upd = sum(u != 0 for u in updates)
    print(f"{case}: {n_upd} of {iters} passes were updates ({sum(u == 2 for u in updates)} with clusters a moved "
          f"label touched recomputed); K history {khist}; NMI {dp._nmi(gt, lu):.4f}")
    assert n_upd >= iters // 3
    if c["init"] == "oversplit":
        ev = su.events()
        print("events", ev)
        assert ev["resets"] >= 1 and ev["merges"] + ev["removed"] >= 1
    for d, s in runs:
        s.close()
        d.close()


def test_update_of_a_cluster_whose_segment_moved_later():
    """Clusters 0 and 2 overlap, so every noisy sweep moves labels between them; cluster 1 lies far away
    and keeps its points.  When cluster 0 gains points, cluster 1's segment of the new partition starts
    later than in the partition swept (its offset into the previous partition is negative), and its
    statistics are still updated from its moved sub-labels.  Fixed parameters, the same seed for the
    updating and the recomputing context: counts equal, sums and scatter within 1e-12."""
    N, D = 60_000, 32
    rng = np.random.default_rng(17)
    means = np.zeros((3, D))
    means[1, :] = 40.0
    means[2, 0] = 0.3   # overlaps cluster 0
    gt = rng.integers(0, 3, N)
    X = (means[gt] + rng.normal(size=(N, D))).astype(np.float32)
    K = 3
    mu = np.stack([X[gt == k].mean(0) for k in range(K)]).astype(np.float64)
    lam = np.stack([np.linalg.inv(np.cov(X[gt == k].T.astype(np.float64))) for k in range(K)])
    ld = np.array([-np.linalg.slogdet(l)[1] for l in lam])
    w = np.bincount(gt, minlength=K) / N
    smu = np.empty((2 * K, D))
    for k in range(K):
        v = rng.normal(size=D)
        v *= 0.5 / np.linalg.norm(v)
        smu[2 * k], smu[2 * k + 1] = mu[k] + v, mu[k] - v
    params = (mu, lam, ld, w, smu, np.repeat(lam, 2, axis=0), np.repeat(ld, 2), np.full(2 * K, 0.5))
    devs = []
    for every in (8, 1):
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=9)
        dev.set_precision(dp.Precision.MIXED)
        dev.set_label_pruning(True)
        dev.set_statistics_refresh(every)
        dev.set_gaussian_params(*params)
        dev.set_labels(gt.astype(np.int32), None)
        devs.append(dev)
    du, df = devs
    say("updated statistics: a cluster whose segment moved later")
    later = 0
    prev0 = None
    for t in range(1, 9):
        for d in devs:
            d.sample_cluster_labels(t)
            d.sample_sub_labels(t)
        Nu, su_, Su = du.sufficient_statistics(K)
        Nf, sf_, Sf = df.sufficient_statistics(K)
        lu, uu = du.get_labels()
        lf, uf = df.get_labels()
        np.testing.assert_array_equal(lu, lf, err_msg=f"iteration {t}: labels")
        np.testing.assert_array_equal(uu, uf, err_msg=f"iteration {t}: sub-labels")
        np.testing.assert_array_equal(Nu, Nf, err_msg=f"iteration {t}: sub-cluster counts")
        _close(su_, sf_, f"iteration {t}: sums")
        _close(Su, Sf, f"iteration {t}: scatter")
        n0 = int((lu == 0).sum())
        if prev0 is not None and n0 > prev0 and du.last_statistics_update() == 2:
            later += 1   # cluster 1 was updated with a negative offset
        prev0 = n0
    print(f"passes that updated cluster 1 after cluster 0 grew: {later}")
    assert later >= 1
    for d in devs:
        d.close()
