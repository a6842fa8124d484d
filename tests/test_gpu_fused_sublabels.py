"""The sub-label draw fused into the pruned mixed label sweep (dpmm_set_fused_sub_labels) against the
separate sub-label sweep (cudaKernel.cu:1299-1398 draws sub-labels in a pass of its own, after the
labels).  The fused sweep screens a point's sub-clusters in the label pass when its label stays the
cluster it was filed under, and draws every other point in FP32 by its final label; both forms use
the same Philox draw per point, so labels, sub-labels and statistics must be identical (bit for
bit), and sub-labels must equal the FP64 oracle's wherever its margin exceeds 1e-6."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture
import oracle_ctypes as orc
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _params(X, gt, K, rng, jitter):
    """Cluster parameters fitted to the generating labels, sub-clusters split along a random direction."""
    D = X.shape[1]
    mu = np.stack([X[gt == k].mean(0) for k in range(K)]).astype(np.float64)
    cov = [np.cov(X[gt == k].T.astype(np.float64)) + 1e-3 * np.eye(D) for k in range(K)]
    lam = np.stack([np.linalg.inv(c) for c in cov])
    ld = np.array([-np.linalg.slogdet(l)[1] for l in lam])
    w = np.bincount(gt, minlength=K) / len(gt)
    smu = np.empty((2 * K, D))
    for k in range(K):
        v = rng.normal(size=D)
        v *= jitter / np.linalg.norm(v)
        smu[2 * k], smu[2 * k + 1] = mu[k] + v, mu[k] - v
    return mu, lam, ld, w, smu, np.repeat(lam, 2, axis=0), np.repeat(ld, 2), np.full(2 * K, 0.5)


def _sweeps(X, params, init, fused, iters):
    dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=5)
    dev.set_precision(dp.Precision.MIXED)
    dev.set_label_pruning(True)
    dev.set_fused_sub_labels(fused)
    dev.set_gaussian_params(*params)
    dev.set_labels(init, None)
    K = len(params[0])
    out = []
    for t in iters:
        dev.sample_cluster_labels(t)
        dev.sample_sub_labels(t)
        lab, sub = dev.get_labels()
        Nsub, s, S = dev.sufficient_statistics(K)
        out.append((lab, sub, Nsub, s, S, dev.last_sub_recheck_count()))
    dev.close()
    return out


# (D, K, spread, jitter): D = 20 pads to 32; close clusters move labels between the two sweeps
CASES = [(32, 10, 100.0, 1.0), (64, 12, 100.0, 2.0), (64, 8, 6.0, 0.5), (20, 6, 8.0, 1.0), (48, 5, 100.0, 0.2)]


@pytest.mark.parametrize("D,K,spread,jitter", CASES)
def test_fused_equals_separate_and_oracle(D, K, spread, jitter):
    N = 50_000
    rng = np.random.default_rng(D * 7 + K)
    X, gt, _ = gaussian_mixture(N, D, K, seed=D + 3 * K, spread=spread, dtype=np.float32)
    params = _params(X, gt, K, rng, jitter)
    # start from the generating labels with 2 % of them scrambled, so that some labels move
    init = gt.astype(np.int32).copy()
    flip = rng.random(N) < 0.02
    init[flip] = rng.integers(0, K, int(flip.sum()))
    say(f"fused sub-labels D={D} K={K} spread={spread}")
    # two iterations each: the second sweeps the partition the first left, with its sub-labels
    fz = _sweeps(X, params, init, True, (3, 4))
    sp = _sweeps(X, params, init, False, (3, 4))
    for (lf, uf, nf, sf, Sf, rf), (ls, us, ns, ss, Ss, rs), t in zip(fz, sp, (3, 4)):
        np.testing.assert_array_equal(lf, ls, err_msg=f"iteration {t}: labels")
        np.testing.assert_array_equal(uf, us, err_msg=f"iteration {t}: sub-labels")
        np.testing.assert_array_equal(nf, ns)
        np.testing.assert_array_equal(sf, ss)
        np.testing.assert_array_equal(Sf, Ss)
        print(f"D={D} K={K} t={t}: FP32 sub-label draws fused {rf}, separate {rs} of {N}; "
              f"labels moved {(lf != (init if t == 3 else fz[0][0])).sum()}")
    # the oracle's sub-labels for the final labels of the last iteration
    lab, sub = fz[1][0], fz[1][1]
    osub, margin = orc.sublabel_sweep(0, X, lab, params[4], params[5], params[6], lr_weights=params[7], seed=5, it=4)
    bad = (sub != osub) & (margin >= 1e-6)
    assert not bad.any(), f"{bad.sum()} sub-labels differ from the oracle above margin 1e-6"


LOCKSTEP = {
    "d64_splits": dict(N=120_000, D=64, K=8, iters=60, init=None, graph=False),
    "d32_oversplit": dict(N=100_000, D=32, K=8, iters=60, init="oversplit", graph=False),
    "d64_oversplit_graph": dict(N=100_000, D=64, K=8, iters=50, init="oversplit", graph=True),
}


@pytest.mark.parametrize("case", sorted(LOCKSTEP))
def test_fused_sampler_lockstep(case):
    c = LOCKSTEP[case]
    N, D, K, iters = c["N"], c["D"], c["K"], c["iters"]
    X, gt, _ = gaussian_mixture(N, D, K, seed=31, a=0.5, ridge=1.0, dtype=np.float32)
    kw = {}
    if c["init"] == "oversplit":
        rng = np.random.default_rng(7)
        kw = dict(init_labels=(2 * gt + rng.integers(0, 2, gt.size)).astype(np.int32))
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    runs = []
    for fused in (True, False):
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=8)
        dev.set_label_pruning(True)
        dev.set_fused_sub_labels(fused)
        dev.set_graph_mode(c["graph"])
        smp = dp.Sampler(dev, hyper, 10.0, iters, seed=8, precision=dp.Precision.MIXED, **kw)
        if not c["graph"]:
            smp.trace_sweeps(True)
        runs.append((dev, smp))
    (df, sf), (ds, ss) = runs
    say(f"fused sub-labels lockstep: {case}")
    khist = []
    for t in range(1, iters + 1):
        sf.step(t)
        ss.step(t)
        if not c["graph"]:
            lf, uf = sf.sweep_trace()
            ls, us = ss.sweep_trace()
            np.testing.assert_array_equal(lf, ls, err_msg=f"iteration {t}: labels")
            np.testing.assert_array_equal(uf, us, err_msg=f"iteration {t}: sub-labels")
        assert sf.num_clusters() == ss.num_clusters(), f"iteration {t}"
        np.testing.assert_array_equal(sf.weights()[1], ss.weights()[1], err_msg=f"iteration {t}: counts")
        khist.append(sf.num_clusters())
    lf, uf = df.get_labels()
    ls, us = ds.get_labels()
    np.testing.assert_array_equal(lf, ls)
    np.testing.assert_array_equal(uf, us)
    print(f"{case}: K history {khist}, events {sf.events()}, NMI {dp._nmi(gt, lf):.4f}")
    for d, s in runs:
        s.close()
        d.close()
