"""Graph mode: every iteration recorded into a hipGraph and replayed through one executable graph,
updated in place.  With the decisions on the device the graph holds the whole iteration (draw, sweeps,
statistics, log MLs, decisions, relabel passes, the summary copy); with host decisions, the sweep
segment (label sweep, sub-label sweep, statistics).  Two samplers with the same data and seed run in
lockstep in one process, one eager and one in graph mode, so every replay runs between the other
sampler's eager launches and the host synchronisations of its own iteration; after every iteration
both must hold the same clusters, weights and counts, and at the end the same labels and sub-labels,
bit for bit."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture, multinomial_mixture
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _lockstep(make_dev, hyper, alpha, iters, prec, prune=None, host_decisions=False):
    runs = []
    for graph in (False, True):
        dev = make_dev()
        if prune is not None:
            dev.set_label_pruning(prune)
        dev.set_graph_mode(graph)
        smp = dp.Sampler(dev, hyper, alpha, iters, seed=5, precision=prec, host_decisions=host_decisions)
        runs.append((dev, smp))
    for t in range(1, iters + 1):
        for _, smp in runs:
            smp.step(t)
        (d0, s0), (d1, s1) = runs
        assert s0.num_clusters() == s1.num_clusters(), f"iteration {t}"
        w0, n0 = s0.weights()
        w1, n1 = s1.weights()
        np.testing.assert_array_equal(n0, n1, err_msg=f"counts, iteration {t}")
        np.testing.assert_array_equal(w0, w1, err_msg=f"weights, iteration {t}")
    (d0, s0), (d1, s1) = runs
    l0, u0 = d0.get_labels()
    l1, u1 = d1.get_labels()
    np.testing.assert_array_equal(l0, l1)
    np.testing.assert_array_equal(u0, u1)
    inst, upd = d1.graph_counts()
    print(f"K={s1.num_clusters()} graphs instantiated {inst}, updated in place {upd}; events {s1.events()}")
    assert s0.events() == s1.events()
    assert inst + upd == iters
    assert upd > 0
    for d, s in runs:
        s.close()
        d.close()


@pytest.mark.parametrize("mode", ["device", "host"])
@pytest.mark.parametrize("prec,prune", [("MIXED", True), ("MIXED", False), ("FP32", None)])
def test_graph_mode_gaussian_matches_eager(prec, prune, mode):
    N, D, K = 150_000, 32, 12
    X, _, _ = gaussian_mixture(N, D, K, seed=8, a=1.0, ridge=0.5, weight=0.5, dtype=np.float32)
    say(f"graph mode gaussian {prec} prune={prune} ({mode} decisions)")
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    _lockstep(lambda: dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=5), hyper, 10.0, 50,
              getattr(dp.Precision, prec), prune, host_decisions=mode == "host")


@pytest.mark.parametrize("mode", ["device", "host"])
def test_graph_mode_multinomial_matches_eager(mode):
    N, D, K = 60_000, 200, 15
    X, _, _ = multinomial_mixture(N, D, K, seed=2)
    say(f"graph mode multinomial ({mode} decisions)")
    hyper = dp.multinomial_hyper(np.ones(D))
    _lockstep(lambda: dp.DeviceSweep(X, dp.prior_type.Multinomial, seed=5), hyper, 10.0, 40, dp.Precision.MIXED,
              host_decisions=mode == "host")
