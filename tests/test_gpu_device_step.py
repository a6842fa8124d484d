"""The device-resident iteration against the host-decision path.

In the default mode every iteration runs on the device from the parameter draw to the relabel
passes: the sampling of the weights and the burn-out history (sample_clusters,
local_clusters_actions.cpp:341-371), the reset, split and merge decisions with the greedy merge
resolution and the removal of empty clusters (reset_bad_clusters, check_and_split, check_and_merge,
remove_empty_clusters; local_clusters_actions.cpp:244-321, 400-447) run in device kernels, and the
host synchronises once per iteration.  With host_decisions the host takes those decisions between
device launches.  Both read the same Philox counters and evaluate the same expressions, so two
samplers with the same data and seed, one per mode, run in lockstep in one process: after every
iteration the sweeps' labels and sub-labels, the cluster count, the counts and the decision totals
must be identical, the mixture weights within 1e-12 relative (their gamma draws' transcendentals
come from the device and the host math libraries, which may differ in the last bit), and at the end
the labels and sub-labels bit for bit.  The cases cover splits from one cluster, merges from an
over-split start, resets (a cluster of a single point), the cluster cap, mixed precision and the
multinomial prior."""
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


def _pair(X, prior, hyper, alpha, iters, seed, prec, cap=None, **kw):
    out = []
    for host in (False, True):
        dev = dp.DeviceSweep(X, prior, seed=seed)
        if cap is not None:
            dev.limit_clusters(cap)
        smp = dp.Sampler(dev, hyper, alpha, iters, seed=seed, precision=prec, host_decisions=host, **kw)
        smp.trace_sweeps(True)
        out.append((dev, smp))
    return out


def _run_lockstep(name, runs, iters):
    (dd, sd), (dh, sh) = runs
    khist = []
    for t in range(1, iters + 1):
        sd.step(t)
        sh.step(t)
        ld, ud = sd.sweep_trace()
        lh, uh = sh.sweep_trace()
        np.testing.assert_array_equal(ld, lh, err_msg=f"{name}: iteration {t}: sweep labels")
        np.testing.assert_array_equal(ud, uh, err_msg=f"{name}: iteration {t}: sweep sub-labels")
        assert sd.num_clusters() == sh.num_clusters(), f"{name}: iteration {t}: cluster counts differ"
        wd, nd = sd.weights()
        wh, nh = sh.weights()
        np.testing.assert_array_equal(nd, nh, err_msg=f"{name}: iteration {t}: counts")
        np.testing.assert_allclose(wd, wh, rtol=1e-12, atol=0, err_msg=f"{name}: iteration {t}: weights")
        assert sd.events() == sh.events(), f"{name}: iteration {t}: {sd.events()} vs {sh.events()}"
        khist.append(sd.num_clusters())
    ld, ud = dd.get_labels()
    lh, uh = dh.get_labels()
    np.testing.assert_array_equal(ld, lh)
    np.testing.assert_array_equal(ud, uh)
    ev = sd.events()
    print(f"{name}: identical for {iters} iterations; K history {khist}; events {ev}")
    lp_d, lp_h = sd.log_posterior(), sh.log_posterior()
    assert lp_d == pytest.approx(lp_h, rel=1e-12)
    for d, s in runs:
        s.close()
        d.close()
    return khist, ev, ld


def test_device_step_fp64_readme_splits():
    """The README example (module_tests.cpp:19-49): one initial cluster split into the true two."""
    X, gt, _ = gaussian_mixture(100_000, 2, 2, seed=0, dtype=np.float64)
    say("device step vs host decisions: readme_d2 FP64")
    hyper = dp.niw_hyperparams(1.0, np.zeros(2), 5.0, np.eye(2))
    khist, ev, lab = _run_lockstep("readme_d2", _pair(X, dp.prior_type.Gaussian, hyper, 1.0, 100, 2,
                                                      dp.Precision.FP64), 100)
    assert ev["splits"] >= 1 and khist[-1] == 2 and dp._nmi(gt, lab) > 0.99


def test_device_step_merges_resets_from_an_oversplit_start():
    """Ten true clusters (D = 64, the C5 mixture's covariances) labelled as twenty, each true cluster cut
    in two at random, plus a cluster of a single point: fragments merge back once their burn-out allows,
    others empty and are removed, and the single point's cluster is reset (a sub-cluster is empty)."""
    D, K = 64, 10
    X, gt, _ = gaussian_mixture(100_000, D, K, seed=11, a=0.5, ridge=1.0, dtype=np.float32)
    rng = np.random.default_rng(3)
    init = (2 * gt + rng.integers(0, 2, gt.size)).astype(np.int32)
    init[0] = 2 * K   # a cluster of one point
    say("device step vs host decisions: over-split start, D=64 mixed")
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    khist, ev, lab = _run_lockstep("oversplit", _pair(X, dp.prior_type.Gaussian, hyper, 10.0, 80, 7,
                                                      dp.Precision.MIXED, init_labels=init), 80)
    assert ev["merges"] >= 1 and ev["resets"] >= 1 and ev["removed"] >= 1
    assert khist[0] > khist[-1]


def test_device_step_mixed_precision():
    """The mixed-precision sweeps (BF16 MFMA screens with FP32 rechecks), D = 32, twelve clusters."""
    N, D, K = 150_000, 32, 12
    X, gt, _ = gaussian_mixture(N, D, K, seed=8, a=1.0, ridge=0.5, weight=0.5, dtype=np.float32)
    say("device step vs host decisions: mixed D=32")
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    khist, ev, lab = _run_lockstep("mixed_d32", _pair(X, dp.prior_type.Gaussian, hyper, 10.0, 60, 5,
                                                      dp.Precision.MIXED), 60)
    assert ev["splits"] >= 3 and khist[-1] == ev["splits"] + 1


def test_device_step_cluster_cap():
    """A cap of four clusters: splits beyond the room left are dropped, the highest indices first."""
    D = 8
    X, _, _ = gaussian_mixture(60_000, D, 8, seed=5, dtype=np.float32)
    say("device step vs host decisions: cluster cap 4")
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    khist, ev, _ = _run_lockstep("cap4", _pair(X, dp.prior_type.Gaussian, hyper, 10.0, 60, 4, dp.Precision.FP32,
                                               cap=4, init_clusters=3), 60)
    assert max(khist) == 4


def test_device_step_multinomial():
    N, D, K = 40_000, 100, 8
    X, gt, _ = multinomial_mixture(N, D, K, seed=6, words=80, concentration=0.1)
    say("device step vs host decisions: multinomial")
    hyper = dp.multinomial_hyper(np.ones(D))
    khist, ev, lab = _run_lockstep("mult", _pair(X, dp.prior_type.Multinomial, hyper, 10.0, 60, 3,
                                                 dp.Precision.FP32), 60)
    assert ev["splits"] >= 1


def test_switch_to_host_decisions_mid_run():
    """A sampler that leaves the device step (a merge screener is set after some iterations) carries its
    cluster table back to the host: it continues exactly as a sampler that decided on the host all along."""
    D = 8
    X, _, _ = gaussian_mixture(40_000, D, 5, seed=9, dtype=np.float32)
    say("device step, then host decisions")
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    runs = _pair(X, dp.prior_type.Gaussian, hyper, 10.0, 60, 6, dp.Precision.FP32)
    (dd, sd), (dh, sh) = runs
    for t in range(1, 61):
        if t == 31:
            sd.set_screener(None, log_pairs=True)   # pair logging keeps the decisions on the host
            sh.set_screener(None, log_pairs=True)
        sd.step(t)
        sh.step(t)
        assert sd.num_clusters() == sh.num_clusters(), f"iteration {t}"
        np.testing.assert_array_equal(sd.weights()[1], sh.weights()[1], err_msg=f"iteration {t}")
    np.testing.assert_array_equal(dd.get_labels()[0], dh.get_labels()[0])
    for d, s in runs:
        s.close()
        d.close()
