"""The merge bound of the device-resident iteration (dpmm_sampler_set_merge_bound) against the exact merge
test of every candidate pair (local_clusters_actions.cpp:287-321 computes every pair's log ML): two
samplers with the same data and seed, one with the bound, run side by side.  Their decisions must be
identical (per-iteration cluster counts, counts per cluster, final labels), and the bound must reject
most pairs of well-separated clusters without their log ML."""
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


CASES = {
    # well separated: splits from one cluster, almost every pair hopeless
    "d32_separated": dict(N=120_000, D=32, K=10, iters=70, spread=100.0, init=None),
    # started over-split into 2 fragments per cluster: merges happen and must not be rejected
    "d32_oversplit": dict(N=100_000, D=32, K=8, iters=60, spread=100.0, init="oversplit", pieces=2),
    "d64_oversplit": dict(N=60_000, D=64, K=8, iters=50, spread=100.0, init="oversplit", pieces=3),
    # overlapping clusters over-split into 3: few pairs are hopeless (lockstep only)
    "d16_oversplit_close": dict(N=80_000, D=16, K=6, iters=60, spread=6.0, init="oversplit", pieces=3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_merge_bound_lockstep(case):
    c = CASES[case]
    N, D, K, iters = c["N"], c["D"], c["K"], c["iters"]
    X, gt, _ = gaussian_mixture(N, D, K, seed=41, spread=c["spread"], dtype=np.float32)
    kw = {}
    if c["init"] == "oversplit":
        rng = np.random.default_rng(5)
        pc = c["pieces"]
        kw = dict(init_labels=(pc * gt + rng.integers(0, pc, gt.size)).astype(np.int32))
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    runs = []
    for bound in (True, False):
        dev = dp.DeviceSweep(X, dp.prior_type.Gaussian, seed=6)
        smp = dp.Sampler(dev, hyper, 10.0, iters, seed=6, precision=dp.Precision.MIXED, **kw)
        smp.set_merge_bound(bound)
        runs.append((dev, smp))
    (db, sb), (de, se) = runs
    say(f"merge bound lockstep: {case}")
    for t in range(1, iters + 1):
        sb.step(t)
        se.step(t)
        assert sb.num_clusters() == se.num_clusters(), f"iteration {t}"
        np.testing.assert_array_equal(sb.weights()[1], se.weights()[1], err_msg=f"iteration {t}: counts")
    lb, ub = db.get_labels()
    le, ue = de.get_labels()
    np.testing.assert_array_equal(lb, le)
    np.testing.assert_array_equal(ub, ue)
    rej = sb.merge_bound_rejects()
    exact_b, exact_e = sb.screen_stats()[0], se.screen_stats()[0]
    print(f"{case}: K {sb.num_clusters()}, events {sb.events()}, pairs rejected by the bound {rej}, exact tests "
          f"{exact_b} (all {exact_e}), NMI {dp._nmi(gt, lb):.4f}")
    assert exact_b + rej == exact_e
    assert se.merge_bound_rejects() == 0
    if case == "d32_separated":
        assert rej > 0.9 * exact_e
    if c["init"] == "oversplit" and c["spread"] >= 100.0:
        ev = sb.events()
        assert ev["merges"] + ev["removed"] >= 1
    for d, s in runs:
        s.close()
        d.close()
