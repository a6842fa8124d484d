"""The command-line path (main.cpp:6-129) end to end: dpmm_main with --model_path (an N x D .npy),
--params_path (a JSON params file with comments, global_params.cpp:8-70) and --result_path, reading
result.json back (labels 1-based, the model's weights, seconds per iteration).  The file path runs the
same sampler as the API: with the same settings and seed its labels equal the Python API's run.  The
params file's defaults are the reference's (outlier_mod 0.05 turns the outlier cluster on), and a
failure is reported in result.json with exit status 1."""
import json
import os
import subprocess

import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from dpmmsubclusters_gpu_amd.synthetic import gaussian_mixture, multinomial_mixture
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "..", "dpmmsubclusters_gpu_amd", "lib", "dpmm_main")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")
    if not os.path.exists(BIN):
        pytest.fail("dpmm_main not built (dpmmsubclusters_gpu_amd.build)")


def _run(tmp_path, X, params, prior="Gaussian", name="run"):
    npy = tmp_path / f"{name}.npy"
    np.save(npy, X)
    pj = tmp_path / f"{name}.json"
    pj.write_text("// written by the test\n" + json.dumps(params, indent=1))
    res = tmp_path / f"{name}_result.json"
    say(f"dpmm_main {name}: N={X.shape[0]} D={X.shape[1]} prior={prior}")
    r = subprocess.run([BIN, f"--params_path={pj}", f"--model_path={tmp_path / name}", f"--prior_type={prior}",
                        f"--result_path={res}"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    return r.returncode, json.loads(res.read_text())


def test_cli_gaussian_equals_api_run(tmp_path):
    X, gt, _ = gaussian_mixture(50_000, 2, 2, seed=4, dtype=np.float64)
    params = {"iterations": 100, "alpha": 1.0, "burnout_period": 15, "outlier_mod": 0, "randomSeed": 2,
              "hyper_params": {"k": 1.0, "m": [0.0, 0.0], "v": 5.0, "psi": [[1.0, 0.0], [0.0, 1.0]]}}
    rc, res = _run(tmp_path, X, params)
    assert rc == 0
    labels = np.array(res["labels"])
    assert labels.shape == (50_000,) and labels.min() >= 1
    assert len(res["iter_count"]) == 100
    # the same run through the API (data D x N, the file's settings, the same seed)
    model = dp.dp_parallel_sampling_class(X.shape[0], X.T, 2, dp.prior_type.Gaussian)
    mi = model.dp_parallel(dp.niw_hyperparams(1.0, np.zeros(2), 5.0, np.eye(2)), 1.0, iters=100, verbose=False,
                           burnout=15)
    np.testing.assert_array_equal(labels, mi.labels)
    np.testing.assert_allclose(res["weights"], mi.weights, rtol=1e-15)
    print("clusters", labels.max(), "NMI", dp._nmi(gt, labels))


def test_cli_defaults_turn_the_outlier_cluster_on(tmp_path):
    """No outlier_mod in the file: the reference's default 0.05 (global_params.h:54) with the default
    outlier prior niw(1, 0, 5, I): cluster 0 is the outlier cluster, weight 0.05."""
    X, gt, _ = gaussian_mixture(20_000, 2, 2, seed=5, dtype=np.float64)
    params = {"iterations": 60, "alpha": 1.0, "randomSeed": 3}
    rc, res = _run(tmp_path, X, params, name="defaults")
    assert rc == 0
    assert res["weights"][0] == 0.05
    assert len(res["iter_count"]) == 60


def test_cli_multinomial(tmp_path):
    X, gt, _ = multinomial_mixture(20_000, 40, 4, seed=6, words=60, concentration=0.1)
    params = {"iterations": 80, "alpha": 10.0, "outlier_mod": 0.0, "burnout_period": 15, "randomSeed": 5,
              "hyper_params": {"alpha": [1.0] * 40}}
    rc, res = _run(tmp_path, X.astype(np.int32), params, prior="Multinomial", name="mult")
    assert rc == 0
    labels = np.array(res["labels"])
    print("clusters", labels.max(), "NMI", dp._nmi(gt, labels))
    assert labels.max() == 4 and dp._nmi(gt, labels) > 0.99


def test_cli_reports_errors_in_the_result_file(tmp_path):
    res = tmp_path / "err.json"
    say("dpmm_main with a missing model file")
    r = subprocess.run([BIN, f"--model_path={tmp_path / 'missing'}", f"--params_path={tmp_path / 'none.json'}",
                        f"--result_path={res}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    out = json.loads(res.read_text())
    assert "missing.npy" in out["error"][0]
