"""The C++ drop-in header (include/dpmm/dp_parallel_sampling.hpp) used as a reference user would: the
README example (2-D Gaussian mixture, niw_hyperparams(1, 0, 5, I), alpha 1, 400 iterations) and a
small multinomial mixture through dp_parallel_sampling_class::dp_parallel, 1-based labels,
ModelInfo histories and weights."""
import os
import re
import subprocess

import pytest

from helpers import say

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "..", "dpmmsubclusters_gpu_amd", "lib", "dropin_readme")


def test_dropin_readme_example():
    if not os.path.exists(BIN):
        pytest.fail("drop-in example not built (dpmmsubclusters_gpu_amd.build)")
    say("C++ drop-in README example")
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    g = re.search(r"gaussian K=(\d+) nmi=([\d.]+) labels=(\d+)\.\.(\d+) weights=([\d.]+) iters=(\d+)", out.stdout)
    m = re.search(r"multinomial K=(\d+) nmi=([\d.]+) labels=(\d+)\.\.(\d+) iters=(\d+)", out.stdout)
    assert g and m
    assert int(g.group(1)) == 2 and float(g.group(2)) > 0.99
    assert (int(g.group(3)), int(g.group(4))) == (1, 2)   # 1-based labels
    assert 0.0 < float(g.group(5)) <= 1.0 and int(g.group(6)) == 400
    assert "nmi_history=0" in out.stdout   # the reference records NMI only in verbose mode
    assert int(m.group(1)) == 3 and float(m.group(2)) > 0.99 and int(m.group(5)) == 80
