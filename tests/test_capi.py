"""CPU checks of the C ABI library: it loads, exports every entry point include/dpmm_hip.h declares,
and reports the absence of a GPU as an error code instead of exiting (no compute calls here)."""
import ctypes
import os
import re

import pytest

import dpmmsubclusters_gpu_amd as dp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "dpmm_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dpmm_\w+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "dpmm_create" in names and "dpmm_sample_cluster_labels" in names and "dpmm_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    L = dp.load_library()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_errors_are_codes_not_exits():
    L = dp.load_library()
    # null output pointer -> EINVAL, with a message
    assert L.dpmm_create(None, 10, 2, 0, 0, 0, None) == -1
    assert L.dpmm_last_error()
    assert L.dpmm_destroy(None) == 0


def test_python_mirror_validates_before_device():
    with pytest.raises(ValueError):
        dp.dp_parallel_sampling_class(5, [[1.0, 2.0]], 0, dp.prior_type.Gaussian)


def test_dropin_header_compiles():
    """The C++ drop-in header is self-contained C++17 over the C ABI header."""
    import subprocess
    root = os.path.join(os.path.dirname(__file__), "..")
    src = "#include \"dpmm/dp_parallel_sampling.hpp\"\nint main() { dpmm::niw_hyperparams h(1.0, std::vector<double>{0.0}, 3.0, std::vector<double>{1.0}); return h.m.size() == 1 ? 0 : 1; }\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        "-x", "c++", "-"], input=src, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
