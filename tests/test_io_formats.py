"""The file formats of the command-line path, on the CPU: .npy arrays of every dtype the reference's
loader takes (utils.cpp:12-57: 1-, 2-, 4- and 8-byte words), in C and Fortran order, and JSON params
files with jsoncpp's comments (global_params.cpp:8-24 reads them with collectComments), read by
include/dpmm/io.hpp and compared with numpy's and Python's own readers."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def io_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "io_check")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "io_check.cpp"), "-o", exe], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("dtype", ["<f8", "<f4", "<i4", "<i8", "<i2", "|u1", "<u2"])
@pytest.mark.parametrize("order", ["C", "F"])
def test_npy_reader(io_check, tmp_path, dtype, order):
    rng = np.random.default_rng(3)
    a = (rng.normal(size=(7, 3)) * 40).astype(dtype) if dtype[1] == "f" else rng.integers(0, 200, (7, 3)).astype(dtype)
    a = np.asfortranarray(a) if order == "F" else np.ascontiguousarray(a)
    f = tmp_path / "a.npy"
    np.save(f, a)
    out = subprocess.run([io_check, "npy", str(f)], capture_output=True, text=True, timeout=60).stdout.split("\n")
    rows, cols, descr = out[0].split()
    assert (int(rows), int(cols)) == a.shape and descr == np.dtype(dtype).str
    got = np.array([float(x) for x in out[1:1 + a.size]]).reshape(a.shape)
    np.testing.assert_array_equal(got, a.astype(np.float64))


def test_npy_reader_rejects_3d(io_check, tmp_path):
    f = tmp_path / "b.npy"
    np.save(f, np.zeros((2, 2, 2)))
    r = subprocess.run([io_check, "npy", str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "2-D" in r.stdout


def test_json_reader_with_comments(io_check, tmp_path):
    text = """{
      "iterations": 200,   // a line comment
      "alpha": 10.0, "use_verbose": true, "outlier_mod": 0,
      /* a block
         comment */
      "hyper_params": {
        //Example for the Gaussian prior
        "k": 1.0, "m": [0.0, -1.5e-3], "v": 5, "psi": [[1.0, 0.25], [0.25, 2.0]]
      },
      "gt": [1, 2, 2], "name": "a \\"quoted\\" path\\\\x", "nothing": null
    }"""
    f = tmp_path / "p.json"
    f.write_text(text)
    out = subprocess.run([io_check, "json", str(f)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout
    want = json.loads("\n".join(l.split("//")[0] for l in text.replace("/* a block\n         comment */", "")
                                .split("\n")))
    assert json.loads(out.stdout) == want


def test_json_reader_reports_errors(io_check, tmp_path):
    f = tmp_path / "bad.json"
    f.write_text('{"alpha": 10,, }')
    r = subprocess.run([io_check, "json", str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "json:" in r.stdout
