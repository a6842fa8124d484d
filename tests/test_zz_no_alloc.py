"""dpmm_label_contingency_static, the allocation-free variant of dpmm_label_contingency, called
through the C ABI on device arrays and checked against the allocating entry point: counts are
integers, so the tables must be identical.  Two calls in a row, the second on fewer points and a
smaller table, so that anything the first call left in the shared scratch array would show."""
import ctypes

import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from helpers import have_gpu, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not have_gpu():
        pytest.fail("no GPU visible to the HIP library")


def _hip():
    """The HIP runtime the library itself loaded (the test allocates its device arrays with it)."""
    dp.load_library()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln and "/torch/" not in ln}
    assert paths, "libamdhip64 not loaded"
    h = ctypes.CDLL(sorted(paths)[0])
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _dev_array(h, a):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
    return p


def _static_table(L, labels, classes, K, C):
    h = _hip()
    dl, dc = _dev_array(h, labels), _dev_array(h, classes)
    out = np.full(K * C, -1, np.int64)
    dt = _dev_array(h, out)
    rc = L.dpmm_label_contingency_static(dl, dc, ctypes.c_int64(labels.size), K, C, dt, None)
    assert rc == 0, L.dpmm_last_error()
    assert h.hipDeviceSynchronize() == 0
    assert h.hipMemcpy(out.ctypes.data, dt, out.nbytes, 2) == 0   # device to host
    for p in (dl, dc, dt):
        h.hipFree(p)
    return out.reshape(K, C)


def test_static_contingency_equals_allocating_one():
    L = dp.load_library()
    L.dpmm_label_contingency_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(7)
    for N, K, C in ((1_000_003, 120, 130), (20_011, 7, 5)):
        say(f"static contingency N={N} K={K} C={C}")
        labels = rng.integers(-1, K + 2, N).astype(np.int32)   # some outside [0, K): left out by both
        classes = rng.integers(0, C, N).astype(np.int32)
        classes[0] = C - 1
        dev = dp.DeviceSweep(np.zeros((N, 2), np.float32), dp.prior_type.Gaussian, seed=1)
        dev.set_labels(labels, None)
        dev.set_ground_truth(classes)
        ref = dev.label_contingency(K, C)
        dev.close()
        got = _static_table(L, labels, classes, K, C)
        np.testing.assert_array_equal(got, ref)
        inside = (labels >= 0) & (labels < K)
        assert got.sum() == inside.sum()


def test_static_contingency_rejects_oversize_tables():
    L = dp.load_library()
    L.dpmm_label_contingency_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    dummy = ctypes.c_void_p(16)   # never dereferenced: the size check comes first
    assert L.dpmm_label_contingency_static(dummy, dummy, 10, 200, 100, dummy, None) == -1
    assert b"16384" in L.dpmm_last_error()
