"""Shared test helpers: random cluster parameters and GPU gating."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def have_gpu():
    try:
        import dpmmsubclusters_gpu_amd as d
        import ctypes
        L = d.load_library()
        c = ctypes.c_int(0)
        return L.dpmm_device_count(ctypes.byref(c)) == 0 and c.value > 0
    except Exception:
        return False


def say(msg):
    """Name the case before each GPU step, flushed, so a fault can be attributed."""
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def random_spd(rng, D, scale=1.0, ridge=0.2):
    A = rng.normal(size=(D, D)) * scale
    return A @ A.T / D + ridge * np.eye(D)


def gaussian_params(rng, X, K, sub=False):
    """Cluster parameters in the reference's mv_gaussian form around data points."""
    N, D = X.shape
    n = 2 * K if sub else K
    idx = rng.choice(N, size=n, replace=False)
    mu = X[idx].astype(np.float64) + rng.normal(size=(n, D))
    lam = np.stack([np.linalg.inv(random_spd(rng, D, 3.0, 1.0) * 4.0) for _ in range(n)])
    logdet = np.array([-np.linalg.slogdet(l)[1] for l in lam])
    return mu, lam, logdet


def chol_lower(lam):
    return np.stack([np.linalg.cholesky(l) for l in lam])
