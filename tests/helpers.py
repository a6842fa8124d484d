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


def adversarial_mixture(N, D, K, seed):
    """Clusters that defeat a screen whose error is taken proportional to |W (x - mu)|^2: covariances
    with one dominant direction (variance 400), a rank D/4 factor and a ridge of 1e-3 (condition
    number ~4e5), and means that differ only along the high-variance directions, so the clusters
    overlap (near-ties) while y = W (x - mu) comes from heavy cancellation of large terms.
    Returns float32 points, generating labels, means, covariances."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=D)
    v /= np.linalg.norm(v)
    B = rng.normal(size=(D, D // 4))
    base = rng.uniform(-5, 5, D)
    mus, sig = [], []
    for _ in range(K):
        R = rng.normal(scale=0.05, size=(D, D // 8))
        sig.append(400.0 * np.outer(v, v) + B @ B.T + R @ R.T + 1e-3 * np.eye(D))
        mus.append(base + B @ rng.normal(scale=0.15, size=D // 4) + v * rng.normal(scale=2.0))
    lab = rng.integers(0, K, N)
    X = np.empty((N, D))
    for k in range(K):
        i = lab == k
        X[i] = mus[k] + rng.normal(size=(int(i.sum()), D)) @ np.linalg.cholesky(sig[k]).T
    return X.astype(np.float32), lab.astype(np.int32), np.array(mus), np.array(sig)
