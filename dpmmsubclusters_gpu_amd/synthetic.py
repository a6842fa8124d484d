"""Seeded synthetic mixtures shaped like the benchmark configurations.

gaussian_mixture: means ~ U(-spread, spread)^D, covariances A A^T + c I with A_ij ~ N(0, a^2),
mixture weights equal or ~ Dirichlet(w * 1); multinomial_mixture: theta_k ~ Dirichlet(0.1 * 1_D),
a uniform topic per document, `words` draws from theta_k, stored as int32 counts.
"""
import numpy as np

CONFIGS = {
    # name: (N, D, K, a, ridge, dirichlet weight (None = equal), dtype)
    "c1_readme_2d": dict(N=100_000, D=2, K=2, a=1.0, ridge=1.0, weight=None, dtype=np.float64),
    "c2_d32": dict(N=1_000_000, D=32, K=20, a=1.0, ridge=1.0, weight=1.0, dtype=np.float32),
    "c3_d64": dict(N=10_000_000, D=64, K=50, a=1.0, ridge=0.5, weight=0.5, dtype=np.float32),
    "c5_d128": dict(N=5_000_000, D=128, K=200, a=0.5, ridge=1.0, weight=None, dtype=np.float32),
}


def gaussian_mixture(N, D, K, seed=0, a=1.0, ridge=1.0, weight=None, dtype=np.float32, spread=100.0,
                     chunk=1 << 20):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-spread, spread, size=(K, D))
    facs = []
    for _ in range(K):
        A = rng.normal(0.0, a, size=(D, D))
        C = A @ A.T + ridge * np.eye(D)
        facs.append(np.linalg.cholesky(C))
    w = np.full(K, 1.0 / K) if weight is None else rng.dirichlet(np.full(K, weight))
    gt = rng.choice(K, size=N, p=w).astype(np.int32)
    X = np.empty((N, D), dtype=dtype)
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        z = rng.standard_normal((e - s, D))
        g = gt[s:e]
        out = np.empty((e - s, D))
        for k in np.unique(g):
            m = g == k
            out[m] = z[m] @ facs[k].T + means[k]
        X[s:e] = out.astype(dtype)
    return X, gt, means


def multinomial_mixture(N, D, K, seed=0, words=100, concentration=0.1):
    rng = np.random.default_rng(seed)
    theta = rng.dirichlet(np.full(D, concentration), size=K)
    gt = rng.integers(0, K, N).astype(np.int32)
    X = np.empty((N, D), dtype=np.int32)
    for k in range(K):
        m = gt == k
        X[m] = rng.multinomial(words, theta[k], size=int(m.sum()))
    return X, gt, theta


def config_mixture(name, seed=0, N=None):
    c = dict(CONFIGS[name])
    if N is not None:
        c["N"] = N
    return gaussian_mixture(c["N"], c["D"], c["K"], seed=seed, a=c["a"], ridge=c["ridge"], weight=c["weight"],
                            dtype=c["dtype"])
