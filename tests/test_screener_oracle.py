"""The screener's CPU oracle pinned by mathematics: the MLP gradient against central finite
differences of its loss, the forward pass against a direct numpy evaluation, and the pair
features against their definition's invariants (identical clusters, swapped pairs)."""
import numpy as np

import oracle_ctypes as orc

F, H = 16, 128


def _unpack(p):
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    return W1, b1, W2, b2, w3, p[o]


def _params(rng):
    n = orc.mlp_param_count()
    assert n == H * F + H + H * H + H + H + 1
    p = rng.normal(size=n)
    p[:H * F] *= 0.4
    p[H * F + H:H * F + H + H * H] *= 0.12
    return p


def test_forward_matches_numpy():
    rng = np.random.default_rng(0)
    p = _params(rng)
    X = rng.normal(size=(9, F))
    W1, b1, W2, b2, w3, b3 = _unpack(p)
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    np.testing.assert_allclose(orc.mlp_forward(p, X), h2 @ w3 + b3, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    X = rng.normal(size=(7, F))
    y = np.array([1, 0, 0, 1, 0, 1, 0], float)
    loss, g = orc.mlp_loss_grad(p, X, y, pos_weight=3.0)
    idx = np.r_[rng.choice(p.size, 300, replace=False), np.arange(p.size - 1 - H, p.size)]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (orc.mlp_loss_grad(pp, X, y, 3.0)[0] - orc.mlp_loss_grad(pm, X, y, 3.0)[0]) / (2 * eps)
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # loss itself: weighted binary cross-entropy with logits
    z = orc.mlp_forward(p, X)
    w = np.where(y > 0.5, 3.0, 1.0)
    want = np.mean(w * np.where(y > 0.5, np.logaddexp(0, -z), np.logaddexp(0, z)))
    np.testing.assert_allclose(loss, want, rtol=1e-12)


def _summaries(rng, K, D):
    N = rng.integers(50, 5000, K).astype(float)
    Nl = np.floor(N * rng.uniform(0.2, 0.8, K))
    return (N, Nl, N - Nl, -rng.uniform(1e3, 1e5, K), -rng.uniform(1e3, 1e5, K), -rng.uniform(1e3, 1e5, K),
            rng.normal(size=K) * 3, rng.normal(size=(K, D)) * 10, rng.uniform(0.5, 4.0, (K, D)))


def test_pair_features_invariants():
    rng = np.random.default_rng(2)
    K, D = 6, 5
    S = list(_summaries(rng, K, D))
    for a in range(len(S)):   # cluster 5 is a copy of cluster 4
        S[a] = S[a].copy()
        S[a][5] = S[a][4]
    f = orc.pair_features(*S, [4, 0], [5, 3])
    np.testing.assert_allclose(f[0, [7, 8, 9, 14, 15]], 0.0, atol=1e-15)
    np.testing.assert_allclose(f[0, [0, 3, 5, 10, 12]], f[0, [1, 4, 6, 11, 13]])
    g = orc.pair_features(*S, [3], [0])   # swapped pair: per-cluster features swap, pair ones stay
    np.testing.assert_allclose(g[0, [0, 1, 3, 4, 5, 6, 7, 8, 10, 11, 12, 13]],
                               f[1, [1, 0, 4, 3, 6, 5, 8, 7, 11, 10, 13, 12]], rtol=1e-14)
    np.testing.assert_allclose(g[0, [2, 9, 14, 15]], f[1, [2, 9, 14, 15]], rtol=1e-14)
    N, Nl, Nr, L, Ll, Lr, ld, m, dpsi = S
    d = m[0] - m[3]
    np.testing.assert_allclose(f[1, 9], np.log1p(np.sum(d * d / (dpsi[0] + dpsi[3])) / D), rtol=1e-14)
    np.testing.assert_allclose(f[1, 10], (Ll[0] + Lr[0] - L[0]) / (N[0] * D), rtol=1e-14)
    np.testing.assert_allclose(f[1, 3], L[0] / (N[0] * D), rtol=1e-14)   # per point and dimension
