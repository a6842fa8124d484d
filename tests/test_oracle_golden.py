"""Pins the FP64 oracle against the reference's own golden vectors (tests/golden/
reference_vectors.json, read from the reference's gtest files) and against closed forms,
invariants and brute force for what those leave open.  CPU only."""
import json
import math
import os

import numpy as np
import pytest
from scipy import special, stats

import oracle_ctypes as orc

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def pts(entry, key="points_DxN"):
    """Reference D x N row-major fill -> N x D."""
    return np.array(entry[key]).reshape(entry["D"], entry["N"]).T.copy()


# ----------------------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Known-answer vectors of Philox4x32-10 from the Random123 distribution (kat_vectors).
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_stream_uniform_and_normal_moments():
    u = orc.stream("uniform", 7, 3, 1, 6, 200000)
    assert 0 < u.min() and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.005 and abs(u.var() - 1 / 12) < 0.002
    z = orc.stream("normal", 7, 3, 1, 6, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.02
    assert stats.kstest(z, "norm").pvalue > 1e-4


@pytest.mark.parametrize("a", [0.3, 1.0, 2.5, 31.0, 1000.0])
def test_gamma_sampler_distribution(a):
    g = orc.stream("gamma", 11, 5, 2, 6, 60000, shape=a)
    assert stats.kstest(g, "gamma", args=(a,)).pvalue > 1e-4


def test_dirichlet_mean():
    alpha = np.array([2.0, 5.0, 0.5])
    draws = np.array([orc.stream("dirichlet", 3, i, 1, 7, 3, alpha=alpha) for i in range(20000)])
    np.testing.assert_allclose(draws.sum(1), 1.0, rtol=1e-12)
    np.testing.assert_allclose(draws.mean(0), alpha / alpha.sum(), atol=0.01)


# ----------------------------------------------------------------------------- likelihoods
@pytest.mark.parametrize("key", ["gauss_ll_w05", "gauss_ll_w1"])
def test_gaussian_loglik_golden(key):
    e = G[key]
    X = np.array(e["x_DxN"]).reshape(2, 10).T
    r = orc.gauss_loglik(X, e["mu"], np.array(e["inv_sigma"]).reshape(2, 2), e["logdet_sigma"], e["weight"])
    np.testing.assert_allclose(r, e["expected"], atol=e["tol"], rtol=0)


def test_gaussian_loglik_identity_closed_form():
    # test/cudaKernel_gaussian_test.cpp:831-872: Sigma = I, constant uses sigma.size() = D^2.
    e = G["gauss_ll_w1"]
    X = np.array(e["x_DxN"]).reshape(2, 10).T
    mu = np.array(e["mu"])
    r = orc.gauss_loglik(X, mu, np.eye(2), e["logdet_sigma"], 1.0)
    expect = -(4 * math.log(2 * math.pi) + e["logdet_sigma"]) / 2 - 0.5 * ((X - mu) ** 2).sum(1)
    np.testing.assert_allclose(r, expect, rtol=1e-14)


def test_multinomial_loglik_golden():
    e = G["mult_ll"]
    X = np.array(e["x_DxN"]).reshape(2, 10).T
    r = orc.mult_loglik(X, e["logtheta"])
    np.testing.assert_allclose(r, e["expected"], atol=e["tol"], rtol=0)


# ----------------------------------------------------------------------------- statistics
def test_niw_stats_golden():
    e = G["niw_stats"]
    X = pts(e)
    Nsub, s, S, Nc, sc, Sc = orc.suff_stats(0, X, np.zeros(10), np.zeros(10), 1)
    assert Nc[0] == 10
    np.testing.assert_allclose(sc[0], e["expected_sum"], atol=e["tol"])
    np.testing.assert_allclose(Sc[0].ravel(), e["expected_S"], atol=e["tol"])


def test_lca_stats_gather_golden():
    e = G["lca_stats_gather"]
    X = pts(e)
    labels = np.array(e["labels"]) - 1
    subs = np.array(e["sub_labels"]) - 1
    Nsub, s, S, Nc, sc, Sc = orc.suff_stats(0, X, labels, subs, 1)
    assert Nsub[0] == e["expected_Nl"] and Nsub[1] == e["expected_Nr"]
    got = [sc[0][0], sc[0][1], s[0][0], s[0][1], s[1][0], s[1][1]]
    np.testing.assert_allclose(got, [x[0] for x in e["expected"]], atol=e["tol"])


def test_stats_brute_force_and_invariants():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(500, 5)) * 10
    labels = rng.integers(0, 4, 500)
    subs = rng.integers(0, 2, 500)
    Nsub, s, S, Nc, sc, Sc = orc.suff_stats(0, X, labels, subs, 4)
    for k in range(4):
        for j in range(2):
            m = (labels == k) & (subs == j)
            assert Nsub[2 * k + j] == m.sum()
            np.testing.assert_allclose(s[2 * k + j], X[m].sum(0), rtol=1e-12)
            np.testing.assert_allclose(S[2 * k + j], X[m].T @ X[m], rtol=1e-12)
        assert Nc[k] == Nsub[2 * k] + Nsub[2 * k + 1]
        np.testing.assert_allclose(Sc[k], S[2 * k] + S[2 * k + 1], rtol=1e-12)
        np.testing.assert_array_equal(Sc[k], Sc[k].T)
    assert Nc.sum() == 500


def test_multinomial_stats_golden():
    e = G["mult_stats"]
    X = pts(e)
    Nsub, s, S, Nc, sc, Sc = orc.suff_stats(1, X, np.zeros(10), np.zeros(10), 1)
    assert Nc[0] == 10 and S is None
    np.testing.assert_allclose(sc[0], e["expected_sum"], atol=e["tol"])


# ----------------------------------------------------------------------------- NIW math
def test_niw_posterior_golden():
    e = G["niw_posterior"]
    k, m, v, psi = orc.niw_posterior(e["k0"], e["m0"], e["v0"], np.array(e["psi0"]).reshape(2, 2), e["N"],
                                     e["sum"], np.array(e["S"]).reshape(2, 2))
    np.testing.assert_allclose(psi.ravel(), e["expected_psi"], atol=e["tol"])
    assert k == e["k0"] + e["N"] and v == e["v0"] + e["N"]


def test_lca_update_stats_golden():
    e = G["lca_update_stats"]
    X = pts(e)
    labels = np.array(e["labels"]) - 1
    subs = np.array(e["sub_labels"]) - 1
    Nsub, s, S, Nc, sc, Sc = orc.suff_stats(0, X, labels, subs, 1)
    psi0 = np.array(e["psi0"]).reshape(2, 2)
    got = []
    for N, su, SS in [(Nc[0], sc[0], Sc[0]), (Nsub[0], s[0], S[0]), (Nsub[1], s[1], S[1])]:
        k, m, v, psi = orc.niw_posterior(e["k0"], e["m0"], e["v0"], psi0, N, su, SS)
        got += [N, su[0], su[1], SS[0, 0], SS[0, 1], SS[1, 0], SS[1, 1], k, m[0], m[1], v,
                psi[0, 0], psi[0, 1], psi[1, 0], psi[1, 1]]
    np.testing.assert_allclose(got, [x[0] for x in e["expected"]], atol=e["tol"])


def test_niw_posterior_sequential_equals_batch():
    rng = np.random.default_rng(1)
    D = 4
    X = rng.normal(size=(40, D)) * 3 + 5
    m0 = rng.normal(size=D)
    A = rng.normal(size=(D, D))
    psi0 = A @ A.T + np.eye(D)
    k, m, v, psi = orc.niw_posterior(2.0, m0, D + 3.0, psi0, 40, X.sum(0), X.T @ X)
    k1, m1, v1, psi1 = orc.niw_posterior(2.0, m0, D + 3.0, psi0, 15, X[:15].sum(0), X[:15].T @ X[:15])
    k2, m2, v2, psi2 = orc.niw_posterior(k1, m1, v1, psi1, 25, X[15:].sum(0), X[15:].T @ X[15:])
    np.testing.assert_allclose([k, v], [k2, v2])
    np.testing.assert_allclose(m, m2, rtol=1e-12)
    np.testing.assert_allclose(psi, psi2, rtol=1e-10)


@pytest.mark.parametrize("key", ["niw_logml_d3", "niw_logml_d50"])
def test_niw_log_marginal_golden(key):
    e = G[key]
    D = e["D"]
    r = orc.niw_log_marginal(e["k0"], e["v0"], np.eye(D), e["kn"], e["vn"], np.array(e["psin"]).reshape(D, D), e["N"])
    assert abs(r - e["expected"]) < e["tol"]


def _mvt_logpdf(x, mu, scale, df):
    D = len(x)
    L = np.linalg.cholesky(scale)
    z = np.linalg.solve(L, x - mu)
    return (special.gammaln((df + D) / 2) - special.gammaln(df / 2) - D / 2 * math.log(df * math.pi)
            - np.log(np.diag(L)).sum() - (df + D) / 2 * math.log1p(z @ z / df))


@pytest.mark.parametrize("D", [1, 3])
def test_niw_log_marginal_chain_rule(D):
    """log p(x_1..x_N) = sum of Student-t posterior predictives (rank-one NIW updates),
    an independent route to the closed form of niw.cpp:122-138 (psi stored as Psi / v)."""
    rng = np.random.default_rng(D)
    N = 30
    X = rng.normal(size=(N, D)) * 2 + 1
    k0, v0 = 1.5, D + 2.0
    m0 = rng.normal(size=D)
    A = rng.normal(size=(D, D))
    psi0 = A @ A.T / 3 + np.eye(D)
    k, v, m, Psi = k0, v0, m0.copy(), v0 * psi0
    total = 0.0
    for x in X:
        df = v - D + 1
        total += _mvt_logpdf(x, m, Psi * (k + 1) / (k * df), df)
        Psi = Psi + k / (k + 1) * np.outer(x - m, x - m)
        m = (k * m + x) / (k + 1)
        k += 1
        v += 1
    kn, mn, vn, psin = orc.niw_posterior(k0, m0, v0, psi0, N, X.sum(0), X.T @ X)
    r = orc.niw_log_marginal(k0, v0, psi0, kn, vn, psin, N)
    assert abs(r - total) < 1e-8 * abs(total)
    np.testing.assert_allclose(psin * vn, Psi, rtol=1e-10)


def test_multinomial_log_marginal_golden_and_chain_rule():
    e = G["mult_logml"]
    assert abs(orc.mult_log_marginal(e["alpha0"], e["alphan"]) - e["expected"]) < e["tol"]
    # token-by-token Polya urn predictive
    rng = np.random.default_rng(3)
    a0 = np.array([0.5, 1.0, 2.0, 0.7])
    tokens = rng.integers(0, 4, 60)
    a = a0.copy()
    lp = 0.0
    for w in tokens:
        lp += math.log(a[w] / a.sum())
        a[w] += 1
    assert abs(orc.mult_log_marginal(a0, a) - lp) < 1e-10 * abs(lp)


def test_lgamma_and_multivariate_gamma():
    for x in [0.5, 1.0, 3.7, 150.25, 5e6]:
        assert abs(orc.lib().orc_lgamma(orc.c_dbl(x)) - special.gammaln(x)) < 1e-12 * max(1, abs(special.gammaln(x)))
    for D in [1, 2, 5]:
        x = 7.3
        assert abs(orc.lib().orc_log_multivariate_gamma(orc.c_dbl(x), D) - special.multigammaln(x, D)) < 1e-10


# ----------------------------------------------------------------------------- Hastings
def _logml_part(e, p):
    D = 2
    return orc.niw_log_marginal(e["k0"], e["v0"], np.array(e["psi0"]).reshape(D, D), p["kn"], p["vn"],
                                np.array(p["psin"]).reshape(D, D), p["N"])


def test_should_split_golden():
    e = G["lca_should_split"]
    c, l, r = e["parts"]
    # should_split_local recomputes posteriors from the prior and stats (local_clusters_actions.cpp:222-224)
    psi0 = np.array(e["psi0"]).reshape(2, 2)
    Ls = []
    for p in (c, l, r):
        k, m, v, psi = orc.niw_posterior(e["k0"], e["m0"], e["v0"], psi0, p["N"], p["sum"], np.array(p["S"]).reshape(2, 2))
        np.testing.assert_allclose(psi.ravel(), p["psin"], atol=2e-5)
        np.testing.assert_allclose(m, p["mn"], atol=2e-6)
        Ls.append(orc.niw_log_marginal(e["k0"], e["v0"], psi0, k, v, psi, p["N"]))
    lhr = orc.split_log_hr(e["alpha"], c["N"], l["N"], r["N"], Ls[0], Ls[1], Ls[2])
    assert abs(lhr - e["expected_log_hr"]) < 1e-4
    assert (lhr > math.log(e["uniform"])) == bool(e["expected_split"])


def test_burnout_history_golden():
    """sample_cluster_params pushes L_l + L_r into the burn-out ring (shared_actions.cpp:27-36)."""
    e = G["sa_hist"]
    vals = [_logml_part(e, s[0]) + _logml_part(e, s[1]) for s in e["states"]]
    got = [vals[0], vals[0], vals[1], vals[2]]
    np.testing.assert_allclose(got, e["expected_hist"], atol=2e-4)


def test_merge_plus_split_is_beta_ratio():
    """merge log H (shared_actions.cpp:97-99) + split log H (local_clusters_actions.cpp:230-233) of the
    same partition = log B(Ni + a/2, Nj + a/2) / B(a/2, a/2)."""
    for a, Ni, Nj, L, Li, Lj in [(10.0, 17, 40, -300.0, -120.0, -170.0), (1.0, 3, 2, -5.0, -2.5, -1.0),
                                 (0.3, 1000, 12345, -1e5, -2e4, -7e4)]:
        m = orc.merge_log_hr(a, Ni, Nj, L, Li, Lj)
        s = orc.split_log_hr(a, Ni + Nj, Ni, Nj, L, Li, Lj)
        beta = special.betaln(Ni + a / 2, Nj + a / 2) - special.betaln(a / 2, a / 2)
        assert abs((m + s) - beta) < 1e-9 * max(1, abs(beta))
        assert abs(orc.merge_log_hr(a, Nj, Ni, L, Lj, Li) - m) < 1e-9 * max(1, abs(m))


# ----------------------------------------------------------------------------- draws
def test_niw_draw_moments():
    """Lambda ~ Wishart((v psi)^-1, v) has mean psi^-1 (niw.cpp:82-107 with the Bartlett
    factor of rinvwish2); mu has mean m."""
    D = 3
    rng = np.random.default_rng(5)
    A = rng.normal(size=(D, D))
    psi = A @ A.T + np.eye(D)
    m = np.array([1.0, -2.0, 3.0])
    lams, mus = [], []
    for s in range(4000):
        mu, lam, ld = orc.niw_draw(4.0, m, 9.0, psi, seed=s, slot=0, it=1)
        lams.append(lam)
        mus.append(mu)
        if s < 5:
            assert abs(ld + np.linalg.slogdet(lam)[1]) < 1e-9
    np.testing.assert_allclose(np.mean(lams, 0), np.linalg.inv(psi), atol=0.05 * np.abs(np.linalg.inv(psi)).max())
    np.testing.assert_allclose(np.mean(mus, 0), m, atol=0.1)


def test_multinomial_draw_golden_distribution():
    e = G["mult_draw"]
    a = np.array(e["alpha"])
    d = orc.stream("dirichlet", 0, 0, 1, 6, 2, alpha=a)
    np.testing.assert_allclose(np.log(d), e["expected_logtheta"], atol=e["tol"])


# ----------------------------------------------------------------------------- categorical draws
def test_label_draw_distribution_matches_softmax():
    """Gumbel-max over the log-likelihoods draws label k with probability softmax(ll)_k, the
    distribution of the reference's softmax + alias draw (cudaKernel.cu:374-403, 27-124).
    Uses the reference's (0.4, 0.6) case of test/cudaKernel_gaussian_test.cpp:18-70."""
    e = G["cat_draw_counts"]
    n = e["n"]
    X = np.zeros((n, 1))
    # multinomial prior with x = 0 makes every per-cluster log-likelihood exactly log w
    labels, margin = orc.label_sweep(1, X, logtheta=np.zeros((2, 1)), weights=np.exp(e["log_likelihood"]),
                                     mult_use_weight=True, seed=123, it=1)
    p1 = 1 / (1 + math.exp(0.2))
    c1 = int((labels == 0).sum())
    sd = math.sqrt(n * p1 * (1 - p1))
    assert abs(c1 - n * p1) < 5 * sd
    assert abs(e["reference_counts"][0] - n * p1) < 5 * sd   # the reference's count is a draw of the same law
    # three clusters: chi-square against softmax
    w = np.array([0.2, 0.5, 0.3])
    labels, _ = orc.label_sweep(1, np.zeros((60000, 1)), logtheta=np.zeros((3, 1)), weights=w,
                                mult_use_weight=True, seed=9, it=4)
    cnt = np.bincount(labels, minlength=3)
    assert stats.chisquare(cnt, w * 60000).pvalue > 1e-4


def test_label_final_is_argmax_first_wins():
    X = np.zeros((10, 1))
    labels, _ = orc.label_sweep(1, X, logtheta=np.zeros((3, 1)), weights=np.array([0.2, 0.4, 0.4]),
                                mult_use_weight=True, final=True)
    assert (labels == 1).all()   # ties resolved to the first maximum (cudaKernel.cu:405-424)


def test_sublabel_draw_distribution():
    n = 80000
    X = np.zeros((n, 1))
    lrw = np.array([0.3, 0.7])
    subs, _ = orc.sublabel_sweep(1, X, np.zeros(n, np.int32), logtheta=np.zeros((2, 1)), lr_weights=lrw, seed=4)
    p0 = 0.3
    assert abs((subs == 0).mean() - p0) < 5 * math.sqrt(p0 * (1 - p0) / n)


# ----------------------------------------------------------------------------- end to end
def _gauss_mixture(seed, N, D, K, spread=100.0):
    rng = np.random.default_rng(seed)
    means = rng.uniform(-spread, spread, size=(K, D))
    gt = rng.integers(0, K, N)
    X = np.empty((N, D))
    for k in range(K):
        A = rng.normal(size=(D, D))
        C = A @ A.T + np.eye(D)
        idx = gt == k
        X[idx] = rng.multivariate_normal(means[k], C, size=idx.sum())
    return X, gt


def test_end_to_end_small_gaussian_finds_clusters():
    """The README basic example in miniature (module_tests.cpp:19-49): 2 well separated clusters."""
    X, gt = _gauss_mixture(0, 4000, 2, 2)
    out = orc.run(X, prior=0, alpha=1.0, iters=60, seed=1, k0=1.0, v0=5.0)
    assert out["K"] == 2
    assert orc.nmi(gt, out["labels"]) > 0.99


def test_end_to_end_multinomial():
    rng = np.random.default_rng(2)
    D, K, N = 50, 3, 3000
    theta = rng.dirichlet(np.full(D, 0.1), size=K)
    gt = rng.integers(0, K, N)
    X = np.stack([rng.multinomial(100, theta[g]) for g in gt]).astype(np.float64)
    out = orc.run(X, prior=1, alpha=10.0, iters=60, seed=3, alpha0=np.ones(D))
    assert out["K"] == 3
    assert orc.nmi(gt, out["labels"]) > 0.95


def test_dirichlet_substreams_moments():
    """Multinomial parameter draws use one sub-stream per coordinate; the draws are Dirichlet(alpha):
    means alpha / sum, variances a_d (A - a_d) / (A^2 (A + 1))."""
    alpha = np.array([0.5, 1.0, 2.0, 4.0, 8.5])
    A = alpha.sum()
    draws = np.stack([orc.stream("dirichlet_substreams", 3, c0, 1, 6, len(alpha), alpha=alpha) for c0 in range(4000)])
    np.testing.assert_allclose(draws.sum(1), 1.0, rtol=1e-12)
    mean, var = alpha / A, alpha * (A - alpha) / (A * A * (A + 1))
    assert np.all(np.abs(draws.mean(0) - mean) < 4 * np.sqrt(var / len(draws)))
    np.testing.assert_allclose(draws.var(0), var, rtol=0.15)
