"""The merge bound's mathematics (k_step_pair_bound), checked on the CPU against exact determinants.

For the NIW posterior Psi_n = v0 psi0 + S + P with S the scatter about the data mean and
P = (k0 N / kappa_n)(xbar - m0)(xbar - m0)^T (niw.cpp:23-49, unnormalised: v_n times the reference's
psi), merging clusters i and j gives Psi_m = W_i + W_j + c d d^T + P_m with W_k = v0 psi0 / 2 + S_k,
d = xbar_i - xbar_j, c = N_i N_j / (N_i + N_j).  The identity and the bound
  log det Psi_m >= D log(det W_i^(1/D) + det W_j^(1/D)) + log(1 + c |d|^2 / tr(W_i + W_j))
are checked with numpy's slogdet on random clusters, near and far apart, unequal in size, with a prior
mean away from zero; a mistake in c, in d, in the halving of the prior scatter or in the trace fails
them."""
import numpy as np
import pytest


def _posterior(X, k0, m0, Psi0):
    N = len(X)
    s = X.sum(0)
    kn = k0 + N
    mu = (k0 * m0 + s) / kn
    return Psi0 + k0 * np.outer(m0, m0) + X.T @ X - kn * np.outer(mu, mu)


def _scatter(X):
    return (X - X.mean(0)).T @ (X - X.mean(0))


@pytest.mark.parametrize("D,sep,seed", [(4, 0.5, 0), (8, 20.0, 1), (16, 3.0, 2), (32, 100.0, 3), (6, 0.0, 4)])
def test_merge_identity_and_bound(D, sep, seed):
    rng = np.random.default_rng(seed)
    k0, v0, m0 = 1.0 + rng.random(), D + 3.0, rng.normal(size=D) * 5.0
    A = rng.normal(size=(D, D))
    Psi0 = v0 * (A @ A.T / D + np.eye(D))
    Xi = rng.normal(size=(200 + seed * 50, D)) @ rng.normal(size=(D, D)) + rng.normal(size=D) * sep
    Xj = rng.normal(size=(120, D)) * 2.0 + rng.normal(size=D) * sep
    Ni, Nj = len(Xi), len(Xj)
    Pm = _posterior(np.vstack([Xi, Xj]), k0, m0, Psi0)
    Wi, Wj = Psi0 / 2 + _scatter(Xi), Psi0 / 2 + _scatter(Xj)
    d = Xi.mean(0) - Xj.mean(0)
    c = Ni * Nj / (Ni + Nj)
    xm = np.vstack([Xi, Xj]).mean(0)
    P = k0 * (Ni + Nj) / (k0 + Ni + Nj) * np.outer(xm - m0, xm - m0)
    np.testing.assert_allclose(Pm, Wi + Wj + c * np.outer(d, d) + P, rtol=1e-9, atol=1e-9 * np.abs(Pm).max())
    ld = lambda M: np.linalg.slogdet(M)[1]
    lwi, lwj = ld(Wi), ld(Wj)
    mink = D * np.logaddexp(lwi / D, lwj / D)
    assert ld(Wi + Wj) >= mink - 1e-9
    bound = mink + np.log1p(c * (d @ d) / np.trace(Wi + Wj))
    assert ld(Pm) >= bound - 1e-9
    # without the halving of the prior scatter the first term would not be a bound in general; with
    # the full prior scatter in each W it overstates det(W_i + W_j) when the data are few
    if seed == 4:
        W2i, W2j = Psi0 + _scatter(Xi[:3]), Psi0 + _scatter(Xj[:3])
        assert ld(Psi0 + _scatter(Xi[:3]) + _scatter(Xj[:3])) < D * np.logaddexp(ld(W2i) / D, ld(W2j) / D)


def test_stirling_upper_bound_of_lgamma():
    """k_step_pair_bound bounds log Gamma_D(v_m / 2) term by term with Robbins' remainder bound
    lgamma(x) < (x - 1/2) log x - x + log(2 pi) / 2 + 1 / (12 x), x > 0; it must exceed scipy's lgamma
    everywhere and by less than 1 / (12 x) - 1 / (360 x^3) + ... (tight for the large x of the bound)."""
    from scipy.special import gammaln
    x = np.concatenate([np.linspace(0.05, 5, 400), np.geomspace(5, 1e8, 400)])
    ub = (x - 0.5) * np.log(x) - x + 0.91893853320467274 + 1.0 / (12.0 * x)
    exact = gammaln(x)
    assert (ub > exact - 1e-12 * np.abs(exact)).all()
    big = x > 10
    assert (ub[big] - exact[big] < 1.0 / (300.0 * x[big] ** 3) + 1e-9 * np.abs(exact[big])).all()
