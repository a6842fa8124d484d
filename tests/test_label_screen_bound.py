"""The error bound of the mixed label screen (csrc/dpmm_mixed.h), checked on the CPU against an
emulation of its arithmetic.  The screen forms y = bf16(W) bf16(x - mu) (products of BF16 values are
exact in FP32; here they are summed in FP64, within the bound's accumulation term of the MFMA's FP32
sum) and takes q = |y|^2; the FP32 sweep's q is |W (x - mu)|^2 with W and mu in FP32 (here in FP64 from
those FP32 values).  With a_j = sum_l |bf16(W_jl)| |bf16(d_l)|, A = sum_j a_j^2 and the library's
factor eps = (2^-7 + 2^-16 + (2 DP + 16) 2^-24)(1 + 2^-6) (dpmm_screen_error_factors), half the
difference of the two q must stay below
eps (sqrt(q A) + eps A / 2), whatever the covariance.  The heuristic the screen used before, q / 32,
fails on strongly correlated, ill-conditioned covariances; the case is kept as the reason for the
derived bound."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp
from helpers import adversarial_mixture


def bf16(a):
    """Round-to-nearest-even to BF16 (as the packers and v_cvt do), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def screen_errors(X, mu, lam, D):
    eps = dp.screen_error_factors(D)[0]
    W32 = np.linalg.cholesky(lam).T.astype(np.float32)   # W = R^T, invSigma = R R^T
    Wb = bf16(W32).astype(np.float64)
    mu32 = mu.astype(np.float32)
    d32 = (X - mu32).astype(np.float32)
    db = bf16(d32).astype(np.float64)
    ys = db @ Wb.T
    yr = d32.astype(np.float64) @ W32.astype(np.float64).T
    a = np.abs(db) @ np.abs(Wb).T
    qs, qr, A = (ys ** 2).sum(1), (yr ** 2).sum(1), (a ** 2).sum(1)
    err = 0.5 * np.abs(qs - qr)
    bound = eps * (np.sqrt(qs * A) + 0.5 * eps * A)
    return err, bound, qs / 32


@pytest.mark.parametrize("D", [64, 128])
def test_bound_holds_where_the_heuristic_fails(D):
    X, _, mus, sig = adversarial_mixture(3000, D, 4, seed=1)
    fails = 0
    for k in range(len(mus)):
        err, bound, old = screen_errors(X, mus[k], np.linalg.inv(sig[k]), D)
        assert (err <= bound).all(), f"D={D} cluster {k}: error {err.max():.3g} above the bound"
        fails += int((err > old).sum())
        print(f"D={D} cluster {k}: max error / bound {(err / bound).max():.3f}, max error / (q/32) "
              f"{(err / old).max():.2f}")
    assert fails > 0.1 * X.shape[0] * len(mus)   # the old heuristic is wrong on many of these points


@pytest.mark.parametrize("D", [32, 64, 128])
def test_bound_holds_on_well_conditioned_covariances(D):
    rng = np.random.default_rng(D)
    A = rng.normal(size=(D, D))
    S = A @ A.T + 0.5 * np.eye(D)   # the covariances of the N = 1e7, D = 64 workload
    mu = rng.uniform(-100, 100, D)
    X = (mu + rng.normal(size=(3000, D)) @ np.linalg.cholesky(S).T).astype(np.float32)
    err, bound, old = screen_errors(X, mu, np.linalg.inv(S), D)
    assert (err <= bound).all()
    print(f"D={D}: median bound / (q/32) {np.median(bound / old):.2f}")
