"""The error factors of the mixed (BF16 MFMA) screens against what BF16 rounding attains.

A screen keeps a label or sub-label only when its score, widened by its error bound, decides it; the
bound is eps a_j per row of y = W d with a_j = sum_l |W_jl| |d_l|.  BF16 keeps 8 significant bits, so
rounding to nearest is relative 2^-8, and the split v = bf16(v) + bf16(v - bf16(v)) leaves at most
2^-17 of v.  The checks:
  - those rounding facts on sampled FP32 values, edge cases near rounding midpoints included;
  - inputs built to attain the neglected terms: one product bf16(W) bf16(d) is off by ~2^-7 a_j, the
    three split products W_hi d_hi + W_hi d_lo + W_lo d_hi by ~2^-15 a_j; the library's factors
    (dpmm_screen_error_factors) must cover them with the FP32 sums' share on top;
  - an emulation of the screens' arithmetic (BF16 splits, FP32 sums) against the FP32 sweep's FMA
    chain on random and adversarial whitening rows: every row within eps a_j.
No GPU is used."""
import numpy as np
import pytest

import dpmmsubclusters_gpu_amd as dp


def bf16(x):
    """FP32 -> BF16 (round to nearest even), returned as FP32."""
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def split(x):
    x = np.asarray(x, np.float32)
    hi = bf16(x)
    return hi, bf16(x - hi)   # x - hi is exact in FP32


def test_bf16_rounding_facts():
    rng = np.random.default_rng(0)
    v = (rng.uniform(1, 2, 1_000_000) * np.exp2(rng.integers(-30, 30, 1_000_000))).astype(np.float32)
    # values just below and above the BF16 midpoints of [1, 2) and of the second split
    k = np.arange(128, dtype=np.float64)
    edge = np.concatenate([1 + (k + 0.5) * 2.0 ** -7 + s * 2.0 ** -23 for s in (-1, 1)]
                          + [1 + 2.0 ** -9 + (k + 0.5) * 2.0 ** -16 + s * 2.0 ** -23 for s in (-1, 1)])
    v = np.concatenate([v, edge.astype(np.float32), -v[:1000]])
    hi, lo = split(v)
    av = np.abs(v.astype(np.float64))
    assert (np.abs(v - hi.astype(np.float64)) / av).max() <= 2.0 ** -8
    assert (np.abs(lo.astype(np.float64)) / av).max() <= 2.0 ** -8 * (1 + 2.0 ** -8)
    res = v.astype(np.float64) - hi - lo
    assert (np.abs(res) / av).max() <= 2.0 ** -17
    # the bounds are attained (to within 2^-6): a tighter unit roundoff would be wrong
    assert (np.abs(v - hi.astype(np.float64)) / av).max() > 2.0 ** -8 * (1 - 2.0 ** -6)
    assert (np.abs(res) / av).max() > 2.0 ** -17 * (1 - 2.0 ** -6)


def _neglected(W, d, products):
    """|W d - screen products| / sum |W| |d| per row, in exact (FP64) arithmetic."""
    Wh, Wl = split(W)
    dh, dl = split(d)
    f = lambda a: a.astype(np.float64)
    exact = f(W) @ f(d)
    if products == 1:
        s = f(Wh) @ f(dh)
    elif products == 3:
        s = f(Wh) @ f(dh) + f(Wh) @ f(dl) + f(Wl) @ f(dh)
    else:
        s = f(Wh) @ f(dh) + f(Wh) @ f(dl) + f(Wl) @ f(dh) + f(Wl) @ f(dl)
    return np.abs(exact - s) / (np.abs(f(W)) @ np.abs(f(d)))


@pytest.mark.parametrize("DP", [32, 64, 96, 128])
def test_factors_cover_attained_rounding(DP):
    label, spl, rel_a, rel_g = dp.screen_error_factors(DP)
    D = DP
    # one product: v just below the midpoint 1 + 2^-8 rounds to 1, so W d - bf16(W) bf16(d) ~ 2^-7 W d
    v1 = np.float32(1 + 2.0 ** -8 - 2.0 ** -23)
    e1 = _neglected(np.full((1, D), v1), np.full(D, v1), 1).max()
    assert e1 > 0.99 * 2.0 ** -7
    # three split products: v = 1 + 2^-8 - 1.5 2^-16 - 2^-23 rounds to hi = 1, lo = 2^-8 - 2^-15 and leaves
    # a second residual of ~2^-17 v, all three neglected terms of one sign
    v3 = np.float32(1 + 2.0 ** -8 - 1.5 * 2.0 ** -16 - 2.0 ** -23)
    e3 = _neglected(np.full((1, D), v3), np.full(D, v3), 3).max()
    assert e3 > 0.97 * 2.0 ** -15
    # a fourth product, W_lo d_lo, would leave about half of it
    e4 = _neglected(np.full((1, D), v3), np.full(D, v3), 4).max()
    assert e4 < 0.6 * e3
    # the FP32 sums: the screen's (DP or 3 DP terms) and the sweep's (DP terms), first order
    u = 2.0 ** -24
    assert label >= e1 + 2 * DP * u
    assert spl >= e3 + 4 * DP * u
    assert rel_a >= e3 + 4 * DP * u
    assert rel_g >= (3 * DP + 2) * u


def _fp32_sum(terms):
    """Sequential FP32 sum over the last axis."""
    acc = np.zeros(terms.shape[0], np.float32)
    for l in range(terms.shape[1]):
        acc = (acc + terms[:, l]).astype(np.float32)
    return acc


def _fp32_fma_chain(W, d):
    acc = np.zeros(W.shape[0], np.float32)
    for l in range(W.shape[1]):
        acc = (W[:, l].astype(np.float64) * np.float64(d[l]) + acc.astype(np.float64)).astype(np.float32)
    return acc


def _screens(W, d):
    """y of the single-product screen, y of the three-product split screen, and a_j."""
    Wh, Wl = split(W)
    dh, dl = split(d)
    y1 = _fp32_sum((Wh * dh[None, :]).astype(np.float32))
    terms = np.concatenate([Wh * dh[None, :], Wh * dl[None, :], Wl * dh[None, :]], axis=1)
    y3 = _fp32_sum(terms.astype(np.float32))
    a = _fp32_sum((np.abs(Wh) * np.abs(dh)[None, :]).astype(np.float32))
    return y1, y3, a


@pytest.mark.parametrize("DP", [32, 64, 128])
def test_emulated_screens_within_bound(DP):
    label, spl, _, _ = dp.screen_error_factors(DP)
    rng = np.random.default_rng(DP)
    D = DP
    cases = []
    for _ in range(20):   # whitening-like rows: upper triangular, wide dynamic range
        W = np.triu(rng.normal(size=(D, D)) * np.exp(rng.normal(size=(D, 1)) * 2)).astype(np.float32)
        d = (rng.normal(size=D) * np.exp(rng.normal() * 3)).astype(np.float32)
        cases.append((W, d))
    # adversarial: signs arranged so that the rounding residues add up while y cancels
    for v in (np.float32(1 + 2.0 ** -8 - 2.0 ** -23), np.float32(1 + 2.0 ** -8 - 1.5 * 2.0 ** -16 - 2.0 ** -23)):
        s = np.where(rng.random(D) < 0.5, 1.0, -1.0).astype(np.float32)
        W = np.outer(np.ones(D), s).astype(np.float32) * v
        W[:, ::2] *= np.float32(1.0 + 2.0 ** -10)
        cases.append((W, (s * v).astype(np.float32)))
    worst1 = worst3 = 0.0
    for W, d in cases:
        y1, y3, a = _screens(W, d)
        yf = _fp32_fma_chain(W, d)
        ok = a > 0
        r1 = np.abs(y1.astype(np.float64) - yf)[ok] / a[ok]
        r3 = np.abs(y3.astype(np.float64) - yf)[ok] / a[ok]
        worst1, worst3 = max(worst1, r1.max()), max(worst3, r3.max())
    print(f"DP={DP}: worst |y_screen - y_fp32| / a_j: one product {worst1:.3e} (factor {label:.3e}), "
          f"split {worst3:.3e} (factor {spl:.3e})")
    assert worst1 <= label
    assert worst3 <= spl
