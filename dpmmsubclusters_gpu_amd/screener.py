"""Learned merge-pair screener: training data, training and threshold choice.

The screener (an MLP 16 -> 128 -> 128 -> 1 whose forward, backward and Adam kernels are the
library's own HIP kernels) scores candidate merge pairs from 16 features of their two clusters.
Its training data are the exact merge decisions of the plain sampler: every candidate pair of every
iteration, its features, its log Hastings ratio log H and its uniform's log U (shared_actions.cpp:
97-111).  Accepted merges are rare (most pairs have log H in the hundreds below zero), so a pair's
label is whether the exact test could plausibly accept it, log H > -30 (the band of pairs near
acceptance); the accepted decisions themselves (log H > log U) measure the screener's recall.  The
sampler then gives the exact test only to pairs scored above a threshold, and still runs the full
all-pairs test every few iterations as fallback, counting the accepted merges it would have missed.
A screener that misses no accepted merge leaves the sampler's trajectory unchanged: the skipped
pairs would all have been rejected.
"""
import ctypes

import numpy as np

from . import DeviceSweep, Sampler, _check, _ptr, load_library, niw_hyperparams, prior_type
from .synthetic import gaussian_mixture

N_FEATURES = 16


class Screener:
    """The device MLP (dpmm_screener_* in include/dpmm_hip.h)."""

    def __init__(self, seed=0):
        L = load_library()
        self.handle = ctypes.c_void_p()
        _check(L.dpmm_screener_create(int(seed), ctypes.byref(self.handle)))
        self.n_params = L.dpmm_screener_param_count()

    def close(self):
        if self.handle:
            load_library().dpmm_screener_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, mean=None, scale=None):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters")
        m = None if mean is None else np.ascontiguousarray(mean, np.float32)
        s = None if scale is None else np.ascontiguousarray(scale, np.float32)
        _check(load_library().dpmm_screener_set_params(self.handle, _ptr(p), _ptr(m), _ptr(s)))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        m, s = np.empty(N_FEATURES, np.float32), np.empty(N_FEATURES, np.float32)
        _check(load_library().dpmm_screener_get_params(self.handle, _ptr(p), _ptr(m), _ptr(s)))
        return p, m, s

    def loss_grad(self, X, y, pos_weight=1.0):
        X, y = np.ascontiguousarray(X, np.float32), np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_float()
        g = np.empty(self.n_params, np.float32)
        _check(load_library().dpmm_screener_loss_grad(self.handle, X.shape[0], _ptr(X), _ptr(y), float(pos_weight),
                                                       ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def train_step(self, X, y, pos_weight=1.0, lr=1e-3):
        X, y = np.ascontiguousarray(X, np.float32), np.ascontiguousarray(y, np.float32)
        loss = ctypes.c_float()
        _check(load_library().dpmm_screener_train_step(self.handle, X.shape[0], _ptr(X), _ptr(y), float(pos_weight),
                                                        float(lr), ctypes.byref(loss)))
        return loss.value

    def logits(self, X):
        X = np.ascontiguousarray(X, np.float32)
        z = np.empty(X.shape[0], np.float32)
        _check(load_library().dpmm_screener_forward(self.handle, X.shape[0], _ptr(X), _ptr(z)))
        return z

    def save(self, path, threshold):
        p, m, s = self.get_params()
        np.savez(path, params=p, mean=m, scale=s, threshold=np.float32(threshold))

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        sc = cls()
        sc.set_params(d["params"], d["mean"], d["scale"])
        return sc, float(d["threshold"])


def over_split_labels(gt, pieces, seed):
    """Each true cluster cut into `pieces` random fragments: the over-split start of a merge-heavy run."""
    rng = np.random.default_rng(seed)
    return (np.asarray(gt) * pieces + rng.integers(0, pieces, len(gt))).astype(np.int32)


def collect_pairs(n_points, D, K_true, pieces, iters, seed, alpha=10.0, spread=100.0, precision=None,
                  dtype=np.float32, a=1.0, ridge=1.0, weight=None):
    """Run the plain sampler (every pair gets the exact test) on a synthetic Gaussian mixture from
    an over-split start and log every candidate pair of every iteration: features (rows x 16),
    log H, log U."""
    X, gt, _ = gaussian_mixture(n_points, D, K_true, seed=seed, spread=spread, dtype=dtype, a=a, ridge=ridge,
                                weight=weight)
    dev = DeviceSweep(X, prior_type.Gaussian, seed=seed)
    hyper = niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    s = Sampler(dev, hyper, alpha, iters, seed=seed, precision=precision,
                init_labels=over_split_labels(gt, pieces, seed))
    s.set_screener(None, log_pairs=True)
    feats, lh, lu = [], [], []
    for t in range(1, iters + 1):
        s.step(t)
        f, h, u = s.pair_log()
        if len(h):
            feats.append(f)
            lh.append(h)
            lu.append(u)
    s.close()
    dev.close()
    if not feats:
        return np.empty((0, N_FEATURES), np.float32), np.empty(0), np.empty(0)
    return np.concatenate(feats), np.concatenate(lh), np.concatenate(lu)


# Mixtures of the two small benchmark configurations' families (means ~ U(-100, 100)^D, covariances
# A A^T + I with A_ij ~ N(0, 1)): the README example's (D = 2, 2 clusters, equal weights, float64) and
# the D = 32 one's (20 clusters, weights ~ Dirichlet(1), float32), each started over-split.
CONFIG_STYLES = {
    "c1": dict(D=2, K_true=2, weight=None, dtype=np.float64, alpha=1.0),
    "c2": dict(D=32, K_true=20, weight=1.0, dtype=np.float32, alpha=10.0),
}


def collect_config_style_pairs(runs, n_points=60_000, pieces=3, iters=60, precision=None, log=None):
    """Training data from the plain sampler on configs-1-2-style mixtures: `runs` is a list of
    (style, seed) with style "c1" or "c2"; returns features, log H, log U of every logged pair.
    `precision` applies to the float32 mixtures; the float64 ones sweep in FP64."""
    feats, hs, us = [], [], []
    for style, seed in runs:
        c = CONFIG_STYLES[style]
        if log:
            log(f"screener training data: {style}-style mixture, seed {seed}")
        f, h, u = collect_pairs(n_points, c["D"], c["K_true"], pieces, iters, seed, alpha=c["alpha"],
                                precision=precision if c["dtype"] == np.float32 else None, dtype=c["dtype"],
                                weight=c["weight"])
        feats.append(f)
        hs.append(h)
        us.append(u)
    return np.concatenate(feats), np.concatenate(hs), np.concatenate(us)


def train(X, y, epochs=30, batch=4096, lr=2e-3, pos_weight=None, seed=0):
    """Train a screener on the device (Adam, FP32 kernels); features are standardised with the
    training set's mean and scale.  Returns the screener and the per-epoch mean loss."""
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    mean = X.mean(0)
    scale = X.std(0)
    scale[scale < 1e-6] = 1.0
    if pos_weight is None:
        npos = max(1.0, float(y.sum()))
        pos_weight = float(np.clip((len(y) - npos) / npos, 1.0, 100.0))
    sc = Screener(seed)
    p, _, _ = sc.get_params()
    sc.set_params(p, mean, scale)
    rng = np.random.default_rng(seed)
    hist = []
    for _ in range(epochs):
        order = rng.permutation(len(y))
        losses = []
        for b0 in range(0, len(y), batch):
            idx = order[b0:b0 + batch]
            losses.append(sc.train_step(X[idx], y[idx], pos_weight, lr))
        hist.append(float(np.mean(losses)))
    return sc, hist


PLAUSIBLE_LOG_H = -30.0


def plausible(log_h):
    """Training label of logged pairs: log H > -30."""
    return (np.asarray(log_h) > PLAUSIBLE_LOG_H).astype(np.float32)


def threshold_for_recall(logits, y, recall=0.995, margin=0.0):
    """Largest logit threshold that keeps at least `recall` of the positive rows above it, lowered
    by `margin` logits for pairs unlike the training set."""
    pos = np.sort(np.asarray(logits)[np.asarray(y) > 0.5])
    if pos.size == 0:
        return float(np.min(logits)) - 1.0
    k = int(np.floor((1.0 - recall) * pos.size))
    return float(pos[k]) - 1e-3 - margin
