"""ctypes binding of the FP64 CPU oracle (oracle/build/liboracle_dpmm.so).

Used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle_dpmm.so")

_lib = None

c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_u32 = ctypes.c_uint32
c_dbl = ctypes.c_double
c_int = ctypes.c_int
P = ctypes.c_void_p


class RunOpts(ctypes.Structure):
    _fields_ = [
        ("prior", c_int), ("alpha", c_dbl), ("iters", c_int), ("init_clusters", c_int),
        ("burnout", c_int), ("argmax_sample_stop", c_int), ("split_stop", c_int),
        ("max_clusters", c_dbl), ("seed", c_u64), ("full_mu_cov", c_int),
        ("mult_use_weight", c_int), ("verbose", c_int), ("outlier_mod", c_dbl),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_lgamma.restype = c_dbl
        L.orc_log_multivariate_gamma.restype = c_dbl
        L.orc_logdet_spd.restype = c_dbl
        L.orc_niw_log_marginal.restype = c_dbl
        L.orc_mult_log_marginal.restype = c_dbl
        L.orc_split_log_hr.restype = c_dbl
        L.orc_split_log_hr.argtypes = [c_dbl, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl]
        L.orc_merge_log_hr.restype = c_dbl
        L.orc_merge_log_hr.argtypes = [c_dbl, c_i64, c_i64, c_dbl, c_dbl, c_dbl]
        L.orc_nmi.restype = c_dbl
        L.orc_run.restype = c_int
        L.orc_mlp_loss_grad.restype = c_dbl
        L.orc_model_create.restype = P
        L.orc_model_create_outlier.restype = P
        L.orc_model_destroy.argtypes = [P]
        L.orc_model_trace.argtypes = [P, c_int]
        L.orc_model_step.argtypes = [P, c_int]
        L.orc_model_num_clusters.argtypes = [P]
        L.orc_model_state.argtypes = [P, P, P, P, P]
        L.orc_model_sweep_trace.argtypes = [P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(P) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def philox(ctr, key):
    c = (c_u32 * 4)(*ctr)
    k = (c_u32 * 2)(*key)
    o = (c_u32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def stream(kind, seed, c0, c2, c3, n, shape=None, alpha=None):
    out = np.empty(n, np.float64)
    L = lib()
    if kind == "uniform":
        L.orc_stream_uniforms(c_u64(seed), c_u32(c0), c_u32(c2), c_u32(c3), c_int(n), _p(out))
    elif kind == "normal":
        L.orc_stream_normals(c_u64(seed), c_u32(c0), c_u32(c2), c_u32(c3), c_int(n), _p(out))
    elif kind == "gamma":
        L.orc_stream_gammas(c_u64(seed), c_u32(c0), c_u32(c2), c_u32(c3), c_dbl(shape), c_int(n), _p(out))
    elif kind == "dirichlet":
        a = f64(alpha)
        L.orc_stream_dirichlet(c_u64(seed), c_u32(c0), c_u32(c2), c_u32(c3), c_int(n), _p(a), _p(out))
    elif kind == "dirichlet_substreams":
        a = f64(alpha)
        L.orc_dirichlet_substreams(c_u64(seed), c_u32(c0), c_u32(c2), c_u32(c3), c_int(n), _p(a), _p(out))
    return out


def gauss_loglik(X, mu, inv_sigma, logdet, weight=1.0, use_weight=True):
    X = f64(X)
    out = np.empty(X.shape[0])
    lib().orc_gauss_loglik(_p(X), c_i64(X.shape[0]), c_int(X.shape[1]), _p(f64(mu)), _p(f64(inv_sigma)),
                           c_dbl(logdet), c_dbl(weight), c_int(int(use_weight)), _p(out))
    return out


def mult_loglik(X, logtheta):
    X = f64(X)
    out = np.empty(X.shape[0])
    lib().orc_mult_loglik(_p(X), c_i64(X.shape[0]), c_int(X.shape[1]), _p(f64(logtheta)), _p(out))
    return out


def niw_posterior(k0, m0, v0, psi0, N, s, S):
    D = len(m0)
    k, v = c_dbl(), c_dbl()
    m = np.empty(D)
    psi = np.empty((D, D))
    lib().orc_niw_posterior(c_int(D), c_dbl(k0), _p(f64(m0)), c_dbl(v0), _p(f64(psi0)), c_i64(N), _p(f64(s)),
                            _p(f64(S)), ctypes.byref(k), _p(m), ctypes.byref(v), _p(psi))
    return k.value, m, v.value, psi


def niw_log_marginal(k0, v0, psi0, kn, vn, psin, N):
    psi0 = f64(psi0)
    D = psi0.shape[0]
    return lib().orc_niw_log_marginal(c_int(D), c_dbl(k0), c_dbl(v0), _p(psi0), c_dbl(kn), c_dbl(vn),
                                      _p(f64(psin)), c_i64(N))


def mult_log_marginal(a0, an):
    a0 = f64(a0)
    return lib().orc_mult_log_marginal(c_int(a0.size), _p(a0), _p(f64(an)))


def split_log_hr(alpha, N, Nl, Nr, L, Ll, Lr):
    return lib().orc_split_log_hr(alpha, N, Nl, Nr, L, Ll, Lr)


def merge_log_hr(alpha, Ni, Nj, L, Li, Lj):
    return lib().orc_merge_log_hr(alpha, Ni, Nj, L, Li, Lj)


def niw_draw(k, m, v, psi, seed, slot, it, full_mu_cov=False):
    D = len(m)
    mu = np.empty(D)
    lam = np.empty((D, D))
    ld = c_dbl()
    lib().orc_niw_draw(c_int(D), c_dbl(k), _p(f64(m)), c_dbl(v), _p(f64(psi)), c_u64(seed), c_u32(slot),
                       c_u32(it), c_int(int(full_mu_cov)), _p(mu), _p(lam), ctypes.byref(ld))
    return mu, lam, ld.value


def label_sweep(prior, X, mu=None, inv_sigma=None, logdet=None, logtheta=None, weights=None, seed=0, it=1,
                final=False, mult_use_weight=False, want_scores=False):
    X = f64(X)
    N, D = X.shape
    K = len(weights)
    labels = np.empty(N, np.int32)
    margin = np.empty(N)
    scores = np.empty((N, K)) if want_scores else None
    lib().orc_label_sweep(c_int(prior), _p(X), c_i64(N), c_int(D), c_int(K),
                          _p(f64(mu) if mu is not None else None),
                          _p(f64(inv_sigma) if inv_sigma is not None else None),
                          _p(f64(logdet) if logdet is not None else None),
                          _p(f64(logtheta) if logtheta is not None else None),
                          _p(f64(weights)), c_int(int(mult_use_weight)), c_u64(seed), c_u32(it), c_int(int(final)),
                          _p(labels), _p(margin), _p(scores))
    return (labels, margin, scores) if want_scores else (labels, margin)


def sublabel_sweep(prior, X, labels, mu=None, inv_sigma=None, logdet=None, logtheta=None, lr_weights=None, seed=0,
                   it=1):
    X = f64(X)
    N, D = X.shape
    K = len(lr_weights) // 2
    subs = np.empty(N, np.int32)
    margin = np.empty(N)
    lib().orc_sublabel_sweep(c_int(prior), _p(X), c_i64(N), c_int(D), c_int(K), _p(i32(labels)),
                             _p(f64(mu) if mu is not None else None),
                             _p(f64(inv_sigma) if inv_sigma is not None else None),
                             _p(f64(logdet) if logdet is not None else None),
                             _p(f64(logtheta) if logtheta is not None else None),
                             _p(f64(lr_weights)), c_u64(seed), c_u32(it), _p(subs), _p(margin))
    return subs, margin


def suff_stats(prior, X, labels, subs, K, want_S=True):
    X = f64(X)
    N, D = X.shape
    Nsub = np.empty(2 * K, np.int64)
    s = np.empty((2 * K, D))
    S = np.empty((2 * K, D, D)) if (want_S and prior == 0) else None
    Nc = np.empty(K, np.int64)
    sc = np.empty((K, D))
    Sc = np.empty((K, D, D)) if (want_S and prior == 0) else None
    lib().orc_suff_stats(c_int(prior), _p(X), c_i64(N), c_int(D), c_int(K), _p(i32(labels)), _p(i32(subs)),
                         _p(Nsub), _p(s), _p(S), _p(Nc), _p(sc), _p(Sc))
    return Nsub, s, S, Nc, sc, Sc


def nmi(a, b):
    a = i32(a)
    return lib().orc_nmi(_p(a), _p(i32(b)), c_i64(a.size))


def run(X, prior=0, alpha=1.0, iters=100, init_clusters=1, burnout=15, seed=0, k0=1.0, m0=None, v0=None, psi0=None,
        alpha0=None, max_clusters=float("inf"), full_mu_cov=False, mult_use_weight=False, verbose=False):
    X = f64(X)
    N, D = X.shape
    o = RunOpts(prior=prior, alpha=alpha, iters=iters, init_clusters=init_clusters, burnout=burnout,
                argmax_sample_stop=5, split_stop=5, max_clusters=max_clusters, seed=seed,
                full_mu_cov=int(full_mu_cov), mult_use_weight=int(mult_use_weight), verbose=int(verbose))
    labels = np.empty(N, np.int32)
    khist = np.zeros(iters, np.int32)
    weights = np.zeros(4096)
    secs = np.zeros(iters)
    if prior == 0:
        m0 = np.zeros(D) if m0 is None else m0
        v0 = float(D + 3) if v0 is None else v0
        psi0 = np.eye(D) if psi0 is None else psi0
        K = lib().orc_run(ctypes.byref(o), _p(X), c_i64(N), c_int(D), c_dbl(k0), _p(f64(m0)), c_dbl(v0),
                          _p(f64(psi0)), None, _p(labels), _p(khist), _p(weights), c_int(4096), _p(secs))
    else:
        alpha0 = np.ones(D) if alpha0 is None else alpha0
        K = lib().orc_run(ctypes.byref(o), _p(X), c_i64(N), c_int(D), c_dbl(0), None, c_dbl(0), None,
                          _p(f64(alpha0)), _p(labels), _p(khist), _p(weights), c_int(4096), _p(secs))
    return {"K": K, "labels": labels, "k_hist": khist, "weights": weights[:max(K, 0)], "iter_seconds": secs}


class Model:
    """The oracle sampler one iteration at a time (orc_model_*): same options and defaults as run()."""

    def __init__(self, X, prior=0, alpha=1.0, iters=100, init_clusters=1, burnout=15, seed=0, k0=1.0, m0=None,
                 v0=None, psi0=None, alpha0=None, max_clusters=float("inf"), full_mu_cov=False,
                 mult_use_weight=False, trace=False, outlier_weight=0.0, outlier_params=None):
        """outlier_params: (k0, m0, v0, psi0) for the Gaussian prior, alpha0 for the multinomial one."""
        X = f64(X)
        self.N, D = X.shape
        o = RunOpts(prior=prior, alpha=alpha, iters=iters, init_clusters=init_clusters, burnout=burnout,
                    argmax_sample_stop=5, split_stop=5, max_clusters=max_clusters, seed=seed,
                    full_mu_cov=int(full_mu_cov), mult_use_weight=int(mult_use_weight), verbose=0,
                    outlier_mod=float(outlier_weight))
        L = lib()
        if outlier_weight > 0:
            m0 = np.zeros(D) if m0 is None else m0
            v0 = float(D + 3) if v0 is None else v0
            psi0 = np.eye(D) if psi0 is None else psi0
            alpha0 = np.ones(D) if alpha0 is None else alpha0
            if prior == 0:
                ok0, om0, ov0, opsi0 = outlier_params
                self._m = L.orc_model_create_outlier(ctypes.byref(o), _p(X), c_i64(self.N), c_int(D), c_dbl(k0),
                                                     _p(f64(m0)), c_dbl(v0), _p(f64(psi0)), None, c_dbl(ok0),
                                                     _p(f64(om0)), c_dbl(ov0), _p(f64(opsi0)), None)
            else:
                self._m = L.orc_model_create_outlier(ctypes.byref(o), _p(X), c_i64(self.N), c_int(D), c_dbl(0), None,
                                                     c_dbl(0), None, _p(f64(alpha0)), c_dbl(0), None, c_dbl(0), None,
                                                     _p(f64(outlier_params)))
        elif prior == 0:
            m0 = np.zeros(D) if m0 is None else m0
            v0 = float(D + 3) if v0 is None else v0
            psi0 = np.eye(D) if psi0 is None else psi0
            self._m = L.orc_model_create(ctypes.byref(o), _p(X), c_i64(self.N), c_int(D), c_dbl(k0), _p(f64(m0)),
                                         c_dbl(v0), _p(f64(psi0)), None)
        else:
            alpha0 = np.ones(D) if alpha0 is None else alpha0
            self._m = L.orc_model_create(ctypes.byref(o), _p(X), c_i64(self.N), c_int(D), c_dbl(0), None, c_dbl(0),
                                         None, _p(f64(alpha0)))
        L.orc_model_trace(self._m, c_int(int(trace)))

    def step(self, t):
        return lib().orc_model_step(self._m, c_int(t))

    def num_clusters(self):
        return lib().orc_model_num_clusters(self._m)

    def state(self):
        """(labels, sub-labels, per-cluster counts, mixture weights)."""
        K = self.num_clusters()
        lab, sub = np.empty(self.N, np.int32), np.empty(self.N, np.int32)
        cnt, w = np.empty(K, np.int64), np.empty(K)
        lib().orc_model_state(self._m, _p(lab), _p(sub), _p(cnt), _p(w))
        return lab, sub, cnt, w

    def sweep_trace(self):
        """(labels, sub-labels, label margin, sub-label margin) of the last step's sweeps."""
        lab, sub = np.empty(self.N, np.int32), np.empty(self.N, np.int32)
        m, sm = np.empty(self.N), np.empty(self.N)
        if lib().orc_model_sweep_trace(self._m, _p(lab), _p(sub), _p(m), _p(sm)) != 0:
            raise RuntimeError("trace is off")
        return lab, sub, m, sm

    def close(self):
        if self._m:
            lib().orc_model_destroy(self._m)
            self._m = None

    def __del__(self):
        self.close()


def mlp_param_count():
    return lib().orc_mlp_param_count()


def mlp_forward(params, X):
    X = f64(X)
    z = np.empty(X.shape[0])
    lib().orc_mlp_forward(_p(f64(params)), _p(X), c_int(X.shape[0]), _p(z))
    return z


def mlp_loss_grad(params, X, y, pos_weight=1.0):
    X, y = f64(X), f64(y)
    g = np.empty(mlp_param_count())
    loss = lib().orc_mlp_loss_grad(_p(f64(params)), _p(X), _p(y), c_int(X.shape[0]), c_dbl(pos_weight), _p(g))
    return loss, g


def pair_features(N, Nl, Nr, L, Ll, Lr, ld, m, dpsi, pi, pj):
    K, D = np.asarray(m).shape
    out = np.empty((len(pi), 16))
    lib().orc_pair_features(c_int(K), c_int(D), _p(f64(N)), _p(f64(Nl)), _p(f64(Nr)), _p(f64(L)), _p(f64(Ll)),
                            _p(f64(Lr)), _p(f64(ld)), _p(f64(m)), _p(f64(dpsi)), c_int(len(pi)), _p(i32(pi)),
                            _p(i32(pj)), _p(out))
    return out
