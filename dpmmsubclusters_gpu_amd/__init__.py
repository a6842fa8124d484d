"""MI355X-native sub-cluster sweep of the DPMM split/merge sampler (Chang & Fisher).

Python mirror of the reference's interface (BGU-CS-VIL/DPMMSubClusters_GPU):
``dp_parallel_sampling_class(num_labels, all_data, random_seed, prior_type)`` with
``dp_parallel(...)`` returning a ``ModelInfo`` (dp_parallel_sampling.h:11-43, modelInfo.h),
``niw_hyperparams`` / ``multinomial_hyper`` (priors/*_hyperparams.h) and ``prior_type``
(global_params.h:14-18), over the C ABI of lib/libdpmm_hip.so (include/dpmm_hip.h).

``DeviceSweep`` exposes the device layer itself (the reference's cudaKernel plugin,
cudaKernel.cuh:31-138) for callers that drive iterations themselves.

The HIP library is required: every entry point raises if lib/libdpmm_hip.so is missing or no
GPU is visible.  There is no CPU fallback.
"""
import ctypes
import enum
import os
import time

import numpy as np

__all__ = ["prior_type", "Precision", "niw_hyperparams", "multinomial_hyper", "ModelInfo",
           "dp_parallel_sampling_class", "DeviceSweep", "DPMMError", "library_path", "load_library"]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "lib", "libdpmm_hip.so")

DPMM_OK = 0
_ERRNAMES = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "ENOTSPD", -5: "ESTATE"}


class DPMMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dpmm error {code} ({_ERRNAMES.get(code, '?')}): {msg}")
        self.code = code


class prior_type(enum.IntEnum):
    """prior_type (global_params.h:14-18)."""
    Gaussian = 0
    Multinomial = 1


class Precision(enum.IntEnum):
    FP64 = 0
    FP32 = 1
    MIXED = 2


class _Options(ctypes.Structure):
    _fields_ = [("prior", ctypes.c_int), ("alpha", ctypes.c_double), ("iters", ctypes.c_int),
                ("init_clusters", ctypes.c_int), ("burnout", ctypes.c_int), ("argmax_sample_stop", ctypes.c_int),
                ("split_stop", ctypes.c_int), ("max_clusters", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("full_mu_cov", ctypes.c_int), ("mult_use_weight", ctypes.c_int), ("precision", ctypes.c_int),
                ("verbose", ctypes.c_int), ("init_from_labels", ctypes.c_int), ("hard_clustering", ctypes.c_int),
                ("outlier_mod", ctypes.c_double), ("host_decisions", ctypes.c_int)]


_lib = None
_P = ctypes.c_void_p


def load_library():
    """Load lib/libdpmm_hip.so; raises if it is missing (build with dpmmsubclusters_gpu_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(library_path):
        raise RuntimeError(f"HIP library not built: {library_path} (run `python -m dpmmsubclusters_gpu_amd.build`)")
    L = ctypes.CDLL(library_path)
    L.dpmm_last_error.restype = ctypes.c_char_p
    i64, u64, u32, dbl, i = ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double, ctypes.c_int
    sig = {
        "dpmm_device_count": [_P],
        "dpmm_debug_fail_allocations": [ctypes.c_int64],
        "dpmm_create": [_P, i64, i, i, i, u64, _P],
        "dpmm_create_device": [_P, i64, i, i, i, u64, _P],
        "dpmm_destroy": [_P],
        "dpmm_set_precision": [_P, i],
        "dpmm_max_clusters": [_P, _P],
        "dpmm_limit_clusters": [_P, i],
        "dpmm_sample_initial_labels": [_P, i],
        "dpmm_set_labels": [_P, _P, _P],
        "dpmm_get_labels": [_P, _P, _P],
        "dpmm_set_gaussian_params": [_P, i, _P, _P, _P, _P, _P, _P, _P, _P],
        "dpmm_set_gaussian_factors": [_P, i, _P, _P, _P, _P, _P, _P, _P, _P],
        "dpmm_set_multinomial_params": [_P, i, _P, _P, _P, _P, i],
        "dpmm_sample_cluster_labels": [_P, u32, i],
        "dpmm_sample_sub_labels": [_P, u32],
        "dpmm_sweep_iteration": [_P, u32, i, _P],
        "dpmm_set_graph_mode": [_P, i],
        "dpmm_set_ground_truth": [_P, _P, i],
        "dpmm_label_contingency": [_P, i, _P],
        "dpmm_label_nmi": [_P, i, _P],
        "dpmm_sampler_log_posterior": [_P, _P],
        "dpmm_graph_counts": [_P, _P, _P],
        "dpmm_log_likelihoods": [_P, i, i, ctypes.c_int64, ctypes.c_int64, _P],
        "dpmm_sufficient_statistics": [_P, i, _P, _P, _P],
        "dpmm_relabel": [_P, u32, i, _P, _P],
        "dpmm_last_kernel_ms": [_P, _P, _P, _P],
        "dpmm_last_sweep_kernel_ms": [_P, _P],
        "dpmm_set_mixed_tolerance": [_P, ctypes.c_float],
        "dpmm_screen_error_factors": [ctypes.c_int, ctypes.POINTER(ctypes.c_float)],
        "dpmm_last_recheck_count": [_P, _P],
        "dpmm_last_sub_recheck_count": [_P, _P],
        "dpmm_set_label_pruning": [_P, i],
        "dpmm_set_fused_sub_labels": [_P, i],
        "dpmm_set_sweep_workgroups": [_P, i],
        "dpmm_set_label_refinement": [_P, i],
        "dpmm_set_statistics_refresh": [_P, i],
        "dpmm_last_statistics_update": [_P, _P],
        "dpmm_last_label_evaluations": [_P, _P],
        "dpmm_synchronize": [_P],
        "dpmm_set_prior": [_P, dbl, _P, dbl, _P, _P],
        "dpmm_set_outlier_prior": [_P, dbl, _P, dbl, _P, _P],
        "dpmm_log_marginal_batch_ex": [_P, i, _P, _P, _P, _P],
        "dpmm_sample_params_ex": [_P, u32, i, _P, _P, _P, _P, i, i],
        "dpmm_set_statistics": [_P, i, _P, _P, _P],
        "dpmm_log_marginal_batch": [_P, i, _P, _P, _P],
        "dpmm_sample_params": [_P, u32, i, _P, _P, _P, i, i],
        "dpmm_get_drawn_params": [_P, _P, _P, _P],
        "dpmm_pair_features": [_P, i, _P, i, _P, _P, _P],
        "dpmm_screener_create": [u64, _P],
        "dpmm_screener_destroy": [_P],
        "dpmm_screener_param_count": [],
        "dpmm_screener_set_params": [_P, _P, _P, _P],
        "dpmm_screener_get_params": [_P, _P, _P, _P],
        "dpmm_screener_loss_grad": [_P, i, _P, _P, ctypes.c_float, _P, _P],
        "dpmm_screener_train_step": [_P, i, _P, _P, ctypes.c_float, ctypes.c_float, _P],
        "dpmm_screener_forward": [_P, i, _P, _P],
        "dpmm_sampler_set_screener": [_P, _P, i, ctypes.c_float, i],
        "dpmm_sampler_screen_stats": [_P, _P, _P, _P, _P],
        "dpmm_sampler_set_merge_bound": [_P, i],
        "dpmm_sampler_merge_bound_rejects": [_P, _P],
        "dpmm_sampler_events": [_P, _P, _P, _P, _P],
        "dpmm_sampler_pair_log": [_P, i, _P, _P, _P, _P],
        "dpmm_sampler_create": [_P, _P, dbl, _P, dbl, _P, _P, _P],
        "dpmm_sampler_step": [_P, i],
        "dpmm_sampler_num_clusters": [_P],
        "dpmm_sampler_weights": [_P, _P, _P, i],
        "dpmm_sampler_host_ms": [_P, _P],
        "dpmm_sampler_trace_sweep": [_P, _P, _P],
        "dpmm_sampler_destroy": [_P],
        "dpmm_run": [_P, _P, dbl, _P, dbl, _P, _P, _P, _P, _P, _P, i, _P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def _check(rc):
    if rc != DPMM_OK:
        raise DPMMError(rc, load_library().dpmm_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


class niw_hyperparams:
    """niw_hyperparams(k, m, v, psi) (priors/niw_hyperparams.h:13); psi is Psi / v as in the reference."""

    def __init__(self, k, m, v, psi):
        self.k = float(k)
        self.m = np.asarray(m, dtype=np.float64).ravel()
        self.v = float(v)
        self.psi = np.asarray(psi, dtype=np.float64).reshape(self.m.size, self.m.size)


class multinomial_hyper:
    """multinomial_hyper(alpha) (priors/multinomial_hyper.h:13)."""

    def __init__(self, alpha):
        self.alpha = np.asarray(alpha, dtype=np.float64).ravel()


class ModelInfo:
    """modelInfo.h:8-56: labels (1-based, as the reference), iter_count (seconds per iteration),
    cluster_count_history, likelihood_history, nmi_score_history, and the final mixture weights."""

    def __init__(self):
        self.labels = None
        self.iter_count = []
        self.cluster_count_history = []
        self.likelihood_history = []
        self.nmi_score_history = []
        self.weights = None
        self.cluster_counts = None


def screen_error_factors(DP):
    """Per-row error factors of the mixed screens at padded width DP: the single-product label screen's,
    the split screens', and the relative sub-cluster screen's eps_a and eps_g (needs no GPU)."""
    out = (ctypes.c_float * 4)()
    _check(load_library().dpmm_screen_error_factors(int(DP), out))
    return [float(v) for v in out]


def _nmi(a, b):
    """2 I(X;Y) / (H(X) + H(Y)) (dp_parallel_sampling.cpp:164-180)."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    n = a.size
    _, ai = np.unique(a, return_inverse=True)
    _, bi = np.unique(b, return_inverse=True)
    cont = np.zeros((ai.max() + 1, bi.max() + 1))
    np.add.at(cont, (ai, bi), 1)
    pa = cont.sum(1) / n
    pb = cont.sum(0) / n
    pab = cont / n
    nz = pab > 0
    mi = (pab[nz] * np.log(pab[nz] / np.outer(pa, pb)[nz])).sum()
    ha = -(pa[pa > 0] * np.log(pa[pa > 0])).sum()
    hb = -(pb[pb > 0] * np.log(pb[pb > 0])).sum()
    return 1.0 if ha + hb <= 0 else 2 * mi / (ha + hb)


class DeviceSweep:
    """The device layer (the reference's cudaKernel, cudaKernel.cuh:31-138) over HBM-resident points.

    X is N x D (each point contiguous): float32 / float64 for the Gaussian prior, int32 counts for
    the multinomial prior.  Labels are 0-based here; sub-labels are 0 / 1.
    """

    def __init__(self, X, prior=prior_type.Gaussian, seed=0, device_ptr=None, N=None, D=None, dtype=None):
        L = load_library()
        self._ctx = ctypes.c_void_p()
        if device_ptr is not None:
            self.N, self.D = int(N), int(D)
            self._dtype = {np.float64: 0, np.float32: 1, np.int32: 2}[np.dtype(dtype).type]
            _check(L.dpmm_create_device(ctypes.c_void_p(device_ptr), self.N, self.D, self._dtype, int(prior),
                                        int(seed), ctypes.byref(self._ctx)))
        else:
            X = np.ascontiguousarray(X)
            if X.dtype not in (np.float64, np.float32, np.int32):
                raise ValueError("X must be float64, float32 or int32")
            self.N, self.D = X.shape
            self._dtype = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int32): 2}[X.dtype]
            _check(L.dpmm_create(_ptr(X), self.N, self.D, self._dtype, int(prior), int(seed), ctypes.byref(self._ctx)))
        self.prior = prior_type(prior)

    def close(self):
        if self._ctx:
            load_library().dpmm_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._ctx

    def set_precision(self, p):
        _check(load_library().dpmm_set_precision(self._ctx, int(p)))

    def max_clusters(self):
        """Largest cluster count the kernels take (lowered by limit_clusters)."""
        v = ctypes.c_int()
        _check(load_library().dpmm_max_clusters(self._ctx, ctypes.byref(v)))
        return v.value

    def limit_clusters(self, kmax):
        """Lower the cluster cap at which the sampler stops splitting (0: the kernels' own)."""
        _check(load_library().dpmm_limit_clusters(self._ctx, int(kmax)))

    def sample_initial_labels(self, init_clusters):
        _check(load_library().dpmm_sample_initial_labels(self._ctx, int(init_clusters)))

    def set_labels(self, labels=None, sub_labels=None):
        lab, sub = _i32(labels), _i32(sub_labels)
        _check(load_library().dpmm_set_labels(self._ctx, _ptr(lab), _ptr(sub)))

    def get_labels(self):
        lab = np.empty(self.N, np.int32)
        sub = np.empty(self.N, np.int32)
        _check(load_library().dpmm_get_labels(self._ctx, _ptr(lab), _ptr(sub)))
        return lab, sub

    def set_gaussian_params(self, mu, inv_sigma, logdet_sigma, weights, sub_mu, sub_inv_sigma, sub_logdet_sigma,
                            lr_weights, factors=False):
        mu, S, ld, w = _f64(mu), _f64(inv_sigma), _f64(logdet_sigma), _f64(weights)
        smu, sS, sld, lrw = _f64(sub_mu), _f64(sub_inv_sigma), _f64(sub_logdet_sigma), _f64(lr_weights)
        K = w.size
        f = load_library().dpmm_set_gaussian_factors if factors else load_library().dpmm_set_gaussian_params
        _check(f(self._ctx, K, _ptr(mu), _ptr(S), _ptr(ld), _ptr(w), _ptr(smu), _ptr(sS), _ptr(sld), _ptr(lrw)))

    def set_multinomial_params(self, logtheta, weights, sub_logtheta, lr_weights, use_weight=False):
        lt, w, slt, lrw = _f64(logtheta), _f64(weights), _f64(sub_logtheta), _f64(lr_weights)
        K = lt.shape[0]
        _check(load_library().dpmm_set_multinomial_params(self._ctx, K, _ptr(lt), _ptr(w), _ptr(slt), _ptr(lrw),
                                                         int(use_weight)))

    def sample_cluster_labels(self, iteration, final=False):
        _check(load_library().dpmm_sample_cluster_labels(self._ctx, int(iteration), int(bool(final))))

    def sample_sub_labels(self, iteration):
        _check(load_library().dpmm_sample_sub_labels(self._ctx, int(iteration)))

    def log_likelihoods(self, K, which, first, n):
        """Noise-free scores of points [first, first + n) in the sweeps' arithmetic against the K
        clusters whose parameters are set: which = 0 the n x K cluster scores, which = 1 the n x 2
        scores of each point's two sub-clusters."""
        out = np.empty((n, K if which == 0 else 2))
        _check(load_library().dpmm_log_likelihoods(self._ctx, int(K), int(which), int(first), int(n), _ptr(out)))
        return out

    def sufficient_statistics(self, K, want_S=True, counts_only=False):
        Nsub = np.empty(2 * K, np.int64)
        s = None if counts_only else np.empty((2 * K, self.D))
        S = np.empty((2 * K, self.D, self.D)) if (want_S and not counts_only and self.prior == prior_type.Gaussian) else None
        _check(load_library().dpmm_sufficient_statistics(self._ctx, K, _ptr(Nsub), _ptr(s), _ptr(S)))
        return Nsub, s, S

    def set_statistics(self, Nsub, sums, S=None):
        """Test hook: statistics of 2K sub-clusters (counts, sums, scatter) uploaded as if computed."""
        n = _i64(Nsub)
        s, SS = _f64(sums), _f64(S)
        _check(load_library().dpmm_set_statistics(self._ctx, n.size // 2, _ptr(n), _ptr(s), _ptr(SS)))

    def pair_features(self, K, cluster_scalars, pi, pj):
        """16 screener features per pair (pi[p], pj[p]) from the current statistics and K x 7
        per-cluster scalars (N, N_left, N_right, L, L_left, L_right, log det psi_n)."""
        sc = _f64(np.asarray(cluster_scalars).reshape(K, 7))
        a, b = _i32(pi), _i32(pj)
        out = np.empty((a.size, 16), np.float32)
        _check(load_library().dpmm_pair_features(self._ctx, int(K), _ptr(sc), a.size, _ptr(a), _ptr(b), _ptr(out)))
        return out

    def set_prior(self, hyper):
        """niw_hyperparams or multinomial_hyper of the batched posterior work."""
        if isinstance(hyper, niw_hyperparams):
            m, psi = _f64(hyper.m), _f64(hyper.psi)
            _check(load_library().dpmm_set_prior(self._ctx, hyper.k, _ptr(m), hyper.v, _ptr(psi), None))
        else:
            a = _f64(hyper.alpha)
            _check(load_library().dpmm_set_prior(self._ctx, 0.0, None, 0.0, None, _ptr(a)))

    def set_outlier_prior(self, hyper):
        """Prior of the outlier cluster (niw_hyperparams or multinomial_hyper)."""
        if isinstance(hyper, niw_hyperparams):
            m, psi = _f64(hyper.m), _f64(hyper.psi)
            _check(load_library().dpmm_set_outlier_prior(self._ctx, hyper.k, _ptr(m), hyper.v, _ptr(psi), None))
        else:
            a = _f64(hyper.alpha)
            _check(load_library().dpmm_set_outlier_prior(self._ctx, 0.0, None, 0.0, None, _ptr(a)))

    def log_marginal_batch(self, members, prior_sel=None):
        """Log marginal likelihood and log det psi_n of each row of members (n x 4 sub-cluster
        entries of the last statistics, -1 unused); prior_sel (n) picks, bit 0, the posterior's
        prior and, bit 1, the marginal's prior terms (0 the model's, 1 the outlier prior)."""
        mem = _i32(np.asarray(members).reshape(-1, 4))
        n = mem.shape[0]
        L, ld = np.empty(n), np.empty(n)
        sel = None if prior_sel is None else _i32(prior_sel)
        _check(load_library().dpmm_log_marginal_batch_ex(self._ctx, n, _ptr(mem), _ptr(sel), _ptr(L), _ptr(ld)))
        return L, ld

    def sample_params(self, iteration, members, weights, lr_weights, full_mu_cov=0, use_weight=0):
        """Device draws of the 3K parts (cluster, left, right) from their posteriors, packed into
        the sweep's parameter slots."""
        mem = _i32(np.asarray(members).reshape(-1, 4))
        K = mem.shape[0] // 3
        w, lrw = _f64(weights), _f64(lr_weights)
        _check(load_library().dpmm_sample_params(self._ctx, int(iteration), K, _ptr(mem), _ptr(w), _ptr(lrw),
                                                  int(full_mu_cov), int(use_weight)))
        return K

    def drawn_params(self, K):
        mu = np.empty((3 * K, self.D))
        R = np.empty((3 * K, self.D, self.D)) if self.prior == prior_type.Gaussian else None
        ld = np.empty(3 * K) if self.prior == prior_type.Gaussian else None
        _check(load_library().dpmm_get_drawn_params(self._ctx, _ptr(mu), _ptr(R), _ptr(ld)))
        return mu, R, ld

    def relabel(self, iteration, new_label, sub_action):
        nl, act = _i32(new_label), _i32(sub_action)
        _check(load_library().dpmm_relabel(self._ctx, int(iteration), act.size, _ptr(nl), _ptr(act)))

    def set_mixed_tolerance(self, tau):
        _check(load_library().dpmm_set_mixed_tolerance(self._ctx, float(tau)))

    def last_recheck_count(self):
        n = ctypes.c_int64()
        _check(load_library().dpmm_last_recheck_count(self._ctx, ctypes.byref(n)))
        return n.value

    def set_label_pruning(self, mode):
        """Exact pruning of the mixed label sweep: False / 0 off, "adaptive" / 1 (the default),
        True / 2 always."""
        if mode is True or mode is False:
            m = 2 if mode else 0
        else:
            m = 1 if mode == "adaptive" else int(mode)
        _check(load_library().dpmm_set_label_pruning(self._ctx, m))

    def set_fused_sub_labels(self, enable):
        """Sub-label draw fused into the pruned mixed label sweep (default on); False: a separate sweep."""
        _check(load_library().dpmm_set_fused_sub_labels(self._ctx, 1 if enable else 0))

    def set_label_refinement(self, enable):
        """Split-BF16 refinement of the pruned sweep's undecided waves (default on)."""
        _check(load_library().dpmm_set_label_refinement(self._ctx, 1 if enable else 0))

    def set_sweep_workgroups(self, per_chunk):
        """Workgroups per partition chunk of the pruned mixed label sweep (0: the default)."""
        _check(load_library().dpmm_set_sweep_workgroups(self._ctx, int(per_chunk)))

    def set_statistics_refresh(self, every):
        """Statistics passes per full recomputation; the passes between update the statistics from the
        points whose sub-label moved (behind a pruned mixed label sweep that moved no label); <= 1: off."""
        _check(load_library().dpmm_set_statistics_refresh(self._ctx, int(every)))

    def last_statistics_update(self):
        """The last statistics pass of a sweep (the one kernel_ms times, not a resets-only pass): 0 recomputed,
        1 updated from the points whose sub-label moved, 2 updated with the clusters a moved label touched
        recomputed."""
        v = ctypes.c_int()
        _check(load_library().dpmm_last_statistics_update(self._ctx, ctypes.byref(v)))
        return v.value

    def last_label_evaluations(self):
        """Fraction of (wave, cluster) scorings the last pruned label sweep performed, -1 if none."""
        v = ctypes.c_double()
        _check(load_library().dpmm_last_label_evaluations(self._ctx, ctypes.byref(v)))
        return v.value

    def last_sub_recheck_count(self):
        n = ctypes.c_int64()
        _check(load_library().dpmm_last_sub_recheck_count(self._ctx, ctypes.byref(n)))
        return n.value

    def kernel_ms(self):
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        _check(load_library().dpmm_last_kernel_ms(self._ctx, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def sweep_kernel_ms(self):
        """The last pruned mixed label sweep's main kernel alone (ms, HIP events), 0 if none."""
        a = ctypes.c_float()
        _check(load_library().dpmm_last_sweep_kernel_ms(self._ctx, ctypes.byref(a)))
        return a.value

    def synchronize(self):
        _check(load_library().dpmm_synchronize(self._ctx))

    def sweep_iteration(self, K, iteration, final=False):
        """Label sweep, sub-label sweep and statistics of the K clusters whose parameters are set,
        in one call; returns the 2K sub-cluster counts."""
        cnt = np.empty(2 * K, np.int64)
        _check(load_library().dpmm_sweep_iteration(self._ctx, int(iteration), 1 if final else 0, _ptr(cnt)))
        return cnt

    def set_graph_mode(self, enable=True):
        """Record the sweep segment of every iteration into a hipGraph and replay it."""
        _check(load_library().dpmm_set_graph_mode(self._ctx, 1 if enable else 0))

    def set_ground_truth(self, classes):
        """Ground-truth class of every point (values 0 .. C-1), for the device NMI."""
        c = np.ascontiguousarray(classes, dtype=np.int32)
        if c.shape != (self.N,):
            raise ValueError("one class per point")
        _check(load_library().dpmm_set_ground_truth(self._ctx, _ptr(c), int(c.max()) + 1))

    def label_contingency(self, K, num_classes):
        t = np.empty((K, num_classes), np.int64)
        _check(load_library().dpmm_label_contingency(self._ctx, int(K), _ptr(t)))
        return t

    def label_nmi(self, K):
        """NMI of the current labels (0 .. K-1) against the ground truth, on the device."""
        v = ctypes.c_double()
        _check(load_library().dpmm_label_nmi(self._ctx, int(K), ctypes.byref(v)))
        return v.value

    def graph_counts(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(load_library().dpmm_graph_counts(self._ctx, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Sampler:
    """Stateful driver: init_model + init_first_clusters, then one group_step per step()."""

    def __init__(self, dev, hyper, alpha, iters, init_clusters=1, burnout=15, seed=0, max_clusters=float("inf"),
                 precision=None, full_mu_cov=False, mult_use_weight=False, verbose=False, init_labels=None,
                 hard_clustering=False, outlier_weight=0.0, outlier_params=None, host_decisions=False):
        """init_labels (N, 0-based) starts from a given clustering, e.g. an over-split one;
        hard_clustering takes the argmax in every label sweep (global_params.h:38); outlier_weight > 0
        with outlier_params (a hyperparameter object of the prior's kind) runs the reference's outlier
        mode, cluster 0 being the outlier cluster (dp_parallel_sampling.cpp:68-72).  host_decisions keeps
        the reset / split / merge decisions on the host between device launches instead of running each
        iteration on the device with one synchronisation (both take the same decisions)."""
        L = load_library()
        self.dev = dev
        if outlier_weight > 0:
            if outlier_params is None:
                raise ValueError("outlier_weight > 0 needs outlier_params")
            dev.set_outlier_prior(outlier_params)
        if precision is None:
            precision = Precision.FP64 if dev._dtype == 0 else Precision.FP32
        if init_labels is not None:
            lab = _i32(init_labels)
            init_clusters = int(lab.max()) + 1
            dev.set_labels(lab, None)
        self.opts = _Options(prior=int(dev.prior), alpha=float(alpha), iters=int(iters),
                             init_clusters=int(init_clusters), burnout=int(burnout), argmax_sample_stop=5,
                             split_stop=5, max_clusters=float(max_clusters), seed=int(seed),
                             full_mu_cov=int(full_mu_cov), mult_use_weight=int(mult_use_weight),
                             precision=int(precision), verbose=int(verbose),
                             init_from_labels=int(init_labels is not None), hard_clustering=int(bool(hard_clustering)),
                             outlier_mod=float(outlier_weight), host_decisions=int(bool(host_decisions)))
        self._s = ctypes.c_void_p()
        if dev.prior == prior_type.Gaussian:
            self._m0, self._psi0 = _f64(hyper.m), _f64(hyper.psi)
            _check(L.dpmm_sampler_create(dev.handle, ctypes.byref(self.opts), hyper.k, _ptr(self._m0), hyper.v,
                                         _ptr(self._psi0), None, ctypes.byref(self._s)))
        else:
            self._a0 = _f64(hyper.alpha)
            _check(L.dpmm_sampler_create(dev.handle, ctypes.byref(self.opts), 0.0, None, 0.0, None, _ptr(self._a0),
                                         ctypes.byref(self._s)))

    def step(self, iteration):
        _check(load_library().dpmm_sampler_step(self._s, int(iteration)))

    def num_clusters(self):
        return load_library().dpmm_sampler_num_clusters(self._s)

    def weights(self):
        K = self.num_clusters()
        w = np.zeros(K)
        n = np.zeros(K, np.int64)
        _check(load_library().dpmm_sampler_weights(self._s, _ptr(w), _ptr(n), K))
        return w, n

    def trace_sweeps(self, enable=True):
        """Keep the labels and sub-labels each step's sweeps drew (before resets, splits, merges)."""
        if enable:
            self._trace = (np.empty(self.dev.N, np.int32), np.empty(self.dev.N, np.int32))
            _check(load_library().dpmm_sampler_trace_sweep(self._s, _ptr(self._trace[0]), _ptr(self._trace[1])))
        else:
            _check(load_library().dpmm_sampler_trace_sweep(self._s, None, None))
            self._trace = None

    def sweep_trace(self):
        """(labels, sub-labels) of the last step's sweeps; trace_sweeps must be on."""
        return self._trace[0].copy(), self._trace[1].copy()

    def host_ms(self):
        v = ctypes.c_double()
        _check(load_library().dpmm_sampler_host_ms(self._s, ctypes.byref(v)))
        return v.value

    def log_posterior(self):
        """calculate_posterior (dp_parallel_sampling.cpp:218-236) after the last step."""
        v = ctypes.c_double()
        _check(load_library().dpmm_sampler_log_posterior(self._s, ctypes.byref(v)))
        return v.value

    def set_screener(self, screener=None, full_scan_every=10, threshold=0.0, log_pairs=False):
        """Merge screening with a trained screener (None: off); log_pairs keeps the last step's pairs."""
        self._screener = screener   # keep it alive while the sampler uses it
        _check(load_library().dpmm_sampler_set_screener(self._s, screener.handle if screener else None,
                                                        int(full_scan_every), float(threshold), int(bool(log_pairs))))

    def screen_stats(self):
        """(exact pair tests, pairs skipped, accepted merges missed on full scans, full scans)."""
        v = [ctypes.c_int64() for _ in range(4)]
        _check(load_library().dpmm_sampler_screen_stats(self._s, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def pair_log(self):
        """Features, log Hastings ratio and log U of every candidate pair of the last step."""
        n = ctypes.c_int()
        _check(load_library().dpmm_sampler_pair_log(self._s, 0, None, None, None, ctypes.byref(n)))
        f = np.empty((n.value, 16), np.float32)
        h, u = np.empty(n.value), np.empty(n.value)
        _check(load_library().dpmm_sampler_pair_log(self._s, n.value, _ptr(f), _ptr(h), _ptr(u), ctypes.byref(n)))
        return f, h, u

    def set_merge_bound(self, enable):
        """Reject hopeless merge pairs by a bound before their log ML (device iterations, NIW; default on)."""
        _check(load_library().dpmm_sampler_set_merge_bound(self._s, 1 if enable else 0))

    def merge_bound_rejects(self):
        """Merge pairs rejected by the bound, without their log ML, since creation."""
        v = ctypes.c_int64()
        _check(load_library().dpmm_sampler_merge_bound_rejects(self._s, ctypes.byref(v)))
        return v.value

    def events(self):
        """Decisions since creation: {"splits", "merges", "resets", "removed"}."""
        v = [ctypes.c_int64() for _ in range(4)]
        _check(load_library().dpmm_sampler_events(self._s, *[ctypes.byref(x) for x in v]))
        return dict(zip(("splits", "merges", "resets", "removed"), (x.value for x in v)))

    def close(self):
        if self._s:
            load_library().dpmm_sampler_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class dp_parallel_sampling_class:
    """dp_parallel_sampling_class(numLabels, all_data, randomSeed, priorType) (dp_parallel_sampling.h:14).

    all_data is D x N as in the reference (each column a point); float64 or float32 for the Gaussian
    prior, integer counts for the multinomial prior.
    """

    def __init__(self, num_labels, all_data, random_seed, prior):
        all_data = np.asarray(all_data)
        if all_data.ndim != 2 or all_data.shape[1] != num_labels:
            raise ValueError("all_data must be D x numLabels")
        prior = prior_type(prior)
        X = all_data.T
        if prior == prior_type.Multinomial:
            X = np.ascontiguousarray(X, dtype=np.int32)
        elif X.dtype != np.float32:
            X = np.ascontiguousarray(X, dtype=np.float64)
        else:
            X = np.ascontiguousarray(X)
        self.prior = prior
        self.seed = int(random_seed) & 0xFFFFFFFFFFFFFFFF
        self.dev = DeviceSweep(X, prior, self.seed)

    def dp_parallel(self, local_hyper_params, alpha_param, iters=100, init_clusters=1, verbose=True, draw_labels=False,
                    save_model=False, burnout=15, gt=None, max_clusters=float("inf"), outlier_weight=0.0,
                    outlier_params=None, precision=None):
        """dp_parallel (dp_parallel_sampling.cpp:132-160 + run_model 175-249)."""
        s = Sampler(self.dev, local_hyper_params, alpha_param, iters, init_clusters, burnout, self.seed, max_clusters,
                    precision, outlier_weight=outlier_weight, outlier_params=outlier_params)
        mi = ModelInfo()
        if verbose and gt is not None:   # classes 0 .. C-1 in order of first appearance
            _, cls = np.unique(np.asarray(gt), return_inverse=True)
            self.dev.set_ground_truth(cls.astype(np.int32))
        for it in range(1, iters + 1):
            t0 = time.perf_counter()
            s.step(it)
            mi.iter_count.append(time.perf_counter() - t0)
            mi.cluster_count_history.append(s.num_clusters())
            # the reference records the log posterior only in verbose mode, else 1 (dp_parallel_sampling.cpp:192)
            mi.likelihood_history.append(s.log_posterior() if verbose else 1.0)
            if verbose and gt is not None:
                mi.nmi_score_history.append(self.dev.label_nmi(s.num_clusters()))
            if verbose:
                nmi = mi.nmi_score_history[-1] if mi.nmi_score_history else -1
                print(f"Iteration: {it} || Clusters count: {mi.cluster_count_history[-1]} || "
                      f"Log posterior: {mi.likelihood_history[-1]:f} || NMI score: {nmi:f} || "
                      f"Iter Time: {mi.iter_count[-1]:.4f} || Total time: {sum(mi.iter_count):.4f}", flush=True)
        lab, _ = self.dev.get_labels()
        mi.labels = lab + 1
        mi.weights, mi.cluster_counts = s.weights()
        s.close()
        return mi
