"""Throughput of the DPMM sub-cluster sampler iteration on one MI355X.

A step is one full sampler iteration (local_clusters_actions::group_step): parameter draws,
label sweep, sub-label sweep, statistics, split / merge / reset / remove, over the whole synthetic
data set resident in HBM.  The metric is point-iterations per second at steady state: an untimed
burn-in (at least 150 iterations, then until the cluster count has not changed for 20 iterations)
lets the sampler grow from one cluster to the mixture's cluster count, whatever --warmup says;
--warmup iterations follow it untimed, then --steps iterations are timed.  The run exits with
status 3 if fewer than 0.9 x the true cluster count are present when the timed steps start.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--config c3_d64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3, "fp64": 78.6, "bf16": 2500.0}   # dense TFLOP/s, MI355X
HBM_PEAK = 8000.0                                       # GB/s
# HBM bytes per launch of each phase's kernels, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# this script (scaled by what the counters read for a copy of known size), with the workload they were
# measured on; a line carries them only when its own workload (config, precision, cluster count) matches.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(config, precision, clusters, phase):
    try:
        with open(TRAFFIC_FILE) as f:
            rows = json.load(f)["rows"]
    except (OSError, ValueError, KeyError):
        return None
    for r in rows:
        if r["config"] == config and r["precision"] == precision and r["clusters"] == clusters:
            return r["bytes_per_launch"].get(phase)
    return None

CONFIGS = {
    # N, D, true clusters, A scale, ridge, Dirichlet weight (None equal), alpha, dtype
    "c1_readme_2d": dict(N=100_000, D=2, K=2, a=1.0, ridge=1.0, weight=None, alpha=1.0, dtype="float64",
                         iters=100),
    "c2_d32": dict(N=1_000_000, D=32, K=20, a=1.0, ridge=1.0, weight=1.0, alpha=10.0, dtype="float32", iters=200),
    "c3_d64": dict(N=10_000_000, D=64, K=50, a=1.0, ridge=0.5, weight=0.5, alpha=10.0, dtype="float32",
                   iters=200),
    # learned merge screener at scale: started over-split into `pieces` fragments per true cluster
    "c5_screener": dict(N=5_000_000, D=128, K=200, a=0.5, ridge=1.0, weight=None, alpha=10.0, dtype="float32",
                        iters=300, pieces=2),
    # multinomial: N documents of `words` words over a vocabulary of D, topics theta ~ Dir(conc)
    "c4_mult": dict(N=1_000_000, D=1000, K=100, words=100, conc=0.1, alpha=10.0, dtype="int32", prior=1.0,
                    iters=150),
}


def gen_on_gpu(cfg, seed):
    """Mixture of Gaussians generated on the GPU (means ~ U(-100,100)^D, covariances A A^T + ridge I)."""
    import torch
    rng = np.random.default_rng(seed)
    N, D, K = cfg["N"], cfg["D"], cfg["K"]
    means = rng.uniform(-100, 100, size=(K, D))
    facs = []
    for _ in range(K):
        A = rng.normal(0.0, cfg["a"], size=(D, D))
        facs.append(np.linalg.cholesky(A @ A.T + cfg["ridge"] * np.eye(D)))
    w = np.full(K, 1.0 / K) if cfg["weight"] is None else rng.dirichlet(np.full(K, cfg["weight"]))
    dt = torch.float64 if cfg["dtype"] == "float64" else torch.float32
    g = torch.Generator(device="cuda").manual_seed(seed)
    gt = torch.multinomial(torch.tensor(w, device="cuda"), N, replacement=True, generator=g)
    X = torch.empty((N, D), device="cuda", dtype=dt)
    for k in range(K):
        idx = (gt == k).nonzero().squeeze(1)
        if idx.numel() == 0:
            continue
        z = torch.randn((idx.numel(), D), device="cuda", dtype=dt, generator=g)
        Lt = torch.tensor(facs[k].T.copy(), device="cuda", dtype=dt).contiguous()
        X[idx] = z @ Lt + torch.tensor(means[k], device="cuda", dtype=dt)
    torch.cuda.synchronize()
    return X, gt, means, np.stack(facs), w


def gen_docs_on_gpu(cfg, seed):
    """Bag-of-words documents generated on the GPU: topic uniform, `words` draws from theta_topic by
    inverse CDF, counted into a dense N x D int32 matrix."""
    import torch
    rng = np.random.default_rng(seed)
    N, D, K, W = cfg["N"], cfg["D"], cfg["K"], cfg["words"]
    theta = rng.dirichlet(np.full(D, cfg["conc"]), size=K)
    g = torch.Generator(device="cuda").manual_seed(seed)
    gt = torch.randint(0, K, (N,), device="cuda", generator=g)
    X = torch.zeros((N, D), device="cuda", dtype=torch.int32)
    for k in range(K):
        idx = (gt == k).nonzero().squeeze(1)
        if idx.numel() == 0:
            continue
        cdf = torch.tensor(np.cumsum(theta[k]), device="cuda", dtype=torch.float64)
        u = torch.rand((idx.numel(), W), device="cuda", dtype=torch.float64, generator=g) * cdf[-1]
        words = torch.searchsorted(cdf, u).clamp_(max=D - 1)
        cnt = torch.zeros((idx.numel(), D), device="cuda", dtype=torch.int32)
        cnt.scatter_add_(1, words, torch.ones_like(words, dtype=torch.int32))
        X[idx] = cnt
    torch.cuda.synchronize()
    return X, gt, theta


def _draw_current(dev, smp, K, it):
    """Draw every cluster's and sub-cluster's parameters from the sampler's current statistics on
    the device (packed into the sweep's slots), and return them with the mixture weights."""
    w = smp.weights()[0]
    lrw = np.full(2 * K, 0.5)
    dev.sufficient_statistics(K, counts_only=True)
    mem = np.array([[2 * k, 2 * k + 1, -1, -1] if p == 0 else [2 * k + p - 1, -1, -1, -1]
                    for k in range(K) for p in range(3)], np.int32)
    dev.sample_params(it, mem, w, lrw)
    mu, R, ld = dev.drawn_params(K)
    return w, lrw, mu, R, ld


def cpu_baseline_and_check(dev, smp, X, N, mult, seed, it, target_s=8.0, score_points=20_000):
    """The FP64 oracle (OpenMP) timed on a bounded sample of the same workload: one iteration's
    label sweep, sub-label sweep and statistics at parameters the sampler draws from its current
    state.  The same parameters and Philox noise drive one sweep of the measured GPU path over all
    N points, and the oracle's labels on the sample give the label-flip rates; the noise-free scores
    of the first points give the log-likelihood error of the arithmetic that decides the labels."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as orc
    K = smp.num_clusters()
    w, lrw, mu, R, ld = _draw_current(dev, smp, K, it)
    c = np.arange(K) * 3
    sub = np.sort(np.r_[c + 1, c + 2])
    dev.sample_cluster_labels(it, False)
    dev.sample_sub_labels(it)
    glab, gsub = dev.get_labels()
    gs = dev.log_likelihoods(K, 0, 0, score_points)
    gss = dev.log_likelihoods(K, 1, 0, score_points)
    prior = 1 if mult else 0
    if mult:
        cl = dict(logtheta=mu[c])
        sl = dict(logtheta=mu[sub])
    else:
        lam = np.einsum("kij,klj->kil", R, R)
        cl = dict(mu=mu[c], inv_sigma=lam[c], logdet=ld[c])
        sl = dict(mu=mu[sub], inv_sigma=lam[sub], logdet=ld[sub])

    def one(n):
        Xs = np.ascontiguousarray(X[:n].cpu().numpy(), dtype=np.float64)
        t0 = time.perf_counter()
        lab, margin = orc.label_sweep(prior, Xs, weights=w, seed=seed, it=it, **cl)
        subs, smargin = orc.sublabel_sweep(prior, Xs, lab, lr_weights=lrw, seed=seed, it=it, **sl)
        orc.suff_stats(prior, Xs, lab, subs, K, want_S=not mult)
        return time.perf_counter() - t0, Xs, lab, margin, subs, smargin

    n = min(N, 5_000 if mult else 20_000)
    dt = one(n)[0]
    n2 = int(min(N, max(n, n * target_s / max(dt, 1e-3))))
    dt2, Xs, olab, margin, osub, smargin = one(n2)
    reps = 1
    if n2 == N and dt2 < target_s:   # the whole data set is under the target: repeat the iteration
        reps = int(min(200, np.ceil(target_s / max(dt2, 1e-4))))
        dt2 = sum(one(N)[0] for _ in range(reps))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    what = "documents" if mult else "points"
    base = {"value": n2 * reps / dt2, "unit": "point*iters/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle label sweep + sub-label sweep + statistics of {reps} iteration(s) over the "
                      f"first {n2} {what} (K={K}, D={Xs.shape[1]}) at sampler-drawn parameters, {dt2:.1f} s"}
    flips = np.nonzero(glab[:n2] != olab)[0]
    same = np.nonzero(glab[:n2] == olab)[0]
    sflips = same[gsub[:n2][same] != osub[same]]
    # noise-free oracle scores of the first points, clusters and the sub-clusters of the GPU labels
    m = min(score_points, n2)
    _, _, os_ = orc.label_sweep(prior, Xs[:m], weights=w, seed=seed, it=it, final=True, want_scores=True, **cl)
    _, _, oss = orc.label_sweep(prior, Xs[:m], weights=lrw, seed=seed, it=it, final=True, want_scores=True,
                                mult_use_weight=True, **sl)
    k = glab[:m]
    oss = np.stack([oss[np.arange(m), 2 * k], oss[np.arange(m), 2 * k + 1]], axis=1)
    err = np.abs(gs[:m] - os_)
    serr = np.abs(gss[:m] - oss)
    check = {
        "points": n2, "iteration": it, "clusters": K,
        "label_flip_rate": flips.size / n2, "label_flips": int(flips.size),
        "label_flips_above_rel_margin_1e-6": int(np.sum(margin[flips] >= 1e-6)),
        "sub_label_flip_rate": sflips.size / max(1, same.size), "sub_label_flips": int(sflips.size),
        "sub_label_flips_above_rel_margin_1e-6": int(np.sum(smargin[sflips] >= 1e-6)),
        "score_points": m,
        "score_max_abs_err": float(err.max()), "score_mean_abs_err": float(err.mean()),
        "score_max_rel_err": float((err / np.maximum(1.0, np.abs(os_))).max()),
        "sub_score_max_abs_err": float(serr.max()),
        "sub_score_max_rel_err": float((serr / np.maximum(1.0, np.abs(oss))).max()),
    }
    return base, check


def run_seeds(args, cfg):
    """Final K and NMI against the generating labels after the configuration's full run, for seeds
    0..S-1 (data and sampler seeded alike), on each requested path; one JSON line."""
    import torch
    import dpmmsubclusters_gpu_amd as dp
    from dpmmsubclusters_gpu_amd import _nmi
    mult = cfg["dtype"] == "int32"
    N, D, iters = cfg["N"], cfg["D"], cfg["iters"]
    paths = [p for p in args.paths.split(",")
             if not (mult and p == "fp64") and not (cfg["dtype"] == "float64" and p != "fp64")]
    res = {p: {"K": [], "nmi": [], "seconds": []} for p in paths}
    for seed in range(args.seeds):
        if mult:
            X, gt, _ = gen_docs_on_gpu(cfg, seed)
        else:
            X, gt, _, _, _ = gen_on_gpu(cfg, seed)
        gtn = gt.cpu().numpy()
        for path in paths:
            Xp = X.double() if path == "fp64" and not mult else X
            dtype = np.int32 if mult else (np.float64 if Xp.dtype == torch.float64 else np.float32)
            prior = dp.prior_type.Multinomial if mult else dp.prior_type.Gaussian
            dev = dp.DeviceSweep(None, prior, seed=seed + 1, device_ptr=Xp.data_ptr(), N=N, D=D, dtype=dtype)
            hyper = (dp.multinomial_hyper(np.full(D, cfg["prior"])) if mult else
                     dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D)))
            prec = {"mixed": dp.Precision.MIXED, "fp32": dp.Precision.FP32, "fp64": dp.Precision.FP64}[path]
            smp = dp.Sampler(dev, hyper, cfg["alpha"], iters, seed=seed + 1, precision=prec)
            t0 = time.perf_counter()
            for t in range(1, iters + 1):
                smp.step(t)
            dev.synchronize()
            dt = time.perf_counter() - t0
            lab, _ = dev.get_labels()
            res[path]["K"].append(smp.num_clusters())
            res[path]["nmi"].append(_nmi(gtn, lab))
            res[path]["seconds"].append(dt)
            print(f"[bench] seed {seed} {path}: K={smp.num_clusters()} NMI={res[path]['nmi'][-1]:.4f} "
                  f"({dt:.1f} s)", file=sys.stderr, flush=True)
            smp.close()
            dev.close()
            del Xp
        del X
        torch.cuda.empty_cache()
    summary = {p: {"K_mean": float(np.mean(r["K"])), "K_std": float(np.std(r["K"])),
                   "nmi_mean": float(np.mean(r["nmi"])), "nmi_min": float(np.min(r["nmi"])), **r}
               for p, r in res.items()}
    print(json.dumps({"metric": "final K and NMI vs generating labels over seeds", "config": args.config,
                      "N": N, "D": D, "true_clusters": cfg["K"], "iterations": iters, "seeds": args.seeds,
                      "paths": summary}), flush=True)


def run_screener(args, cfg):
    """C5: train the merge screener on pairs logged by the plain sampler on mixtures like the two small
    configurations, then run the workload from its over-split start twice, every candidate pair tested
    exactly and screened with a full scan every 10 iterations; one JSON line."""
    import torch
    import dpmmsubclusters_gpu_amd as dp
    from dpmmsubclusters_gpu_amd import _nmi
    from dpmmsubclusters_gpu_amd import screener as scr
    N, D, iters = cfg["N"], cfg["D"], cfg["iters"]
    t0 = time.perf_counter()
    # training data: merge decisions the plain sampler logs on mixtures of the two small configurations'
    # families (D = 2 with 2 clusters, D = 32 with 20 clusters), started over-split into 3 fragments
    train_runs = [("c1", 101), ("c1", 102), ("c2", 103), ("c2", 104), ("c2", 105), ("c2", 106)]
    Xf, h, u = scr.collect_config_style_pairs(train_runs, n_points=200_000, pieces=3, iters=60,
                                              precision=dp.Precision.MIXED,
                                              log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
    a = h > u
    y = scr.plausible(h)
    s, hist = scr.train(Xf, y, epochs=30, batch=4096, lr=2e-3, seed=2)
    z = s.logits(Xf)
    thr = scr.threshold_for_recall(z, y, recall=1.0, margin=8.0)   # wide margin: pairs unlike the training set
    t_train = time.perf_counter() - t0
    print(f"[bench] training logits: positives min {z[y > 0.5].min():.2f}, negatives max {z[y < 0.5].max():.2f}, "
          f"median {np.median(z[y < 0.5]):.2f}", file=sys.stderr, flush=True)
    print(f"[bench] screener trained on {len(y)} pairs ({int(y.sum())} plausible, {int(a.sum())} accepted) "
          f"in {t_train:.1f} s, threshold {thr:.3f}", file=sys.stderr, flush=True)
    print(f"[bench] generating {N} points", file=sys.stderr, flush=True)
    X, gt, _, _, _ = gen_on_gpu(cfg, args.seed)
    gtn = gt.cpu().numpy()
    print("[bench] generated", file=sys.stderr, flush=True)
    init = scr.over_split_labels(gtn, cfg["pieces"], args.seed)
    hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
    runs = {}
    for name, use in (("all_pairs", False), ("screened", True)):
        dev = dp.DeviceSweep(None, dp.prior_type.Gaussian, seed=args.seed + 1, device_ptr=X.data_ptr(), N=N, D=D,
                             dtype=np.float32)
        smp = dp.Sampler(dev, hyper, cfg["alpha"], iters, seed=args.seed + 1, precision=dp.Precision.MIXED,
                         init_labels=init)
        smp.set_screener(s if use else None, full_scan_every=10, threshold=thr)
        dev.synchronize()
        t1 = time.perf_counter()
        for t in range(1, iters + 1):
            smp.step(t)
            if t <= 3 or t % 10 == 0:
                dev.synchronize()
                lab_ms, sub_ms, st_ms = dev.kernel_ms()
                print(f"[bench] {name} iteration {t}/{iters} K={smp.num_clusters()} "
                      f"({time.perf_counter() - t1:.1f}s; label {lab_ms:.1f} ms, sub {sub_ms:.1f} ms, statistics "
                      f"{st_ms:.1f} ms, conjugate {smp.host_ms():.1f} ms, rechecked {dev.last_recheck_count()} "
                      f"+ {dev.last_sub_recheck_count()}, evaluated {dev.last_label_evaluations():.3f})",
                      file=sys.stderr, flush=True)
        dev.synchronize()
        el = time.perf_counter() - t1
        lab, _ = dev.get_labels()
        st = smp.screen_stats()
        runs[name] = {"seconds": el, "ms_per_iteration": 1000 * el / iters, "point_iters_per_s": N * iters / el,
                      "final_K": smp.num_clusters(), "nmi_vs_truth": _nmi(gtn, lab),
                      "exact_merge_tests": int(st[0]), "pairs_skipped": int(st[1]), "missed_merges": int(st[2]),
                      "full_scans": int(st[3]), "labels": lab}
        smp.close()
        dev.close()
    same = bool(np.array_equal(runs["screened"].pop("labels"), runs["all_pairs"].pop("labels")))
    sr = runs["screened"]
    print(json.dumps({
        "metric": "point*iters/sec", "value": sr["point_iters_per_s"], "unit": "point*iters/s", "n_gpus": 1,
        "steps": iters, "warmup": 0, "ms_per_step": sr["ms_per_iteration"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f32", "data": "synthetic",
        "config": {"workload": f"c5_screener: Gaussian DPMM sampler, N={N}, D={D}, {cfg['K']} true clusters "
                               f"(A ~ N(0, {cfg['a']}^2)), started over-split into {cfg['pieces'] * cfg['K']} "
                               f"clusters, alpha={cfg['alpha']}, {iters} iterations, merge screener with a full "
                               f"scan every 10 iterations", "N": N, "D": D, "precision": "mixed",
                   "parallelism": "single GPU"},
        "screener": {"training_data": "pairs logged by the plain sampler on 2 mixtures like the D=2 README "
                                      "example (2 clusters) and 4 like the D=32 configuration (20 clusters, "
                                      "Dirichlet(1) weights), 200,000 points each, over-split into 3 fragments per "
                                      "cluster, 60 iterations; no D=128 data",
                     "training_pairs": int(len(y)), "plausible": int(y.sum()), "accepted": int(a.sum()),
                     "training_seconds": t_train, "threshold": thr, "final_loss": hist[-1]},
        "all_pairs": runs["all_pairs"], "screened": sr,
        "speedup_vs_all_pairs": runs["all_pairs"]["seconds"] / sr["seconds"],
        "labels_identical": same}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--burn-in", type=int, default=150,
                    help="minimum untimed iterations before --warmup; burn-in then continues until the cluster "
                         "count has not changed for 20 iterations (at most 400 in all)")
    ap.add_argument("--config", default="c3_d64", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32", "fp64"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--host-decisions", action="store_true",
                    help="reset / split / merge decisions on the host between device launches (default: on the "
                         "device, one synchronisation per iteration)")
    ap.add_argument("--graph", action="store_true",
                    help="record each iteration's sweep segment into a hipGraph (no per-phase timings)")
    ap.add_argument("--seeds", type=int, default=0,
                    help="instead of timing: final K and NMI over this many seeds on each of --paths")
    ap.add_argument("--paths", default="mixed,fp32,fp64")
    ap.add_argument("--n", type=int, default=0, help="override the configuration's N (a reduced run, not its bench line)")
    ap.add_argument("--iters", type=int, default=0, help="override the screener workload's iteration count")
    ap.add_argument("--no-fused", action="store_true",
                    help="separate sub-label sweep instead of the draw fused into the pruned label sweep")
    ap.add_argument("--sweep-workgroups", type=int, default=0,
                    help="workgroups per partition chunk of the pruned label sweep (0: the library default)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="copy X once (N*D*4 bytes read + written) before the run: a plain copy kernel of known "
                         "size that calibrates the HBM byte counters of a profiler")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (points would shard across GPUs "
              "with one FP64 all-reduce of the sufficient statistics per iteration; not built).", file=sys.stderr)
        sys.exit(2)

    if args.seeds > 0:
        run_seeds(args, CONFIGS[args.config])
        return
    if args.n > 0:
        CONFIGS[args.config]["N"] = args.n
    if args.iters > 0:
        CONFIGS[args.config]["iters"] = args.iters
    if args.config == "c5_screener":
        run_screener(args, CONFIGS[args.config])
        return

    import torch
    import dpmmsubclusters_gpu_amd as dp

    cfg = CONFIGS[args.config]
    N, D = cfg["N"], cfg["D"]
    mult = cfg["dtype"] == "int32"
    print(f"[bench] generating {args.config}: N={N} D={D} K*={cfg['K']}", file=sys.stderr, flush=True)
    if mult:
        X, gt, theta = gen_docs_on_gpu(cfg, args.seed)
        dev = dp.DeviceSweep(None, dp.prior_type.Multinomial, seed=args.seed + 1, device_ptr=X.data_ptr(), N=N, D=D,
                             dtype=np.int32)
        hyper = dp.multinomial_hyper(np.full(D, cfg["prior"]))
        prec = dp.Precision.MIXED if args.precision == "mixed" else dp.Precision.FP32
    else:
        X, gt, means, facs, w = gen_on_gpu(cfg, args.seed)
        dtype = np.float64 if cfg["dtype"] == "float64" else np.float32
        dev = dp.DeviceSweep(None, dp.prior_type.Gaussian, seed=args.seed + 1, device_ptr=X.data_ptr(), N=N, D=D,
                             dtype=dtype)
        hyper = dp.niw_hyperparams(1.0, np.zeros(D), D + 3.0, np.eye(D))
        if dtype == np.float64:
            prec = dp.Precision.FP64
        else:
            prec = dp.Precision.MIXED if args.precision == "mixed" else dp.Precision.FP32
    if args.copy_probe:
        Y = X.clone()
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {X.numel() * X.element_size()} bytes", file=sys.stderr, flush=True)
        del Y
    max_burn = max(400, args.burn_in)
    # the schedule's final / no-more-splits iterations lie beyond every iteration run here
    iters = max_burn + args.warmup + args.steps + 20
    dev.set_graph_mode(args.graph)
    if args.sweep_workgroups:
        dev.set_sweep_workgroups(args.sweep_workgroups)
    if args.no_fused:
        dev.set_fused_sub_labels(False)
    smp = dp.Sampler(dev, hyper, cfg["alpha"], iters, seed=args.seed + 1, precision=prec,
                     host_decisions=args.host_decisions)
    decisions = "host" if args.host_decisions else "device"
    t_w = time.perf_counter()
    t, last_change, k_prev = 0, 0, smp.num_clusters()
    while t < max_burn and (t < args.burn_in or t - last_change < 20):
        t += 1
        smp.step(t)
        k = smp.num_clusters()
        if k != k_prev:
            last_change, k_prev = t, k
        if t % 25 == 0:
            print(f"[bench] burn-in {t} K={k} ({time.perf_counter() - t_w:.1f}s)", file=sys.stderr, flush=True)
    burn_in = t
    print(f"[bench] burn-in done after {burn_in} iterations, K={smp.num_clusters()} (unchanged since iteration "
          f"{last_change})", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        t += 1
        smp.step(t)
    torch.cuda.synchronize()
    dev.synchronize()
    k_start = smp.num_clusters()
    if k_start < 0.9 * cfg["K"]:
        print(f"[bench] only {k_start} clusters after burn-in and warmup ({cfg['K']} true): not the steady-state "
              "workload, no bench line", file=sys.stderr, flush=True)
        sys.exit(3)
    first = t + 1
    lab_ms, sub_ms, st_ms, host_ms, flops, rechk, srechk, evfrac, ks = [], [], [], [], [], [], [], [], []
    main_ms, st_mode = [], []
    t0 = time.perf_counter()
    for t in range(first, first + args.steps):
        kc = smp.num_clusters()
        ks.append(kc)
        smp.step(t)
        a, b, c = dev.kernel_ms()   # HIP events of the phases (the step has synchronised already)
        lab_ms.append(a); sub_ms.append(b); st_ms.append(c); host_ms.append(smp.host_ms())
        main_ms.append(dev.sweep_kernel_ms())   # the pruned sweep's main kernel alone
        flops.append(2.0 * kc * D * (1 if mult else D) * N)
    dev.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = 1000 * elapsed / args.steps
    value = N * args.steps / elapsed
    # after the timed steps, untimed ones for what needs a copy and a synchronisation of its own: the
    # mixed sweeps' recheck and pruning counters, and which statistics passes recomputed every point
    # (the statistics' FP64 rate is taken from those); the timed steps' statistics times stand otherwise
    t = first + args.steps - 1
    st_ms_t, st_ms = st_ms, []
    for u in range(10 if not mult else 5):
        t += 1
        smp.step(t)
        if not mult:
            st_ms.append(dev.kernel_ms()[2])
            st_mode.append(dev.last_statistics_update())
        if prec == dp.Precision.MIXED and u < 5:
            rechk.append(dev.last_recheck_count())
            if not mult:
                srechk.append(dev.last_sub_recheck_count())
                evfrac.append(dev.last_label_evaluations())
    if mult or not st_ms:
        st_ms, st_mode = st_ms_t, [0] * len(st_ms_t)
    if args.graph:   # phase events are not recorded inside a graph: the whole step only
        inst, upd = dev.graph_counts()
        print(json.dumps({"metric": "point*iters/sec", "value": value, "unit": "point*iters/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
                          "higher_is_better": True, "config": {"workload": args.config, "precision": args.precision,
                                                               "graph_mode": True, "decisions": decisions,
                                                               "clusters_timed": int(round(float(np.mean(ks))))},
                          "graphs_instantiated": inst, "graphs_updated_in_place": upd,
                          "final_clusters": smp.num_clusters()}), flush=True)
        return
    avg_lab = float(np.mean(lab_ms))
    avg_sub, avg_st = float(np.mean(sub_ms)), float(np.mean(st_ms_t))   # the timed steps
    pruned = bool(evfrac) and float(np.mean(evfrac)) >= 0.0

    def hbm(kernel, bytes_launch, ms, what):
        return {"bound": "hbm", "kernel": kernel, "achieved": bytes_launch / (ms * 1e-3) / 1e9, "peak": HBM_PEAK,
                "unit": "GB/s", "traffic": None, "algorithmic": what}

    # one roofline per phase; the bench line's `roofline` is the phase that takes the longest
    phase = {}
    if mult:   # HBM-bound: one read of the N x D int32 counts, one label write
        phase["label_sweep"] = hbm("k_label_mult_mx + k_label_mult_list" if prec == dp.Precision.MIXED else
                                   "k_label_mult", N * (4.0 * D + 4.0), avg_lab,
                                   "4 D + 4 bytes per document (counts read once, label written)")
        phase["label_sweep"]["flops_TFLOP/s"] = float(np.mean(flops)) / (avg_lab * 1e-3) / 1e12
        phase["sub_sweep"] = hbm("k_sub_mult", N * (4.0 * D + 12.0), avg_sub,
                                 "4 D + 12 bytes per document (gathered counts, index, label, sub-label)")
        phase["statistics"] = hbm("k_stats_mult", N * (4.0 * D + 8.0), avg_st,
                                  "4 D + 8 bytes per document (gathered counts, index, sub-label)")
    else:
        esz = 8.0 if prec == dp.Precision.FP64 else 4.0
        dpad = 32 * ((D + 31) // 32)
        fused = pruned and dpad <= 64 and not args.no_fused
        if fused:   # the sub-label draw rides in the pruned sweep: one gathered read of each point
            avg_main = float(np.mean(main_ms))
            phase["label_sub_sweep"] = hbm(
                "k_label_gauss_mxp (fused sub-labels)", N * (esz * D + 12.0), avg_main,
                "D coordinates + index + previous and new sub-label per point, rows gathered in partition order "
                f"({float(np.mean(evfrac)):.3f} of the (wave, cluster) scorings done); this kernel's own HIP events")
            phase["label_sub_sweep"]["phase_ms"] = avg_lab
            phase["label_sub_sweep"]["phase_kernels"] = ("k_prune_prep + partition + k_label_gauss_mxp + "
                                                         "k_label_gauss_listp + k_sub_gauss_list")
        elif pruned:   # exact pruning leaves about one cluster scored per point: a gather of the points
            phase["label_sweep"] = hbm("k_prune_prep + partition + k_label_gauss_mxp + k_label_gauss_list",
                                       N * (esz * D + 12.0), avg_lab,
                                       "D coordinates + index + label read + label write per point "
                                       f"({float(np.mean(evfrac)):.3f} of the (wave, cluster) scorings done)")
        else:
            peak = PEAK[{dp.Precision.FP64: "fp64", dp.Precision.FP32: "fp32", dp.Precision.MIXED: "bf16"}[prec]]
            phase["label_sweep"] = {
                "bound": "mfma",
                "kernel": "k_label_gauss_mx + k_label_gauss_list" if prec == dp.Precision.MIXED else "k_label_gauss",
                "achieved": float(np.mean(flops)) / (avg_lab * 1e-3) / 1e12, "peak": peak, "unit": "TFLOP/s",
                "traffic": None,
                "algorithmic": "2 K D^2 flops per point (dense whitening of K clusters)"}
        if not fused:
            phase["sub_sweep"] = hbm("k_sub_gauss_mx + k_sub_gauss_list" if prec == dp.Precision.MIXED else
                                     "k_sub_gauss", N * (esz * D + 12.0), avg_sub,
                                     "D coordinates + index + label + sub-label write per point (gathered)")
        # the statistics' FP64 rate from the full recomputations; the passes between update the statistics
        # from the points whose sub-label moved and are reported by their time only
        full = [m for m, u in zip(st_ms, st_mode) if u == 0]
        upd = [m for m, u in zip(st_ms, st_mode) if u != 0]
        t_full = float(np.mean(full)) if full else float("nan")
        phase["statistics"] = {
            "bound": "mfma", "kernel": ("k_stats_gauss" if D <= 8 else "k_stats_gauss_wave" if D <= 64 else
                                        "k_stats_gauss_mfma") + " + k_stats_reduce",
            "achieved": N * (D * (D + 1.0) + D) / (t_full * 1e-3) / 1e12 if full else None, "peak": PEAK["fp64"],
            "unit": "TFLOP/s", "traffic": None,
            "algorithmic": "D (D + 1) + D flops per point (upper triangle of x x^T and the sum, FP64), full "
                           "recomputations only",
            "full_passes": len(full), "full_pass_ms": t_full if full else None,
            "update_passes": len(upd), "update_pass_ms": float(np.mean(upd)) if upd else None}
    k_timed = int(round(float(np.mean(ks))))
    for name, r in phase.items():
        r["frac"] = r["achieved"] / r["peak"] if r["achieved"] is not None else None
        r["traffic"] = measured_traffic(args.config, args.precision, k_timed, name)
    phase_ms = {"label_sweep": avg_lab, "label_sub_sweep": avg_lab, "sub_sweep": avg_sub, "statistics": avg_st}
    dominant = max(phase, key=lambda k: phase_ms[k])
    roof = dict(phase[dominant], phase=dominant)
    K_final = smp.num_clusters()
    from dpmmsubclusters_gpu_amd import _nmi
    lab, _ = dev.get_labels()
    nmi = _nmi(gt.cpu().numpy(), lab)
    out = {
        "metric": "point*iters/sec", "value": value, "unit": "point*iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": ("i32+f32" if mult else
                  {dp.Precision.FP64: "f64", dp.Precision.FP32: "f32", dp.Precision.MIXED: "bf16+f32"}[prec]),
        "data": "synthetic",
        "config": {"workload": (f"{args.config}: multinomial DPMM sampler iteration, N={N} documents x "
                                f"{cfg['words']} words, vocabulary D={D}, {cfg['K']} topics theta~Dir({cfg['conc']}), "
                                f"alpha={cfg['alpha']}, Dirichlet prior {cfg['prior']}") if mult else
                               (f"{args.config}: Gaussian DPMM sampler iteration, N={N}, D={D}, "
                                f"{cfg['K']} true clusters, alpha={cfg['alpha']}, NIW(k=1, m=0, v=D+3, psi=I)"),
                   "N": N, "D": D, "clusters_timed": k_timed, "clusters_after": K_final,
                   "burn_in_iterations": burn_in, "precision": args.precision, "decisions": decisions,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "roofline_by_phase": phase,
        # the rest of a step: parameter draw, log MLs, decisions, relabel passes, launch gaps and the
        # synchronisation(s); with host decisions, `host_decisions` of it is host time between launches
        "breakdown_ms": dict({("label_sub_sweep" if "label_sub_sweep" in phase else "label_sweep"): avg_lab,
                              ("sub_label_order" if "label_sub_sweep" in phase else "sub_sweep"): avg_sub,
                              "statistics": avg_st, "rest": ms_step - avg_lab - avg_sub - avg_st},
                             **({"host_decisions": float(np.mean(host_ms))} if args.host_decisions else {})),
        "final_nmi_vs_truth": nmi,
        "recheck_fraction": (float(np.mean(rechk)) / N) if rechk else None,
        "sub_recheck_fraction": (float(np.mean(srechk)) / N) if srechk else None,
        "label_evaluation_fraction": float(np.mean(evfrac)) if evfrac else None,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline and flip check (FP64 oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"], out["oracle_check"] = cpu_baseline_and_check(
            dev, smp, X, N, mult, args.seed + 1, t + 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
