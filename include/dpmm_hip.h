/*
 * C ABI of the MI355X (gfx950) sub-cluster sweep: the device layer of the DPMM split/merge
 * sampler, drop-in for the reference's cudaKernel plugin (prior::get_cuda(),
 * priors/prior.h:29; cudaKernel.cuh:31-138) plus the host driver that replaces
 * dp_parallel_sampling_class::dp_parallel() (dp_parallel_sampling.h:17-29).
 *
 * Plain pointers and sizes only.  Every function returns DPMM_OK (0) or a negative
 * DPMM_E* code; dpmm_last_error() describes the last failure.  The library owns no caller
 * memory.  Conventions: points are row-major N x D (each point contiguous, the reference's
 * D x N column-major layout, cudaKernel.cu:623-624); labels are 0-based here (the C++
 * drop-in layer, include/dpmm/, converts to the reference's 1-based labels); sub-labels are
 * 0 (the reference's 1, "left") or 1 (the reference's 2, "right").
 */
#ifndef DPMM_HIP_H
#define DPMM_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DPMM_OK = 0,
    DPMM_EINVAL = -1,   /* bad argument (null pointer, size, K = 0, D mismatch, NaN) */
    DPMM_EHIP = -2,     /* HIP runtime error */
    DPMM_ENOMEM = -3,   /* device or host allocation failed */
    DPMM_ENOTSPD = -4,  /* a covariance / precision matrix is not symmetric positive definite */
    DPMM_ESTATE = -5    /* call out of order (e.g. statistics before labels) */
};
enum { DPMM_PRIOR_GAUSSIAN = 0, DPMM_PRIOR_MULTINOMIAL = 1 };        /* prior_type, global_params.h:14-18 */
enum { DPMM_DTYPE_F64 = 0, DPMM_DTYPE_F32 = 1, DPMM_DTYPE_I32 = 2 };
/* Arithmetic of the label sweep: FP64, FP32, or BF16 MFMA with FP32 near-tie recompute. */
enum { DPMM_PREC_FP64 = 0, DPMM_PREC_FP32 = 1, DPMM_PREC_MIXED = 2 };

typedef struct dpmm_ctx dpmm_ctx;
typedef struct dpmm_sampler dpmm_sampler;

const char* dpmm_last_error(void);
/* Test hook (no counterpart in the reference, whose device layer exits on any failure): after `after`
 * further successful device allocations of the contexts, every allocation fails as out of memory, so
 * that tests can check each entry point reports DPMM_ENOMEM and leaves the context usable; -1 turns the
 * injection off. */
int dpmm_debug_fail_allocations(int64_t after);
int dpmm_device_count(int* count);

/* cudaKernel::init (cudaKernel.cu:537-635): copies X (host) to HBM, N x D of `dtype`. */
int dpmm_create(const void* X, int64_t N, int D, int dtype, int prior, uint64_t seed, dpmm_ctx** out);
/* Same, with X already in device memory (not copied; must outlive the context). */
int dpmm_create_device(const void* dX, int64_t N, int D, int dtype, int prior, uint64_t seed, dpmm_ctx** out);
/* cudaKernel::release (global_params.cpp:108-114). */
int dpmm_destroy(dpmm_ctx* ctx);
int dpmm_set_precision(dpmm_ctx* ctx, int precision);
/* Largest cluster count the context's kernels take (2048 Gaussian, 433 multinomial).  The
 * reference has no such limit; the host driver stops splitting at it (as run_model stops at
 * max_clusters, dp_parallel_sampling.cpp:147) and accepts splits only up to it. */
int dpmm_max_clusters(dpmm_ctx* ctx, int* kmax);
/* Lowers the cap dpmm_max_clusters reports (0 restores the kernels' own), e.g. to bound the
 * memory of a run; no counterpart in the reference. */
int dpmm_limit_clusters(dpmm_ctx* ctx, int kmax);

/* cudaKernel::sample_labels + sample_sub_labels (cudaKernel.cu:1093-1111): uniform labels
 * in 0..init_clusters-1 and uniform sub-labels. */
int dpmm_sample_initial_labels(dpmm_ctx* ctx, int init_clusters);
/* Test hook of the reference's myCudaKernel set_labels / set_sub_labels (test/myCudaKernel.h). */
int dpmm_set_labels(dpmm_ctx* ctx, const int32_t* labels, const int32_t* sub_labels);
/* cudaKernel::get_labels / get_sub_labels (cudaKernel.cu:1113-1143). */
int dpmm_get_labels(dpmm_ctx* ctx, int32_t* labels, int32_t* sub_labels);

/* Per-iteration cluster parameters, the thin_cluster_params broadcast (ds.h:214-227,
 * local_clusters_actions.cpp:449-458), in the reference's mv_gaussian representation
 * (distributions/mv_gaussian.h: mu, invSigma, logdetSigma): K clusters with mixture
 * weights, and 2K sub-clusters (index 2k + s) with lr_weights (K x 2).  The library
 * factors invSigma = R R^T in FP64.  Returns DPMM_ENOTSPD if a matrix is not SPD. */
int dpmm_set_gaussian_params(dpmm_ctx* ctx, int K, const double* mu, const double* inv_sigma,
                             const double* logdet_sigma, const double* weights, const double* sub_mu,
                             const double* sub_inv_sigma, const double* sub_logdet_sigma,
                             const double* lr_weights);
/* Same with the lower Cholesky factor R of invSigma given (the sampler's own draws produce
 * it directly from the Bartlett decomposition). */
int dpmm_set_gaussian_factors(dpmm_ctx* ctx, int K, const double* mu, const double* chol_inv_sigma,
                              const double* logdet_sigma, const double* weights, const double* sub_mu,
                              const double* sub_chol_inv_sigma, const double* sub_logdet_sigma,
                              const double* lr_weights);
/* multinomial_dist (distributions/multinomial_dist.h): log theta per cluster (K x D) and per
 * sub-cluster (2K x D).  use_weight = 0 reproduces the reference's label log-likelihood,
 * which leaves out log pi_k (cudaKernel_multinomial.cu:60-76). */
int dpmm_set_multinomial_params(dpmm_ctx* ctx, int K, const double* logtheta, const double* weights,
                                const double* sub_logtheta, const double* lr_weights, int use_weight);

/* cudaKernel::create_clusters_labels (cudaKernel.cu:1400-1490): one fused pass over X that
 * scores every point against all K clusters and draws its label (Gumbel-max on Philox
 * noise of iteration `iter`), or takes the argmax when final != 0. */
int dpmm_sample_cluster_labels(dpmm_ctx* ctx, uint32_t iter, int final_iter);
/* cudaKernel::create_subclusters_labels (cudaKernel.cu:1299-1398): deterministic partition
 * of the points by label, then one segmented pass drawing each point's sub-label. */
int dpmm_sample_sub_labels(dpmm_ctx* ctx, uint32_t iter);
/* The log-likelihood rows the reference writes into its N x K matrix d_r before the draw
 * (cudaKernel_gaussian.cu:43-113, cudaKernel_multinomial.cu:37-76, called from cudaKernel.cu:1331
 * and 1430-1434), for points [first, first + n), in the sweeps' own arithmetic and without noise:
 * which = 0 the n x K scores log pi_k + log p(x | theta_k) (multinomial: without log pi_k unless
 * use_weight), which = 1 the n x 2 scores log pi_{k,s} + log p(x | theta_{k,s}) of the two
 * sub-clusters of each point's current label.  Gaussian scores carry the reference's constant
 * -D^2 log(2 pi) / 2 (cudaKernel_gaussian.cu:75, 108), added in FP64 to the sweep's value.
 * Row-major into out (FP64).  K must equal the cluster count of the parameters set. */
int dpmm_log_likelihoods(dpmm_ctx* ctx, int K, int which, int64_t first, int64_t n, double* out);
/* cudaKernel::create_sufficient_statistics (cudaKernel.cu:844-941) for all K clusters at
 * once: per sub-cluster q = 2k + s: count Nsub[q], sum[q*D + d], and (Gaussian) the
 * uncentred scatter S[q*D*D + a*D + b] in FP64; S may be NULL.  Deterministic: no
 * atomics on bins; a fixed reduction order. */
int dpmm_sufficient_statistics(dpmm_ctx* ctx, int K, int64_t* Nsub, double* sum, double* S);
/* The device part of local_clusters_actions::group_step's first half (local_clusters_actions.cpp:
 * 462-467: create_clusters_labels, create_subclusters_labels, then create_sufficient_statistics;
 * cudaKernel.cu:1400-1490, 1299-1398, 844-941) in one call: label sweep, sub-label sweep and the
 * statistics of the new labels, with their 2K counts returned in Nsub (the statistics stay on
 * the device for dpmm_log_marginal_batch).  With graph mode on (dpmm_set_graph_mode), the call
 * records the whole segment into a hipGraph, updates one executable graph in place with it while
 * the recorded work keeps its shape (else instantiates it anew), and launches it; results are
 * identical to the eager calls.  Graph mode has no counterpart in the reference; in it the
 * per-phase timings of dpmm_last_kernel_ms are not recorded. */
int dpmm_sweep_iteration(dpmm_ctx* ctx, uint32_t iter, int final_iter, int64_t* Nsub);
int dpmm_set_graph_mode(dpmm_ctx* ctx, int enable);
/* Executable graphs instantiated and updated in place so far in graph mode. */
int dpmm_graph_counts(dpmm_ctx* ctx, int64_t* instantiations, int64_t* updates);
/* One relabel pass for split / merge / reset / remove (cudaKernel.cu:172-263): a point
 * with label k and sub-label s gets label new_label[2k + s]; its sub-label becomes
 * s (sub_action[k] = 0), a Philox draw of tag 4 (reset, 1) or 5 (split, 2), or the
 * constant 0 (3) / 1 (4).  Tables have K_old entries (new_label 2 K_old). */
int dpmm_relabel(dpmm_ctx* ctx, uint32_t iter, int K_old, const int32_t* new_label, const int32_t* sub_action);

/* --------------------------------------------------- batched conjugate FP64 work on the device ----- */
/* Prior of the posteriors below: niw_hyperparams(k0, m0 (D), v0, psi0 (D x D row-major, the
 * reference's psi = Psi / v0)) (priors/niw_hyperparams.h:13, 40-43) for a Gaussian context;
 * multinomial_hyper(alpha0 (D)) (priors/multinomial_hyper.h:13, 29) for a multinomial one.
 * DPMM_ENOTSPD if psi0 is not SPD. */
int dpmm_set_prior(dpmm_ctx* ctx, double k0, const double* m0, double v0, const double* psi0, const double* alpha0);
/* The outlier cluster's prior, outlier_params of dp_parallel (dp_parallel_sampling.h:28-29,
 * local_clusters_actions.cpp:48-61), in the form of dpmm_set_prior: jobs and parts below select it
 * with their prior selectors. */
int dpmm_set_outlier_prior(dpmm_ctx* ctx, double k0, const double* m0, double v0, const double* psi0,
                           const double* alpha0);
/* Test hook beside dpmm_set_labels (the reference's test/myCudaKernel.h): statistics of 2K
 * sub-clusters in the layout of dpmm_sufficient_statistics, for dpmm_log_marginal_batch and
 * dpmm_sample_params; the next dpmm_sufficient_statistics call recomputes them from the labels. */
int dpmm_set_statistics(dpmm_ctx* ctx, int K, const int64_t* Nsub, const double* sum, const double* S);
/* Log marginal likelihoods of n statistic sums in one launch (one workgroup per sum).  Sum j adds the
 * sub-cluster entries members[4j .. 4j+3] (-1: unused; entry q = 2k + s) of the statistics of the
 * last dpmm_sufficient_statistics call, forms the posterior (niw::calc_posterior, niw.cpp:23-49;
 * multinomial_prior::calc_posterior, multinomial_prior.cpp:12-22), factors it in FP64 and
 * evaluates niw::log_marginal_likelihood (niw.cpp:122-138) or the Dirichlet-multinomial one
 * (multinomial_prior.cpp:56-67).  One call serves a whole iteration: the clusters, sub-clusters and
 * merge pairs of update_suff_stats_posterior / should_split_local / should_merge
 * (local_clusters_actions.cpp:103-168, 215-242; shared_actions.cpp:88-111).  logdet_psi (may be
 * NULL) gets log det psi_n; a non-SPD posterior gives NaN. */
int dpmm_log_marginal_batch(dpmm_ctx* ctx, int n, const int32_t* members, double* logml, double* logdet_psi);
/* Same, with a prior selector per sum (prior_sel may be NULL: all 0): bit 0 picks the prior the posterior
 * is formed from, bit 1 the prior whose terms the log marginal likelihood subtracts (0 the model's, 1 the
 * outlier's).  The reference takes both separately (niw::log_marginal_likelihood(prior, posterior, stats),
 * niw.cpp:122-138) and its split test forms a sub-cluster's posterior from the cluster's prior
 * (local_clusters_actions.cpp:221-228). */
int dpmm_log_marginal_batch_ex(dpmm_ctx* ctx, int n, const int32_t* members, const int32_t* prior_sel, double* logml,
                               double* logdet_psi);
/* sample_cluster_params (shared_actions.cpp:9-55; local_clusters_actions.cpp:341-371) for K clusters
 * in one launch: part p = 3k + r (r: 0 cluster, 1 left, 2 right sub-cluster) is drawn from the
 * posterior of the statistic sum members[4p .. 4p+3] (niw::sample_distribution niw.cpp:82-107, Bartlett
 * form; multinomial_prior::sample_distribution multinomial_prior.cpp:24-49) on the Philox stream
 * (seed, p, iter, 6), and the sweeps' parameter slots are written on the device with log weights[k]
 * (clusters) and log lr_weights[2k + s] (sub-clusters).  full_mu_cov = 0 keeps the reference's
 * diagonal-only mu draw (pdflib.cpp:4007-4014); use_weight = 0 its multinomial label score without
 * log pi.  Replaces dpmm_set_*_params for the sampler.  DPMM_ENOTSPD if a draw fails. */
int dpmm_sample_params(dpmm_ctx* ctx, uint32_t iter, int K, const int32_t* members, const double* weights,
                       const double* lr_weights, int full_mu_cov, int use_weight);
/* Same, with 4 control entries per part (ctl may be NULL): [0] a row of the previous call's draws whose
 * parameters the part keeps (-1: draw it), [1], [2] the slot and iteration of its Philox stream (-1: p and
 * iter), [3] the prior of its posterior (0 the model's, 1 the outlier's).  Kept rows serve the reference's
 * outlier mode, whose sample_clusters leaves cluster index 1 undrawn (local_clusters_actions.cpp:346-352),
 * and draws keyed elsewhere reproduce parameters drawn at a split (shared_actions.cpp:57-70). */
int dpmm_sample_params_ex(dpmm_ctx* ctx, uint32_t iter, int K, const int32_t* members, const int32_t* ctl,
                          const double* weights, const double* lr_weights, int full_mu_cov, int use_weight);
/* The FP64 draws of the last dpmm_sample_params, 3K parts: mu (or log theta) 3K x D, the lower
 * factor R of invSigma 3K x D x D (may be NULL), logdet Sigma 3K (may be NULL). */
int dpmm_get_drawn_params(dpmm_ctx* ctx, double* mu, double* chol_inv_sigma, double* logdet_sigma);

/* Mixed precision (DPMM_PREC_MIXED): a point's BF16-screened label is kept when its best score,
 * lowered by its error bound e_b, still beats every other cluster's score raised by e_k; otherwise
 * the point is re-scored exactly in FP32.  e_k bounds the difference between the screened score and
 * the FP32 sweep's score of the same point and cluster, from sum_l |W_jl| |x_l - mu_l| per row (a
 * second MFMA), so the kept labels are the FP32 sweep's whatever the covariance's conditioning.  tau
 * scales the bound: 1 (default) the derived bound, 0 keeps every screened label (a diagnostic of the
 * raw BF16 screen), up to 16 widens it.  No counterpart in the reference (whose sweep is FP64). */
int dpmm_set_mixed_tolerance(dpmm_ctx* ctx, float tau);
/* The per-row error factors of the mixed screens for padded width DP (32, 64, 96 or 128), in out[0..3]:
 * the single-product label screen's, the split screens' (four hi / lo products), and the relative
 * sub-cluster screen's eps_a and eps_g.  Each screen keeps |y_screen,j - y_fp32,j| within its factor
 * times sum_l |W_jl| |d_l| (eps_g: times sum_l |W_jl| |mu_h,l - mu_a,l| + |c_j|).  Needs no GPU.  No
 * counterpart in the reference (its sweeps have no reduced-precision screen). */
int dpmm_screen_error_factors(int DP, float* out);
/* Number of points the last mixed label sweep re-scored in FP32 (-1 if none ran). */
int dpmm_last_recheck_count(dpmm_ctx* ctx, int64_t* n);
/* Number of points whose sub-label the last mixed sub-label sweep drew in FP32 (-1 if none ran).
 * The mixed sub-label sweep screens s_left - s_right with split-BF16 MFMA and a per-point error
 * bound against the FP32 sweep's arithmetic; no counterpart in the reference. */
int dpmm_last_sub_recheck_count(dpmm_ctx* ctx, int64_t* n);
/* Exact pruning of the mixed Gaussian label sweep (on by default; no counterpart in the
 * reference).  Points are visited grouped by their labels before the sweep; for a wave of points
 * of cluster a, a cluster k is skipped when a bound from the triangle inequality in k's whitened
 * metric (||W_k (mu_a - mu_k)|| - ||W_k||_2 max ||x - mu_a||) plus the largest Gumbel noise of a
 * 32-bit uniform shows that no point of the wave can take label k in the FP32 sweep.  Labels are
 * identical with pruning on or off.  Pruning needs D to be a multiple of 4 (other D sweep
 * unpruned).  mode 0: off; 1 (default): adaptive, pruning only sweeps with
 * N K DP^2 >= 4e11 (DP: D padded) and following a sweep that still scored more than a quarter of
 * the (wave, cluster) pairs by four unpruned sweeps; 2: always. */
int dpmm_set_label_pruning(dpmm_ctx* ctx, int mode);
/* Statistics updated from the points that moved (no counterpart in the reference, which recomputes
 * create_sufficient_statistics from every point each iteration, cudaKernel.cu:844-941).  After a pruned
 * mixed label sweep that moved no label, the partition of the labels stands and the statistics of the
 * iteration are the last ones plus the sums of the points whose sub-label moved: (k, 0) gains the points
 * that moved into sub-cluster 0 and loses those that moved out, (k, 1) the reverse.  `every` statistics
 * passes hold one full recomputation, which bounds the FP64 rounding the updates carry (default 8;
 * <= 1: every pass recomputes).  Counts are exact either way. */
int dpmm_set_statistics_refresh(dpmm_ctx* ctx, int every);
/* The last statistics pass of a sweep (the one dpmm_last_kernel_ms times; a pass that only updates the
 * clusters reset by the decisions is not counted): 0 a full recomputation, 1 an update of every cluster
 * from the points whose sub-label moved, 2 such an update with the clusters a moved label touched
 * recomputed.  Synchronises. */
int dpmm_last_statistics_update(dpmm_ctx* ctx, int* moved_only);
/* Sub-label draw fused into the pruned mixed label sweep (on by default, padded D <= 64; no counterpart
 * in the reference, whose sub-label sweep is a pass of its own, cudaKernel.cu:1299-1398).  A point
 * whose label stays the cluster it was filed under is screened against that cluster's two
 * sub-clusters in the same pass over its coordinates, with the split-BF16 screen and error bound of
 * the separate sweep; points whose label moved and undecided ones are drawn in FP32 afterwards by
 * their final labels.  Sub-labels are identical with fusion on or off (same Philox draws); the next
 * dpmm_sample_sub_labels of the same iteration only orders them for the statistics pass.
 * enable 0: the separate sub-label sweep. */
int dpmm_set_fused_sub_labels(dpmm_ctx* ctx, int enable);
/* Split-BF16 refinement in the pruned mixed label sweep (on by default; no counterpart in the
 * reference).  A wave whose BF16 screen leaves points undecided scores the cluster they were filed
 * under and every cluster the prune test keeps again with W and x - mu split into BF16 highs and lows
 * (three MFMAs) and a bound some 2^-7 of the screen's; a point takes the best cluster when its lower
 * bound beats every other upper bound, which makes it the FP32 sweep's label.  Near-ties between
 * sibling clusters (after a split, or an over-split start) then rarely reach the FP32 recheck.
 * enable 0: every undecided point goes to the FP32 recheck. */
int dpmm_set_label_refinement(dpmm_ctx* ctx, int enable);
/* Workgroups per partition chunk of the pruned mixed label sweep (a tuning knob, no counterpart in
 * the reference): each takes an equal share of the chunk's points and stages the chunk's cluster slots
 * in LDS once.  0 (default): 2 with the fused sub-label draw, else 4; at most 64. */
int dpmm_set_sweep_workgroups(dpmm_ctx* ctx, int per_chunk);
/* Fraction of (wave, cluster) scorings the last pruned label sweep performed (-1 if the last
 * sweep was not pruned). */
int dpmm_last_label_evaluations(dpmm_ctx* ctx, double* fraction);

/* Verbose diagnostics of dp_parallel (dp_parallel_sampling.cpp:164-180: the NMI of the labels against
 * the ground truth, which the reference computes on the host with MIToolbox after copying all labels
 * back): the ground-truth class of every point is set once; dpmm_label_contingency returns the K x C
 * table of (label, class) counts (row-major, labels outside [0, K) left out; K x C <= 16384), built on
 * the device with LDS histograms and a fixed-order reduction; dpmm_label_nmi reduces it to
 * 2 I(X;Y) / (H(X) + H(Y)) (natural logarithms; 1 when both entropies are 0). */
int dpmm_set_ground_truth(dpmm_ctx* ctx, const int32_t* classes, int num_classes);
int dpmm_label_contingency(dpmm_ctx* ctx, int K, int64_t* table);
int dpmm_label_nmi(dpmm_ctx* ctx, int K, double* nmi);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The table of
 * dpmm_label_contingency from device arrays: labels and classes (N each, device memory), K x C
 * (<= 16384) int64 counts into d_table (device memory), launched on `stream` (a hipStream_t; NULL for
 * the default stream) without synchronising.  It calls no hipMalloc / hipFree: its partial tables
 * live in a 4 MiB array that the library holds on each GPU it runs on until the process ends.
 * Calls that overlap on one GPU share that array and race; calls in stream order do not. */
int dpmm_label_contingency_static(const int32_t* d_labels, const int32_t* d_classes, int64_t N, int K, int C,
                                  int64_t* d_table, void* stream);

/* Kernel timings of the last calls (ms, HIP events on the library's stream). */
int dpmm_last_kernel_ms(dpmm_ctx* ctx, float* label_sweep_ms, float* sub_sweep_ms, float* stats_ms);
/* Duration of the last pruned mixed label sweep's main kernel alone (the BF16 screen, with the fused
 * sub-label draw when on; ms, HIP events around its launch; 0 when the last sweep was not pruned or ran
 * inside a graph capture).  No counterpart in the reference. */
int dpmm_last_sweep_kernel_ms(dpmm_ctx* ctx, float* ms);
int dpmm_synchronize(dpmm_ctx* ctx);

/* ------------------------------------------------------------------ merge screener ----- */
/* No counterpart in the reference.  A learned screener in front of check_and_merge's all-pairs
 * Hastings test (local_clusters_actions.cpp:287-321; should_merge, shared_actions.cpp:88-111), for
 * callers with many clusters: an MLP 16 -> 128 -> 128 -> 1 scores each candidate pair from 16
 * features of its two clusters, and only pairs scoring above a threshold get the exact test.  It
 * never accepts a merge by itself; the sampler runs the full all-pairs test every `full_scan_every`
 * iterations as fallback and counts the accepted merges the screener would have missed. */
typedef struct dpmm_screener dpmm_screener;
/* Features (16 floats per pair) of cluster pairs (pi[p], pj[p]) from the current statistics of K
 * clusters and per-cluster scalars (K x 7: N, N_left, N_right, log ML of cluster / left / right,
 * log det psi_n of the cluster): log N_i, log N_j, log(N_i + N_j), L_i / (N_i D), L_j / (N_j D),
 * logdet_i / D, logdet_j / D, log1p of the mean distance whitened by diag psi_i, by diag psi_j and
 * by their sum (each / D), the split gains (L_l + L_r - L) / (N D) of i and j, the sub-cluster count
 * imbalances |log(N_l + 1) - log(N_r + 1)| of i and j, |logdet_i - logdet_j| / D, |log N_i - log N_j|.
 * Every feature is per point and per dimension, so a screener trained at one D applies at another. */
int dpmm_pair_features(dpmm_ctx* ctx, int K, const double* cluster_scalars, int npairs, const int32_t* pi,
                       const int32_t* pj, float* features);
int dpmm_screener_create(uint64_t seed, dpmm_screener** out);
int dpmm_screener_destroy(dpmm_screener* s);
/* Number of parameters: [W1 (128 x 16) | b1 (128) | W2 (128 x 128) | b2 (128) | w3 (128) | b3]. */
int dpmm_screener_param_count(void);
/* Parameters (FP32) and, when both are given, the per-feature standardisation (x - mean) / scale. */
int dpmm_screener_set_params(dpmm_screener* s, const float* params, const float* mean, const float* scale);
int dpmm_screener_get_params(dpmm_screener* s, float* params, float* mean, float* scale);
/* Mean binary cross-entropy with logits over n rows (rows with y = 1 weighted pos_weight) and its
 * gradient, FP32 forward and backward kernels. */
int dpmm_screener_loss_grad(dpmm_screener* s, int n, const float* X, const float* y, float pos_weight, float* loss,
                            float* grad);
/* One Adam step (Kingma & Ba) on a batch; loss of the batch before the step. */
int dpmm_screener_train_step(dpmm_screener* s, int n, const float* X, const float* y, float pos_weight, float lr,
                             float* loss);
/* Logits of n rows, BF16 MFMA with FP32 accumulation. */
int dpmm_screener_forward(dpmm_screener* s, int n, const float* X, float* logits);

/* ------------------------------------------------------------------ host driver ----- */
/* dp_parallel() options (dp_parallel_sampling.h:17-29, global_params.h:36-78). */
typedef struct {
    int prior;
    double alpha;
    int iters;
    int init_clusters;
    int burnout;
    int argmax_sample_stop;   /* reference default 5 */
    int split_stop;           /* reference default 5 */
    double max_clusters;
    uint64_t seed;
    int full_mu_cov;          /* 0: the reference's diagonal-only mu draw (niw.cpp:100-103) */
    int mult_use_weight;      /* 0: the reference's multinomial label log-likelihood */
    int precision;            /* DPMM_PREC_* of the label sweep */
    int verbose;
    /* != 0: start from the labels already in the context (dpmm_set_labels, 0 .. init_clusters - 1)
     * instead of uniform ones, as the reference's model-file constructor continues a saved model
     * (dp_parallel_sampling.h:15); sub-labels are drawn uniformly. */
    int init_from_labels;
    /* != 0: hard assignments, every label sweep takes the argmax (global_params.h:38,
     * local_clusters_actions.cpp:466: sample_labels(group, hard_clustering ? true : final, ...));
     * split / merge decisions still use the final flag of the iteration. */
    int hard_clustering;
    /* > 0: the reference's outlier mode (dp_parallel's outlier_weight, dp_parallel_sampling.h:28): cluster 0
     * is the outlier cluster, with the prior set by dpmm_set_outlier_prior and mixture weight outlier_mod, and
     * the reference's index rules hold (local_clusters_actions.cpp:250, 295, 348, 364-370, 409-410). */
    double outlier_mod;
    /* 0 (default): each iteration runs on the device from the parameter draw to the relabel passes (the
     * decisions of reset_bad_clusters, check_and_split, check_and_merge and remove_empty_clusters,
     * local_clusters_actions.cpp:244-321, 400-447, in one device kernel) with one synchronisation, in
     * graph mode as one hipGraph; runs with an outlier cluster or a merge screener keep their decisions on
     * the host.  != 0: decisions on the host between device launches.  Both take the same decisions. */
    int host_decisions;
} dpmm_options;

/* init_model + init_first_clusters (dp_parallel_sampling.cpp:91-130).  Gaussian hyper:
 * niw_hyperparams(k0, m0, v0, psi0) with psi0 = Psi / v0 (priors/niw_hyperparams.h);
 * multinomial: multinomial_hyper(alpha0) (priors/multinomial_hyper.h). */
int dpmm_sampler_create(dpmm_ctx* ctx, const dpmm_options* opt, double k0, const double* m0, double v0,
                        const double* psi0, const double* alpha0, dpmm_sampler** out);
/* One iteration, local_clusters_actions::group_step (local_clusters_actions.cpp:460-490),
 * with the final / no_more_splits flags of run_model (dp_parallel_sampling.cpp:182-193). */
int dpmm_sampler_step(dpmm_sampler* s, int iter);
int dpmm_sampler_num_clusters(dpmm_sampler* s);
/* Mixture weights of the last step (K values) and per-cluster counts. */
int dpmm_sampler_weights(dpmm_sampler* s, double* weights, int64_t* counts, int max_k);
/* Test hook: from the next step on, copy the labels and sub-labels the step's two sweeps drew
 * (before resets, splits and merges relabel them) into the caller's host arrays of N entries
 * (either may be NULL; both NULL turns the copies off).  No counterpart in the reference; it lets a
 * test follow the sampler iteration by iteration against an FP64 restatement. */
int dpmm_sampler_trace_sweep(dpmm_sampler* s, int32_t* labels, int32_t* sub_labels);
/* Host time spent in the last step's FP64 work (draws, posteriors, Hastings), ms. */
int dpmm_sampler_host_ms(dpmm_sampler* s, double* host_ms);
/* dp_parallel_sampling_class::calculate_posterior (dp_parallel_sampling.cpp:218-236) after the last
 * step: lgamma(alpha) - lgamma(N + alpha) + sum over clusters with points of
 * log ML_k + log alpha + lgamma(N_k), the entry of ModelInfo::likelihood_history in verbose mode. */
int dpmm_sampler_log_posterior(dpmm_sampler* s, double* log_posterior);
/* Decisions since creation: accepted splits, merges, reset clusters (reset_bad_clusters), and clusters
 * removed (merged away or emptied, remove_empty_clusters); the out_events of the decision step
 * (local_clusters_actions.cpp:244-321, 400-447).  Any pointer may be NULL. */
int dpmm_sampler_events(dpmm_sampler* s, int64_t* splits, int64_t* merges, int64_t* resets, int64_t* removed);
int dpmm_sampler_destroy(dpmm_sampler* s);

/* Merge screening for a Gaussian sampler (s not owned): from then on, iterations t with
 * t % full_scan_every != 0 give the exact merge test only to candidate pairs whose logit exceeds
 * `threshold`; the others run it on all pairs and count accepted merges scored below it.  s = NULL
 * turns screening off.  log_pairs != 0 keeps every candidate pair's features, log Hastings ratio
 * and log U of the last step (screening off: all pairs), the training data of the screener. */
int dpmm_sampler_set_screener(dpmm_sampler* sampler, dpmm_screener* s, int full_scan_every, float threshold,
                              int log_pairs);
/* Totals since creation: exact pair tests run, pairs the screener skipped, accepted merges the
 * screener scored below threshold on full-scan iterations, full scans. */
int dpmm_sampler_screen_stats(dpmm_sampler* sampler, int64_t* exact_tests, int64_t* skipped, int64_t* missed,
                              int64_t* full_scans);
/* Merge bound of the device-resident iteration (NIW prior; on by default; no counterpart in the
 * reference, which computes the log marginal likelihood of every candidate pair's union,
 * local_clusters_actions.cpp:287-321).  Merging i and j gives Psi_m = Psi_i + S_j + c r r^T exactly
 * (r = mu_i - xbar_j, c = kappa_i N_j / (kappa_i + N_j)), so log det Psi_m >= log det Psi_i +
 * log(1 + c |r|^2 / tr Psi_i), and the larger of that bound from i and from j bounds the merged log ML,
 * and the Hastings ratio, from above; a pair whose bounded ratio is below -100 is rejected without its
 * log ML, as its exact test would decide whatever its uniform.  Decisions are identical with the bound
 * on or off.  dpmm_sampler_merge_bound_rejects: pairs rejected so since creation (exact tests: the
 * other candidates, counted by dpmm_sampler_screen_stats). */
int dpmm_sampler_set_merge_bound(dpmm_sampler* sampler, int enable);
int dpmm_sampler_merge_bound_rejects(dpmm_sampler* sampler, int64_t* rejected);
/* The last step's pair log: up to max_rows rows of features (16 floats), log H and log U. */
int dpmm_sampler_pair_log(dpmm_sampler* sampler, int max_rows, float* features, double* log_h, double* log_u,
                          int* n_rows);

/* dp_parallel_sampling_class::dp_parallel (dp_parallel_sampling.cpp:132-160) end to end:
 * labels (N, 0-based), cluster count history (iters), per-iteration seconds (iters). */
int dpmm_run(dpmm_ctx* ctx, const dpmm_options* opt, double k0, const double* m0, double v0, const double* psi0,
             const double* alpha0, int32_t* labels, int32_t* k_hist, double* iter_seconds, double* weights,
             int max_k, int* final_k);

#ifdef __cplusplus
}
#endif
#endif
