/*
 * FP64 CPU oracle for the DPMM sub-cluster split/merge sampler (Chang & Fisher).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product library
 * (dpmmsubclusters_gpu_amd) never links, loads or calls it.
 *
 * Every function restates the reference's algorithm (BGU-CS-VIL/DPMMSubClusters_GPU)
 * in plain FP64 C++ and cites the reference file:line it follows.  Random numbers come
 * from the counter-based Philox4x32-10 schedule documented in dpmm_oracle.cpp, which the
 * product's HIP kernels consume identically, so labels can be compared point by point.
 *
 * Conventions: points are row-major N x D (each point contiguous, the reference's
 * D x N column-major layout); labels are 0-based, sub-labels are 0 (left, the
 * reference's 1) or 1 (right, the reference's 2).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10, Salmon et al. SC'11. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* First `n` draws of a sequential stream (c0, c2, c3) as uniforms / normals / gammas. */
void orc_stream_uniforms(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, double* out);
void orc_stream_normals(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, double* out);
void orc_stream_gammas(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, double shape, int n, double* out);
void orc_stream_dirichlet(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, const double* alpha, double* out);
/* Dirichlet draw of the multinomial parameter draws: coordinate d's gamma from the stream
 * (c0, n, c2, c3) starting at n = 1024 d. */
void orc_dirichlet_substreams(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, const double* alpha,
                              double* out);

double orc_lgamma(double x);
double orc_log_multivariate_gamma(double x, int D);
double orc_logdet_spd(int D, const double* A);

/* Per-point log-likelihoods, reference cudaKernel_gaussian.cu:82-113 and
 * cudaKernel_multinomial.cu:60-76.  weight <= 0 with use_weight=1 gives -inf. */
void orc_gauss_loglik(const double* X, int64_t N, int D, const double* mu, const double* inv_sigma,
                      double logdet_sigma, double weight, int use_weight, double* out);
void orc_mult_loglik(const double* X, int64_t N, int D, const double* logtheta, double* out);

/* NIW / Dirichlet-multinomial conjugate math (priors/niw.cpp, priors/multinomial_prior.cpp). */
void orc_niw_posterior(int D, double k0, const double* m0, double v0, const double* psi0, int64_t N,
                       const double* sum, const double* S, double* k, double* m, double* v, double* psi);
double orc_niw_log_marginal(int D, double k0, double v0, const double* psi0, double kn, double vn,
                            const double* psin, int64_t N);
double orc_mult_log_marginal(int D, const double* alpha0, const double* alphan);
double orc_split_log_hr(double alpha, int64_t N, int64_t Nl, int64_t Nr, double L, double Ll, double Lr);
double orc_merge_log_hr(double alpha, int64_t Ni, int64_t Nj, double L, double Li, double Lj);

/* NIW parameter draw (niw.cpp:82-107).  Outputs mu (D), inv_sigma (DxD), logdet_sigma. */
void orc_niw_draw(int D, double k, const double* m, double v, const double* psi, uint64_t seed,
                  uint32_t slot, uint32_t iter, int full_mu_cov, double* mu, double* inv_sigma,
                  double* logdet_sigma);

/* Label sweep (cudaKernel.cu:1400-1490): K clusters, Gumbel-max categorical draw on
 * Philox noise, or argmax when final != 0.  prior 0 = Gaussian, 1 = multinomial.
 * Gaussian params: mu K*D, inv_sigma K*D*D, logdet K.  Multinomial: logtheta K*D.
 * weights K (ignored for multinomial unless mult_use_weight).  margin_out (optional)
 * gets (top1 - top2) / |top1| of the perturbed scores. */
void orc_label_sweep(int prior, const double* X, int64_t N, int D, int K, const double* mu,
                     const double* inv_sigma, const double* logdet, const double* logtheta,
                     const double* weights, int mult_use_weight, uint64_t seed, uint32_t iter,
                     int final_iter, int32_t* labels, double* margin_out, double* score_out);

/* Sub-label sweep (cudaKernel.cu:1299-1398): params for 2K sub-clusters, index 2k+s.
 * lr_weights K*2.  margin_out gets |(s0 - s1) - logit(u)| / max(|s0|,|s1|). */
void orc_sublabel_sweep(int prior, const double* X, int64_t N, int D, int K, const int32_t* labels,
                        const double* mu, const double* inv_sigma, const double* logdet,
                        const double* logtheta, const double* lr_weights, uint64_t seed,
                        uint32_t iter, int32_t* sub_labels, double* margin_out);

/* Sufficient statistics (cudaKernel.cu:844-941 + cudaKernel_gaussian.cu:126-167):
 * per cluster k and sub s, Nsub[k*2+s], sum[(k*2+s)*D + d], S[(k*2+s)*D*D] (Gaussian only,
 * S may be NULL).  Also cluster level: Nc[k], sumc[k*D], Sc[k*D*D] computed in their own pass. */
void orc_suff_stats(int prior, const double* X, int64_t N, int D, int K, const int32_t* labels,
                    const int32_t* sub_labels, int64_t* Nsub, double* sum, double* S, int64_t* Nc,
                    double* sumc, double* Sc);

double orc_nmi(const int32_t* a, const int32_t* b, int64_t N);

/* Full sampler, dp_parallel_sampling.cpp:132-249 + local_clusters_actions.cpp:460-490.
 * Gaussian hyper: k0, m0 (D), v0, psi0 (DxD).  Multinomial hyper: alpha0 (D).
 * Outputs: labels (N, 0-based), k_hist (iters), weights_out (max_k_out), returns final K
 * (negative on error).  iter_seconds (optional, iters). */
typedef struct {
    int prior;            /* 0 Gaussian, 1 multinomial */
    double alpha;
    int iters;
    int init_clusters;
    int burnout;
    int argmax_sample_stop;
    int split_stop;
    double max_clusters;
    uint64_t seed;
    int full_mu_cov;      /* 0 reproduces the reference's diagonal-only mu draw */
    int mult_use_weight;  /* 0 reproduces the reference's multinomial label LL without log pi */
    int verbose;
    double outlier_mod;   /* > 0: outlier mode (orc_model_create_outlier); orc_run ignores it */
} orc_run_opts;

int orc_run(const orc_run_opts* opts, const double* X, int64_t N, int D, double k0, const double* m0,
            double v0, const double* psi0, const double* alpha0, int32_t* labels, int32_t* k_hist,
            double* weights_out, int max_k_out, double* iter_seconds);

/* The same sampler one iteration at a time: orc_model_create runs init_model + init_first_clusters
 * (dp_parallel_sampling.cpp:48-87) on a copy of X, orc_model_step(m, t) one run_model iteration t
 * (139-154, group_step local_clusters_actions.cpp:460-490) and returns the cluster count.  With
 * orc_model_trace on, each step keeps the labels and sub-labels its two sweeps drew (before resets,
 * splits and merges) and their relative margins, read by orc_model_sweep_trace. */
typedef struct orc_model orc_model;
orc_model* orc_model_create(const orc_run_opts* opts, const double* X, int64_t N, int D, double k0, const double* m0,
                            double v0, const double* psi0, const double* alpha0);
/* The same with the reference's outlier cluster (opts->outlier_mod > 0; dp_parallel_sampling.cpp:68-72):
 * cluster 0 with the prior (ok0, om0, ov0, opsi0) or oalpha0, mixture weight outlier_mod, and the index
 * rules of local_clusters_actions.cpp:250, 295, 348, 364-370, 409-410. */
orc_model* orc_model_create_outlier(const orc_run_opts* opts, const double* X, int64_t N, int D, double k0,
                                    const double* m0, double v0, const double* psi0, const double* alpha0, double ok0,
                                    const double* om0, double ov0, const double* opsi0, const double* oalpha0);
void orc_model_destroy(orc_model* m);
void orc_model_trace(orc_model* m, int enable);
int orc_model_step(orc_model* m, int t);
int orc_model_num_clusters(const orc_model* m);
/* labels / sub-labels (N), per-cluster counts and mixture weights (K); any may be NULL. */
void orc_model_state(const orc_model* m, int32_t* labels, int32_t* subs, int64_t* counts, double* weights);
int orc_model_sweep_trace(const orc_model* m, int32_t* labels, int32_t* subs, double* margin, double* smargin);

/* Merge-pair screener (mlp_oracle.cpp; no reference counterpart): MLP 16 -> 128 -> 128 -> 1 with
 * parameters [W1 (128x16) | b1 | W2 (128x128) | b2 | w3 | b3], ReLU hidden layers, logit output;
 * mean weighted binary cross-entropy with logits and its gradient; pair features from cluster
 * summaries (counts, log marginal likelihoods, log det psi_n, posterior means, diag psi_n). */
int orc_mlp_param_count(void);
void orc_mlp_forward(const double* params, const double* X, int n, double* z);
double orc_mlp_loss_grad(const double* params, const double* X, const double* y, int n, double pos_weight,
                         double* grad);
void orc_pair_features(int K, int D, const double* N, const double* Nl, const double* Nr, const double* L,
                       const double* Ll, const double* Lr, const double* ld, const double* m, const double* dpsi,
                       int npairs, const int32_t* pi, const int32_t* pj, double* out);

#ifdef __cplusplus
}
#endif
