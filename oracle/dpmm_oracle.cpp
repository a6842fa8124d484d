/*
 * FP64 CPU oracle for the DPMM sub-cluster split/merge sampler.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product (dpmmsubclusters_gpu_amd) never uses it.
 *
 * It restates, in plain FP64 C++ with OpenMP over points, the per-iteration step of
 * BGU-CS-VIL/DPMMSubClusters_GPU.  Each function names the reference file:line it follows.
 * The reference cannot be compiled here (CUDA + cuBLAS + cuRAND), so the restatement is
 * pinned by the reference's own golden vectors (tests/test_oracle_golden.py) plus closed
 * forms and brute force.
 *
 * Random numbers.  The reference draws from cuRAND XORWOW, mt19937 and std::random_device;
 * none of that is reproducible across implementations.  Oracle and product both use the
 * counter-based Philox4x32-10 generator (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11) keyed by the 64-bit seed, with this counter schedule
 * (c0, c1, c2, c3):
 *   init labels       (i, 0, 0, 1)      word0 -> label floor(u*K0), word1 -> sub (w & 1)
 *   label noise       (i, k/4, t, 2)    word k%4 -> Gumbel noise of cluster k
 *   sub-label draw    (i, 0, t, 3)      word0 -> logistic threshold
 *   reset relabel     (i, 0, t, 4)      word0 & 1 -> sub-label
 *   split relabel     (i, 0, t, 5)      word0 & 1 -> sub-label
 *   sequential streams (c0, n, t, tag) with n = 0, 1, 2, ... (4 words per n):
 *     tag 6 parameter draw of slot 3*cluster + {0 cluster, 1 left, 2 right}; a Dirichlet
 *           (multinomial) draw gives coordinate d the stream starting at n = 1024 d, and the
 *           NIW draw's D Bartlett chi variates take the streams starting at n = 1024 (16 + i),
 *           so the D gammas are independent of each other's rejection counts (the NIW draw's
 *           normals, D (D - 1) / 2 of the Bartlett factor then D of mu, stay on n = 0, 1, ...)
 *     tag 7 left/right weights of a cluster, tag 8 mixture weights (c0 = component index, one
 *           gamma per stream),
 *     tag 9 split-test uniform (c0 = cluster), tag 10 merge-test uniform (c0 = i<<16 | j).
 * uniform(w) = (w + 0.5) 2^-32; normal = Box-Muller cos branch of two uniforms;
 * Gamma(a) = Marsaglia-Tsang (ACM TOMS 26(3), 2000), with Gamma(a+1) U^(1/a) for a < 1.
 * The categorical label draw is Gumbel-max over the K log-likelihoods (argmax of
 * ll_k - log(E_k), E_k = -log u_k), which has exactly the distribution of the reference's
 * softmax + Walker alias draw (cudaKernel.cu:27-124, 374-403).
 */
#include "dpmm_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kNegInf = -std::numeric_limits<double>::infinity();
constexpr double kPi = 3.14159265358979323846;

enum : uint32_t {
    TAG_INIT = 1, TAG_LABEL = 2, TAG_SUB = 3, TAG_RESET = 4, TAG_SPLITSUB = 5,
    TAG_PARAM = 6, TAG_LRW = 7, TAG_MIXW = 8, TAG_SPLIT_U = 9, TAG_MERGE_U = 10,
    TAG_SPLIT_LRW = 11, TAG_MERGE_LRW = 12   // sub-cluster weights set by a split / a merge
};

/* ---------------------------------------------------------------- Philox ---------- */
inline void philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

inline void key_of(uint64_t seed, uint32_t key[2]) {
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
}

inline void philox_at(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) {
    uint32_t key[2];
    key_of(seed, key);
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    philox(ctr, key, out);
}

inline double u01(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

/* E = -log(u) for u = (w + 0.5) 2^-32, evaluated without cancellation on both halves. */
inline double exp1_of(uint32_t w) {
    if (w < 0x80000000u) return -std::log(u01(w));
    return -std::log1p(-(((double)(0xFFFFFFFFu - w)) + 0.5) * 0x1p-32);
}
inline double gumbel_of(uint32_t w) { return -std::log(exp1_of(w)); }
/* logit(u) = log(u) - log(1 - u). */
inline double logit_of(uint32_t w) {
    return std::log((double)w + 0.5) - std::log((double)(0xFFFFFFFFu - w) + 0.5);
}

struct Stream {
    uint32_t key[2];
    uint32_t c0, c1, c2, c3;
    uint32_t buf[4];
    int pos;
    Stream(uint64_t seed, uint32_t a, uint32_t t, uint32_t tag, uint32_t first_block = 0)
        : c0(a), c1(first_block), c2(t), c3(tag), pos(4) {
        key_of(seed, key);
    }
    uint32_t next() {
        if (pos == 4) {
            const uint32_t ctr[4] = {c0, c1, c2, c3};
            philox(ctr, key, buf);
            ++c1;
            pos = 0;
        }
        return buf[pos++];
    }
    double uniform() { return u01(next()); }
    double normal() {
        const double a = uniform();
        const double b = uniform();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * kPi * b);
    }
    double gamma(double shape) {
        if (!(shape > 0.0)) return 0.0;
        if (shape < 1.0) {
            const double g = gamma(shape + 1.0);
            const double u = uniform();
            return g * std::pow(u, 1.0 / shape);
        }
        const double d = shape - 1.0 / 3.0;
        const double c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            const double z = normal();
            double v = 1.0 + c * z;
            if (v <= 0.0) continue;
            v = v * v * v;
            const double u = uniform();
            if (std::log(u) < 0.5 * z * z + d - d * v + d * std::log(v)) return d * v;
        }
    }
    /* dirichlet_distribution::operator() (distributions_util/dirichlet.h:38-48). */
    void dirichlet(int n, const double* alpha, double* out) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) { out[i] = gamma(alpha[i]); s += out[i]; }
        for (int i = 0; i < n; ++i) out[i] = out[i] / s;
    }
};

/* ------------------------------------------------------- small dense FP64 LA ------- */
/* Lower Cholesky in place (row-major DxD); upper part zeroed.  Returns false if not SPD. */
bool chol_lower(int D, std::vector<double>& A) {
    for (int j = 0; j < D; ++j) {
        double s = A[(size_t)j * D + j];
        for (int k = 0; k < j; ++k) s -= A[(size_t)j * D + k] * A[(size_t)j * D + k];
        if (!(s > 0.0)) return false;
        const double ljj = std::sqrt(s);
        A[(size_t)j * D + j] = ljj;
        for (int i = j + 1; i < D; ++i) {
            double t = A[(size_t)i * D + j];
            for (int k = 0; k < j; ++k) t -= A[(size_t)i * D + k] * A[(size_t)j * D + k];
            A[(size_t)i * D + j] = t / ljj;
        }
        for (int k = j + 1; k < D; ++k) A[(size_t)j * D + k] = 0.0;
    }
    return true;
}

/* logdet via Cholesky (distributions_util/logdet.h:4-27 with use_cholesky = true). */
double logdet_chol(int D, const double* A) {
    std::vector<double> L(A, A + (size_t)D * D);
    if (!chol_lower(D, L)) return std::numeric_limits<double>::quiet_NaN();
    double ld = 0.0;
    for (int i = 0; i < D; ++i) ld += std::log(L[(size_t)i * D + i]);
    return 2.0 * ld;
}

/* Inverse of an SPD matrix through its Cholesky factor (replaces the reference's
 * Gauss-Jordan kernels, cudaKernel.cu:1558-1598; same result in exact arithmetic). */
bool inv_spd(int D, const std::vector<double>& A, std::vector<double>& Ainv) {
    std::vector<double> L = A;
    if (!chol_lower(D, L)) return false;
    /* Linv = L^-1 (lower) */
    std::vector<double> Li((size_t)D * D, 0.0);
    for (int i = 0; i < D; ++i) {
        Li[(size_t)i * D + i] = 1.0 / L[(size_t)i * D + i];
        for (int j = 0; j < i; ++j) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s += L[(size_t)i * D + k] * Li[(size_t)k * D + j];
            Li[(size_t)i * D + j] = -s / L[(size_t)i * D + i];
        }
    }
    Ainv.assign((size_t)D * D, 0.0);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = i; k < D; ++k) s += Li[(size_t)k * D + i] * Li[(size_t)k * D + j];
            Ainv[(size_t)i * D + j] = s;
            Ainv[(size_t)j * D + i] = s;
        }
    return true;
}

/* ----------------------------------------------------------- model structures ------ */
struct Hyper {           /* niw_hyperparams (priors/niw_hyperparams.h:13-43) / multinomial_hyper */
    double k = 0, v = 0;
    std::vector<double> m, psi, alpha;
};
struct Stats {           /* niw_sufficient_statistics / multinomial_sufficient_statistics */
    int64_t N = 0;
    std::vector<double> sum, S;
};
struct Dist {            /* mv_gaussian (distributions/mv_gaussian.h) / multinomial_dist */
    std::vector<double> mu, lambda, logtheta;
    double logdet_sigma = 0;
};
struct ClusterParams {   /* cluster_parameters (ds.h:39-118) */
    Hyper prior, post;
    Stats st;
    Dist dist;
};
struct Cluster {         /* splittable_cluster_params (ds.h:122-212) */
    ClusterParams c, l, r;
    double lrw[2] = {0.5, 0.5};
    bool splittable = false;
    std::vector<std::pair<bool, double>> hist;
};

struct Model {
    int prior = 0;
    int D = 0;
    int64_t N = 0;
    const double* X = nullptr;
    Hyper hyper;
    double alpha = 1.0;
    uint64_t seed = 0;
    int burnout = 15;
    int full_mu_cov = 0;
    int mult_use_weight = 0;
    std::vector<Cluster> cl;
    std::vector<double> weights;
    std::vector<int32_t> labels, subs;
    /* outlier mode (dp_parallel's outlier_weight > 0): cluster 0 is the outlier cluster */
    double outlier_mod = 0.0;
    bool outlier() const { return outlier_mod > 0.0; }
    /* trace of the last iteration's sweeps (orc_model_trace): labels and sub-labels as the two
     * sweeps drew them, before resets, splits and merges, with the sweeps' relative margins */
    bool trace = false;
    std::vector<int32_t> tr_labels, tr_subs;
    std::vector<double> tr_margin, tr_smargin;
};

Stats empty_stats(int prior, int D) {
    Stats s;
    s.sum.assign(D, 0.0);
    if (prior == 0) s.S.assign((size_t)D * D, 0.0);
    return s;
}

/* niw::calc_posterior (priors/niw.cpp:23-49) and multinomial_prior::calc_posterior
 * (priors/multinomial_prior.cpp:12-22). */
Hyper posterior(int prior, int D, const Hyper& h, const Stats& s) {
    if (s.N == 0) return h;
    Hyper p;
    if (prior == 1) {
        p.alpha.resize(D);
        for (int d = 0; d < D; ++d) p.alpha[d] = h.alpha[d] + s.sum[d];
        return p;
    }
    p.k = h.k + (double)s.N;
    p.v = h.v + (double)s.N;
    p.m.resize(D);
    for (int d = 0; d < D; ++d) p.m[d] = (h.m[d] * h.k + s.sum[d]) / p.k;
    p.psi.assign((size_t)D * D, 0.0);
    for (int a = 0; a < D; ++a)
        for (int b = a; b < D; ++b) {
            const double t = h.v * h.psi[(size_t)a * D + b] + h.k * h.m[a] * h.m[b];
            const double val = (t - p.k * p.m[a] * p.m[b] + s.S[(size_t)a * D + b]) / p.v;
            p.psi[(size_t)a * D + b] = val;   /* selfadjointView<Upper> (niw.cpp:45) */
            p.psi[(size_t)b * D + a] = val;
        }
    return p;
}

double log_mvgamma(double x, long D) {   /* utils::log_multivariate_gamma (utils.cpp:94-102) */
    double res = D * (D - 1) / 4.0 * std::log(kPi);
    for (long d = 1; d <= D; ++d) res += std::lgamma(x + (1 - d) / 2.0);
    return res;
}

/* niw::log_marginal_likelihood (priors/niw.cpp:122-138). */
double niw_log_marginal(int D, const Hyper& h, const Hyper& p, int64_t N) {
    const double logpi = std::log(kPi);
    const double ld0 = logdet_chol(D, h.psi.data());
    const double ldn = logdet_chol(D, p.psi.data());
    return -(double)(N * (int64_t)D) * 0.5 * logpi + log_mvgamma(p.v / 2.0, D) - log_mvgamma(h.v / 2.0, D) +
           (h.v / 2.0) * (D * std::log(h.v) + ld0) - (p.v / 2.0) * (D * std::log(p.v) + ldn) +
           (D / 2.0) * std::log(h.k / p.k);
}

/* multinomial_prior::log_marginal_likelihood (priors/multinomial_prior.cpp:56-67). */
double mult_log_marginal(int D, const double* a0, const double* an) {
    double s0 = 0, sn = 0;
    for (int d = 0; d < D; ++d) { s0 += a0[d]; sn += an[d]; }
    double val = std::lgamma(s0) - std::lgamma(sn);
    for (int d = 0; d < D; ++d) val += std::lgamma(an[d]) - std::lgamma(a0[d]);
    return val;
}

double log_marginal(const Model& M, const Hyper& h, const Hyper& p, const Stats& s) {
    if (M.prior == 1) return mult_log_marginal(M.D, h.alpha.data(), p.alpha.data());
    return niw_log_marginal(M.D, h, p, s.N);
}

/* niw::sample_distribution (priors/niw.cpp:82-107) with inverseWishartNoInverse (61-75):
 * Lambda = (B A)(B A)^T, B = chol(inv(v psi)), A the Bartlett factor of rinvwish2
 * (stats rinvwish.ipp:35-58); Sigma = inv(Lambda); mu = m + diag(chol(Sigma / k)) o z, the
 * reference's r8vec_multinormal_sample reading only the diagonal (pdflib.cpp:4007-4014),
 * or m + chol(Sigma / k) z when full_mu_cov. */
bool niw_draw(int D, const Hyper& h, uint64_t seed, uint32_t slot, uint32_t t, int full_mu_cov, Dist& out) {
    Stream rs(seed, slot, t, TAG_PARAM);
    std::vector<double> Psi((size_t)D * D), PsiInv, B;
    for (size_t i = 0; i < Psi.size(); ++i) Psi[i] = h.v * h.psi[i];
    if (!inv_spd(D, Psi, PsiInv)) return false;
    B = PsiInv;
    if (!chol_lower(D, B)) return false;
    std::vector<double> A((size_t)D * D, 0.0);
    for (int i = 1; i < D; ++i)
        for (int j = 0; j < i; ++j) A[(size_t)i * D + j] = rs.normal();
    for (int i = 0; i < D; ++i) {
        Stream gs(seed, slot, t, TAG_PARAM, 1024u * (16u + (uint32_t)i));
        A[(size_t)i * D + i] = std::sqrt(2.0 * gs.gamma((h.v - i) / 2.0));
    }
    std::vector<double> T((size_t)D * D, 0.0);   /* T = B A (lower) */
    for (int i = 0; i < D; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = j; k <= i; ++k) s += B[(size_t)i * D + k] * A[(size_t)k * D + j];
            T[(size_t)i * D + j] = s;
        }
    out.lambda.assign((size_t)D * D, 0.0);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = 0; k <= j; ++k) s += T[(size_t)i * D + k] * T[(size_t)j * D + k];
            out.lambda[(size_t)i * D + j] = s;
            out.lambda[(size_t)j * D + i] = s;
        }
    std::vector<double> Sigma;
    if (!inv_spd(D, out.lambda, Sigma)) return false;
    std::vector<double> Lk(Sigma);
    for (auto& x : Lk) x /= h.k;
    if (!chol_lower(D, Lk)) return false;
    out.mu.resize(D);
    std::vector<double> z(D);
    for (int i = 0; i < D; ++i) z[i] = rs.normal();
    for (int i = 0; i < D; ++i) {
        double x = h.m[i];
        if (full_mu_cov) {
            for (int j = 0; j <= i; ++j) x += Lk[(size_t)i * D + j] * z[j];
        } else {
            x += Lk[(size_t)i * D + i] * z[i];
        }
        out.mu[i] = x;
    }
    out.logdet_sigma = logdet_chol(D, Sigma.data());
    return true;
}

/* multinomial_prior::sample_distribution (priors/multinomial_prior.cpp:24-49). */
/* Dirichlet(alpha) with coordinate d's gamma from the sub-stream starting at block 1024 d. */
void dirichlet_substreams(uint64_t seed, uint32_t c0, uint32_t t, uint32_t tag, int D, const double* alpha,
                          double* out) {
    double s = 0.0;
    for (int d = 0; d < D; ++d) {
        Stream rs(seed, c0, t, tag, (uint32_t)d * 1024u);
        out[d] = rs.gamma(alpha[d]);
        s += out[d];
    }
    for (int d = 0; d < D; ++d) out[d] /= s;
}

void mult_draw(int D, const Hyper& h, uint64_t seed, uint32_t slot, uint32_t t, Dist& out) {
    out.logtheta.resize(D);
    dirichlet_substreams(seed, slot, t, TAG_PARAM, D, h.alpha.data(), out.logtheta.data());
    for (int d = 0; d < D; ++d) out.logtheta[d] = std::log(out.logtheta[d]);
}

bool draw_dist(const Model& M, const Hyper& h, uint32_t slot, uint32_t t, Dist& out) {
    if (M.prior == 1) { mult_draw(M.D, h, M.seed, slot, t, out); return true; }
    return niw_draw(M.D, h, M.seed, slot, t, M.full_mu_cov, out);
}

/* ------------------------------------------------------------ log-likelihoods ------ */
inline double gauss_ll(int D, const double* x, const double* mu, const double* lam, double logdet) {
    double q = 0.0;
    double z[1024];
    for (int a = 0; a < D; ++a) z[a] = x[a] - mu[a];
    for (int a = 0; a < D; ++a) {
        double s = 0.0;
        const double* row = lam + (size_t)a * D;
        for (int b = 0; b < D; ++b) s += row[b] * z[b];
        q += z[a] * s;
    }
    /* sigma.size() = D*D in the constant (cudaKernel_gaussian.cu:75, 108). */
    return -(((double)D * D) * std::log(2.0 * kPi) + logdet) / 2.0 - q / 2.0;
}

inline double mult_ll(int D, const double* x, const double* lt) {
    double s = 0.0;
    for (int d = 0; d < D; ++d)
        if (x[d] != 0.0) s += lt[d] * x[d];
    return s;
}

/* ------------------------------------------------------------------ sweeps --------- */
void label_sweep(int prior, const double* X, int64_t N, int D, int K, const double* mu, const double* lam,
                 const double* logdet, const double* logtheta, const double* weights, int mult_use_weight,
                 uint64_t seed, uint32_t iter, int final_iter, int32_t* labels, double* margin,
                 double* score_out) {
    std::vector<double> logw(K);
    for (int k = 0; k < K; ++k) logw[k] = std::log(weights[k]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const double* x = X + (size_t)i * D;
        double best = kNegInf, second = kNegInf;
        int bi = 0;
        uint32_t w[4] = {0, 0, 0, 0};
        for (int k = 0; k < K; ++k) {
            if ((k & 3) == 0 && !final_iter) philox_at(seed, (uint32_t)i, (uint32_t)(k >> 2), iter, TAG_LABEL, w);
            double ll;
            if (prior == 0) {
                ll = gauss_ll(D, x, mu + (size_t)k * D, lam + (size_t)k * D * D, logdet[k]) + logw[k];
            } else {
                ll = mult_ll(D, x, logtheta + (size_t)k * D) + (mult_use_weight ? logw[k] : 0.0);
            }
            const double s = final_iter ? ll : ll + gumbel_of(w[k & 3]);
            if (score_out) score_out[(size_t)i * K + k] = s;
            if (s > best || k == 0) { second = best; best = s; bi = k; }
            else if (s > second) second = s;
        }
        labels[i] = bi;
        if (margin) margin[i] = (K > 1) ? (best - second) / std::fabs(best) : std::numeric_limits<double>::infinity();
    }
}

void sublabel_sweep(int prior, const double* X, int64_t N, int D, int K, const int32_t* labels, const double* mu,
                    const double* lam, const double* logdet, const double* logtheta, const double* lrw,
                    uint64_t seed, uint32_t iter, int32_t* subs, double* margin) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const double* x = X + (size_t)i * D;
        const int k = labels[i];
        double s[2];
        for (int j = 0; j < 2; ++j) {
            const int q = 2 * k + j;
            double ll;
            if (prior == 0) ll = gauss_ll(D, x, mu + (size_t)q * D, lam + (size_t)q * D * D, logdet[q]);
            else ll = mult_ll(D, x, logtheta + (size_t)q * D);
            s[j] = ll + std::log(lrw[2 * k + j]);   /* cudaKernel.cu:352-372 */
        }
        uint32_t w[4];
        philox_at(seed, (uint32_t)i, 0, iter, TAG_SUB, w);
        const double lg = logit_of(w[0]);
        const double diff = s[0] - s[1];
        subs[i] = (lg < diff) ? 0 : 1;
        if (margin) margin[i] = std::fabs(diff - lg) / std::max(std::fabs(s[0]), std::fabs(s[1]));
    }
}

/* Stable bucket lists of point indices by key (deterministic order). */
void bucket(int64_t N, int nb, const int32_t* key, std::vector<int64_t>& start, std::vector<int64_t>& idx) {
    start.assign(nb + 1, 0);
    for (int64_t i = 0; i < N; ++i) start[key[i] + 1]++;
    for (int b = 0; b < nb; ++b) start[b + 1] += start[b];
    idx.resize(N);
    std::vector<int64_t> cur(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < N; ++i) idx[cur[key[i]]++] = i;
}

/* Stats over bucketed lists: N, sum, S (upper triangle mirrored), summed in index order. */
void stats_of_buckets(const double* X, int D, int nb, int want_S, const std::vector<int64_t>& start,
                      const std::vector<int64_t>& idx, int64_t* Nout, double* sum, double* S) {
    for (int b = 0; b < nb; ++b) Nout[b] = start[b + 1] - start[b];
#pragma omp parallel for schedule(dynamic)
    for (int64_t task = 0; task < (int64_t)nb * D; ++task) {
        const int b = (int)(task / D), a = (int)(task % D);
        double sa = 0.0;
        std::vector<double> row(want_S ? D : 0, 0.0);
        for (int64_t p = start[b]; p < start[b + 1]; ++p) {
            const double* x = X + (size_t)idx[p] * D;
            sa += x[a];
            if (want_S)
                for (int c = a; c < D; ++c) row[c] += x[a] * x[c];
        }
        sum[(size_t)b * D + a] = sa;
        if (want_S)
            for (int c = a; c < D; ++c) {
                S[((size_t)b * D + a) * D + c] = row[c];
                S[((size_t)b * D + c) * D + a] = row[c];
            }
    }
}

/* create_sufficient_statistics for all clusters (cudaKernel.cu:844-941): sub-cluster
 * statistics from the points of each (label, sub), cluster statistics in their own pass. */
void suff_stats(int prior, const double* X, int64_t N, int D, int K, const int32_t* labels, const int32_t* subs,
                int64_t* Nsub, double* sum, double* S, int64_t* Nc, double* sumc, double* Sc) {
    const int want_S = (prior == 0);
    std::vector<int32_t> key(N);
    for (int64_t i = 0; i < N; ++i) key[i] = labels[i] * 2 + subs[i];
    std::vector<int64_t> start, idx;
    bucket(N, 2 * K, key.data(), start, idx);
    stats_of_buckets(X, D, 2 * K, want_S && S, start, idx, Nsub, sum, S);
    if (Nc) {
        bucket(N, K, labels, start, idx);
        stats_of_buckets(X, D, K, want_S && Sc, start, idx, Nc, sumc, Sc);
    }
}

/* ------------------------------------------------------------- sampler step -------- */
void reset_hist(Model& M, Cluster& c) {   /* Logsublikelihood_hist(burnout + 5, (false, 0)) */
    c.hist.assign(M.burnout + 5, std::make_pair(false, 0.0));
}

/* update_suff_stats_posterior (local_clusters_actions.cpp:109-168) for clusters `which`. */
void update_stats(Model& M, const std::vector<int>& which) {
    const int K = (int)M.cl.size(), D = M.D;
    std::vector<int64_t> Nsub(2 * K), Nc(K);
    std::vector<double> sum((size_t)2 * K * D), sumc((size_t)K * D);
    std::vector<double> S, Sc;
    if (M.prior == 0) { S.resize((size_t)2 * K * D * D); Sc.resize((size_t)K * D * D); }
    suff_stats(M.prior, M.X, M.N, D, K, M.labels.data(), M.subs.data(), Nsub.data(), sum.data(),
               M.prior == 0 ? S.data() : nullptr, Nc.data(), sumc.data(), M.prior == 0 ? Sc.data() : nullptr);
    auto fill = [&](Stats& st, int64_t n, const double* su, const double* s2) {
        st.N = n;
        st.sum.assign(su, su + D);
        if (M.prior == 0) st.S.assign(s2, s2 + (size_t)D * D);
    };
    for (int k : which) {
        Cluster& c = M.cl[k];
        fill(c.c.st, Nc[k], &sumc[(size_t)k * D], M.prior == 0 ? &Sc[(size_t)k * D * D] : nullptr);
        fill(c.l.st, Nsub[2 * k], &sum[(size_t)(2 * k) * D], M.prior == 0 ? &S[(size_t)(2 * k) * D * D] : nullptr);
        fill(c.r.st, Nsub[2 * k + 1], &sum[(size_t)(2 * k + 1) * D],
             M.prior == 0 ? &S[(size_t)(2 * k + 1) * D * D] : nullptr);
        c.c.post = posterior(M.prior, D, c.c.prior, c.c.st);
        c.l.post = posterior(M.prior, D, c.l.prior, c.l.st);
        c.r.post = posterior(M.prior, D, c.r.prior, c.r.st);
    }
}

void dirichlet_mix(Model& M, uint32_t t) {   /* sample_clusters tail (local_clusters_actions.cpp:357-370) */
    const int K = (int)M.cl.size();
    std::vector<double> counts;
    for (int k = 0; k < K; ++k)
        if (!(M.outlier() && k == 1)) counts.push_back((double)M.cl[k].c.st.N);   /* index 1 skipped (348) */
    counts.push_back(M.alpha);
    std::vector<double> out(counts.size());
    if (counts.size() == 2 && counts[0] == 0.0) {   /* get_dirichlet_distribution special case (323-339) */
        out[0] = 0; out[1] = 1;
    } else {   /* component i's gamma from stream (i, t, TAG_MIXW), the sum in index order */
        double s = 0.0;
        for (size_t i = 0; i < counts.size(); ++i) {
            Stream rs(M.seed, (uint32_t)i, t, TAG_MIXW);
            out[i] = rs.gamma(counts[i]);
            s += out[i];
        }
        for (size_t i = 0; i < counts.size(); ++i) out[i] = out[i] / s;
    }
    /* weights scaled by 1 - outlier_mod, the outlier cluster's weight in front (362-370) */
    M.weights.clear();
    if (M.outlier()) M.weights.push_back(M.outlier_mod);
    for (size_t i = 0; i + 1 < out.size(); ++i) M.weights.push_back(out[i] * (1.0 - M.outlier_mod));
}

/* sample_clusters + shared_actions::sample_cluster_params (shared_actions.cpp:9-55); in outlier mode
 * cluster index 1 keeps its parameters, sub-cluster weights and history (348-351). */
bool sample_clusters(Model& M, uint32_t t) {
    const int K = (int)M.cl.size();
    bool ok = true;
    for (int k = 0; k < K; ++k) {
        if (M.outlier() && k == 1) continue;
        Cluster& c = M.cl[k];
        ok &= draw_dist(M, c.c.post, 3 * k + 0, t, c.c.dist);
        ok &= draw_dist(M, c.l.post, 3 * k + 1, t, c.l.dist);
        ok &= draw_dist(M, c.r.post, 3 * k + 2, t, c.r.dist);
        const double a[2] = {(double)c.l.st.N + M.alpha / 2, (double)c.r.st.N + M.alpha / 2};
        Stream rs(M.seed, k, t, TAG_LRW);
        rs.dirichlet(2, a, c.lrw);
        const double Ll = log_marginal(M, c.l.prior, c.l.post, c.l.st);
        const double Lr = log_marginal(M, c.r.prior, c.r.post, c.r.st);
        const int B = M.burnout;
        for (int i = 0; i < B - 1; ++i) c.hist[i] = c.hist[i + 1];
        c.hist[B - 1] = std::make_pair(true, Ll + Lr);
        double now = 0.0;
        bool all = true;
        for (int i = 0; i < B; ++i) {
            if (c.hist[i].first) now += c.hist[i].second * (1 / (B - 0.1));
            else all = false;
        }
        if (all && now - c.hist[B - 1].second < 1e-2) c.splittable = true;
    }
    dirichlet_mix(M, t);
    return ok;
}

void gather_params(const Model& M, std::vector<double>& mu, std::vector<double>& lam, std::vector<double>& ld,
                   std::vector<double>& lt, int sub) {
    const int K = (int)M.cl.size(), D = M.D;
    const int n = sub ? 2 * K : K;
    mu.assign((size_t)n * D, 0.0); ld.assign(n, 0.0);
    if (M.prior == 0) lam.assign((size_t)n * D * D, 0.0); else lt.assign((size_t)n * D, 0.0);
    for (int q = 0; q < n; ++q) {
        const Cluster& c = M.cl[sub ? q / 2 : q];
        const Dist& d = sub ? ((q & 1) ? c.r.dist : c.l.dist) : c.c.dist;
        if (M.prior == 0) {
            std::copy(d.mu.begin(), d.mu.end(), mu.begin() + (size_t)q * D);
            std::copy(d.lambda.begin(), d.lambda.end(), lam.begin() + (size_t)q * D * D);
            ld[q] = d.logdet_sigma;
        } else {
            std::copy(d.logtheta.begin(), d.logtheta.end(), lt.begin() + (size_t)q * D);
        }
    }
}

void group_step(Model& M, uint32_t t, bool no_more_splits, bool final_iter) {
    const int D = M.D;
    sample_clusters(M, t);
    int K = (int)M.cl.size();
    std::vector<double> mu, lam, ld, lt;
    gather_params(M, mu, lam, ld, lt, 0);
    if (M.trace) { M.tr_margin.resize(M.N); M.tr_smargin.resize(M.N); }
    label_sweep(M.prior, M.X, M.N, D, K, mu.data(), lam.data(), ld.data(), lt.data(), M.weights.data(),
                M.mult_use_weight, M.seed, t, final_iter ? 1 : 0, M.labels.data(),
                M.trace ? M.tr_margin.data() : nullptr, nullptr);
    gather_params(M, mu, lam, ld, lt, 1);
    std::vector<double> lrw(2 * K);
    for (int k = 0; k < K; ++k) { lrw[2 * k] = M.cl[k].lrw[0]; lrw[2 * k + 1] = M.cl[k].lrw[1]; }
    sublabel_sweep(M.prior, M.X, M.N, D, K, M.labels.data(), mu.data(), lam.data(), ld.data(), lt.data(),
                   lrw.data(), M.seed, t, M.subs.data(), M.trace ? M.tr_smargin.data() : nullptr);
    if (M.trace) { M.tr_labels = M.labels; M.tr_subs = M.subs; }
    std::vector<int> all(K);
    for (int k = 0; k < K; ++k) all[k] = k;
    update_stats(M, all);

    /* reset_bad_clusters (local_clusters_actions.cpp:429-447; cudaKernel.cu:248-263) */
    std::vector<int> bad;
    std::vector<char> isbad(K, 0);
    for (int k = 0; k < K; ++k)
        if (M.cl[k].l.st.N == 0 || M.cl[k].r.st.N == 0) {
            bad.push_back(k); isbad[k] = 1;
            reset_hist(M, M.cl[k]);
            M.cl[k].splittable = false;
        }
    if (!bad.empty()) {
        for (int64_t i = 0; i < M.N; ++i)
            if (isbad[M.labels[i]]) {
                uint32_t w[4];
                philox_at(M.seed, (uint32_t)i, 0, t, TAG_RESET, w);
                M.subs[i] = (int32_t)(w[0] & 1u);
            }
        update_stats(M, bad);
    }

    if (!no_more_splits) {
        /* check_and_split (local_clusters_actions.cpp:244-285) + should_split_local (215-242) */
        std::vector<int> split_of(K, -1);
        int new_index = K;
        std::vector<int> touched;
        std::vector<int> accepted;
        for (int k = 0; k < K; ++k) {
            Cluster& c = M.cl[k];
            if (M.outlier() && k == 1) continue;   /* local_clusters_actions.cpp:250-253 */
            if (!(c.splittable && c.c.st.N > 1)) continue;
            if (final_iter || c.l.st.N == 0 || c.r.st.N == 0) continue;
            const Hyper post = posterior(M.prior, D, c.c.prior, c.c.st);
            const Hyper lpost = posterior(M.prior, D, c.c.prior, c.l.st);
            const Hyper rpost = posterior(M.prior, D, c.c.prior, c.r.st);
            const double Ll = log_marginal(M, c.l.prior, lpost, c.l.st);
            const double Lr = log_marginal(M, c.r.prior, rpost, c.r.st);
            const double L = log_marginal(M, c.c.prior, post, c.c.st);
            const double lhr = std::log(M.alpha) + std::lgamma((double)c.l.st.N) + Ll + std::lgamma((double)c.r.st.N) +
                               Lr - (std::lgamma((double)c.c.st.N) + L);
            Stream rs(M.seed, k, t, TAG_SPLIT_U);
            if (lhr > std::log(rs.uniform())) accepted.push_back(k);
        }
        for (int k : accepted) {
            /* split_cluster_local (183-194) + create_splittable_from_params (shared_actions.cpp:57-70) */
            Cluster old = M.cl[k];
            Cluster nl, nr;
            nl.c = old.l; nl.l = old.l; nl.r = old.l; reset_hist(M, nl); nl.splittable = false;
            nr.c = old.r; nr.l = old.r; nr.r = old.r; reset_hist(M, nr); nr.splittable = false;
            /* sub-cluster weights ~ Dir(alpha / 2, alpha / 2) (shared_actions.cpp:65-67) */
            const double half[2] = {M.alpha / 2, M.alpha / 2};
            Stream(M.seed, (uint32_t)k, t, TAG_SPLIT_LRW).dirichlet(2, half, nl.lrw);
            Stream(M.seed, (uint32_t)new_index, t, TAG_SPLIT_LRW).dirichlet(2, half, nr.lrw);
            M.cl[k] = nl;
            M.cl.push_back(nr);
            split_of[k] = new_index;
            touched.push_back(k);
            ++new_index;
        }
        for (int k : accepted) touched.push_back(split_of[k]);
        if (!accepted.empty()) {
            /* split_cluster_local_worker (cudaKernel.cu:202-221) */
            for (int64_t i = 0; i < M.N; ++i) {
                const int k = M.labels[i];
                if (k < K && split_of[k] >= 0) {
                    if (M.subs[i] == 1) M.labels[i] = split_of[k];
                    uint32_t w[4];
                    philox_at(M.seed, (uint32_t)i, 0, t, TAG_SPLITSUB, w);
                    M.subs[i] = (int32_t)(w[0] & 1u);
                }
            }
            update_stats(M, touched);
            /* fresh sub-cluster parameters of both halves, drawn from the half's posterior
             * (create_splittable_from_params, shared_actions.cpp:57-64), on the streams of the half's
             * sub-cluster slots; observable only in outlier mode, where cluster index 1 keeps them */
            if (M.outlier())
                for (int k : touched) {
                    Cluster& c = M.cl[k];
                    const Hyper post = posterior(M.prior, D, c.c.prior, c.c.st);
                    draw_dist(M, post, 3 * k + 1, t, c.l.dist);
                    draw_dist(M, post, 3 * k + 2, t, c.r.dist);
                }
        }

        /* check_and_merge (local_clusters_actions.cpp:287-321) + should_merge (shared_actions.cpp:88-111) */
        K = (int)M.cl.size();
        std::vector<std::pair<int, int>> merges;
        for (int i = M.outlier() ? 1 : 0; i < K; ++i)   /* the outlier cluster never merges (295-298) */
            for (int j = i + 1; j < K; ++j) {
                Cluster& a = M.cl[i];
                Cluster& b = M.cl[j];
                if (!(a.splittable && b.splittable && a.c.st.N > 0 && b.c.st.N > 0)) continue;
                Stats agg = a.c.st;
                agg.N += b.c.st.N;
                for (int d = 0; d < D; ++d) agg.sum[d] += b.c.st.sum[d];
                for (size_t e = 0; e < agg.S.size(); ++e) agg.S[e] += b.c.st.S[e];
                const Hyper p = posterior(M.prior, D, a.c.prior, agg);
                const double Li = log_marginal(M, a.c.prior, a.c.post, a.c.st);
                const double Lj = log_marginal(M, b.c.prior, b.c.post, b.c.st);
                const double L = log_marginal(M, a.c.prior, p, agg);
                const double al = M.alpha;
                const double Ni = (double)a.c.st.N, Nj = (double)b.c.st.N, Nm = (double)agg.N;
                const double lhr = (-std::log(al) + std::lgamma(al) - 2 * std::lgamma(0.5 * al) + std::lgamma(Nm) -
                                    std::lgamma(Nm + al) + std::lgamma(Ni + 0.5 * al) - std::lgamma(Ni) -
                                    std::lgamma(Nj) + std::lgamma(Nj + 0.5 * al) + L - Li - Lj);
                Stream rs(M.seed, ((uint32_t)i << 16) | (uint32_t)j, t, TAG_MERGE_U);
                const double u = rs.uniform();
                if ((lhr > std::log(u)) || (final_iter && lhr > std::log(0.1))) {
                    /* merge_clusters (204-213) + merge_clusters_to_splittable (shared_actions.cpp:72-86) */
                    /* sub-cluster weights ~ Dir(N_i + alpha / 2, N_j + alpha / 2) (shared_actions.cpp:77-81) */
                    const double mw[2] = {Ni + 0.5 * al, Nj + 0.5 * al};
                    Stream(M.seed, (uint32_t)i, t, TAG_MERGE_LRW).dirichlet(2, mw, a.lrw);
                    a.c.st = agg;
                    a.c.post = p;
                    a.l = a.c;   /* cluster_params_l = clone of the merged cluster (reference behaviour) */
                    a.r = b.c;
                    a.splittable = false;
                    reset_hist(M, a);
                    b.c.st.N = 0;
                    b.splittable = false;
                    merges.emplace_back(i, j);
                }
            }
        if (!merges.empty()) {
            std::vector<int> into(K, -1);
            std::vector<char> left(K, 0);
            for (auto& m : merges) { into[m.second] = m.first; left[m.first] = 1; }
            for (int64_t i = 0; i < M.N; ++i) {   /* merge_clusters_worker (cudaKernel.cu:223-246) */
                const int k = M.labels[i];
                if (left[k]) M.subs[i] = 0;
                if (into[k] >= 0) { M.subs[i] = 1; M.labels[i] = into[k]; }
            }
        }
    }

    /* remove_empty_clusters (local_clusters_actions.cpp:384-419; cudaKernel.cu:172-187) */
    K = (int)M.cl.size();
    std::vector<int> remap(K, -1);
    std::vector<Cluster> kept;
    for (int k = 0; k < K; ++k)   /* outlier mode keeps the outlier cluster, and cluster 1 while K = 2 (408-411) */
        if (M.cl[k].c.st.N > 0 || (M.outlier() && (k == 0 || (k == 1 && K == 2)))) {
            remap[k] = (int)kept.size();
            kept.push_back(M.cl[k]);
        }
    if ((int)kept.size() != K) {
        for (int64_t i = 0; i < M.N; ++i) M.labels[i] = remap[M.labels[i]];
        M.cl.swap(kept);
    }
}

}  // namespace

/* ================================================================== C ABI ========== */
extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

void orc_stream_uniforms(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, double* out) {
    Stream s(seed, c0, c2, c3);
    for (int i = 0; i < n; ++i) out[i] = s.uniform();
}
void orc_stream_normals(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, double* out) {
    Stream s(seed, c0, c2, c3);
    for (int i = 0; i < n; ++i) out[i] = s.normal();
}
void orc_stream_gammas(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, double shape, int n, double* out) {
    Stream s(seed, c0, c2, c3);
    for (int i = 0; i < n; ++i) out[i] = s.gamma(shape);
}
void orc_stream_dirichlet(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, const double* alpha,
                          double* out) {
    Stream s(seed, c0, c2, c3);
    s.dirichlet(n, alpha, out);
}

void orc_dirichlet_substreams(uint64_t seed, uint32_t c0, uint32_t c2, uint32_t c3, int n, const double* alpha,
                              double* out) {
    dirichlet_substreams(seed, c0, c2, c3, n, alpha, out);
}

double orc_lgamma(double x) { return std::lgamma(x); }
double orc_log_multivariate_gamma(double x, int D) { return log_mvgamma(x, D); }
double orc_logdet_spd(int D, const double* A) { return logdet_chol(D, A); }

void orc_gauss_loglik(const double* X, int64_t N, int D, const double* mu, const double* inv_sigma,
                      double logdet_sigma, double weight, int use_weight, double* out) {
    const double lw = use_weight ? std::log(weight) : 0.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) out[i] = gauss_ll(D, X + (size_t)i * D, mu, inv_sigma, logdet_sigma) + lw;
}

void orc_mult_loglik(const double* X, int64_t N, int D, const double* logtheta, double* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) out[i] = mult_ll(D, X + (size_t)i * D, logtheta);
}

void orc_niw_posterior(int D, double k0, const double* m0, double v0, const double* psi0, int64_t N,
                       const double* sum, const double* S, double* k, double* m, double* v, double* psi) {
    Hyper h;
    h.k = k0; h.v = v0; h.m.assign(m0, m0 + D); h.psi.assign(psi0, psi0 + (size_t)D * D);
    Stats s;
    s.N = N; s.sum.assign(sum, sum + D); s.S.assign(S, S + (size_t)D * D);
    const Hyper p = posterior(0, D, h, s);
    *k = p.k; *v = p.v;
    std::copy(p.m.begin(), p.m.end(), m);
    std::copy(p.psi.begin(), p.psi.end(), psi);
}

double orc_niw_log_marginal(int D, double k0, double v0, const double* psi0, double kn, double vn,
                            const double* psin, int64_t N) {
    Hyper h, p;
    h.k = k0; h.v = v0; h.psi.assign(psi0, psi0 + (size_t)D * D);
    p.k = kn; p.v = vn; p.psi.assign(psin, psin + (size_t)D * D);
    return niw_log_marginal(D, h, p, N);
}

double orc_mult_log_marginal(int D, const double* alpha0, const double* alphan) {
    return mult_log_marginal(D, alpha0, alphan);
}

double orc_split_log_hr(double alpha, int64_t N, int64_t Nl, int64_t Nr, double L, double Ll, double Lr) {
    return std::log(alpha) + std::lgamma((double)Nl) + Ll + std::lgamma((double)Nr) + Lr -
           (std::lgamma((double)N) + L);
}

double orc_merge_log_hr(double al, int64_t Ni_, int64_t Nj_, double L, double Li, double Lj) {
    const double Ni = (double)Ni_, Nj = (double)Nj_, Nm = Ni + Nj;
    return -std::log(al) + std::lgamma(al) - 2 * std::lgamma(0.5 * al) + std::lgamma(Nm) - std::lgamma(Nm + al) +
           std::lgamma(Ni + 0.5 * al) - std::lgamma(Ni) - std::lgamma(Nj) + std::lgamma(Nj + 0.5 * al) + L - Li - Lj;
}

void orc_niw_draw(int D, double k, const double* m, double v, const double* psi, uint64_t seed, uint32_t slot,
                  uint32_t iter, int full_mu_cov, double* mu, double* inv_sigma, double* logdet_sigma) {
    Hyper h;
    h.k = k; h.v = v; h.m.assign(m, m + D); h.psi.assign(psi, psi + (size_t)D * D);
    Dist d;
    if (!niw_draw(D, h, seed, slot, iter, full_mu_cov, d)) { *logdet_sigma = std::numeric_limits<double>::quiet_NaN(); return; }
    std::copy(d.mu.begin(), d.mu.end(), mu);
    std::copy(d.lambda.begin(), d.lambda.end(), inv_sigma);
    *logdet_sigma = d.logdet_sigma;
}

void orc_label_sweep(int prior, const double* X, int64_t N, int D, int K, const double* mu, const double* inv_sigma,
                     const double* logdet, const double* logtheta, const double* weights, int mult_use_weight,
                     uint64_t seed, uint32_t iter, int final_iter, int32_t* labels, double* margin_out,
                     double* score_out) {
    label_sweep(prior, X, N, D, K, mu, inv_sigma, logdet, logtheta, weights, mult_use_weight, seed, iter,
                final_iter, labels, margin_out, score_out);
}

void orc_sublabel_sweep(int prior, const double* X, int64_t N, int D, int K, const int32_t* labels,
                        const double* mu, const double* inv_sigma, const double* logdet, const double* logtheta,
                        const double* lr_weights, uint64_t seed, uint32_t iter, int32_t* sub_labels,
                        double* margin_out) {
    sublabel_sweep(prior, X, N, D, K, labels, mu, inv_sigma, logdet, logtheta, lr_weights, seed, iter, sub_labels,
                   margin_out);
}

void orc_suff_stats(int prior, const double* X, int64_t N, int D, int K, const int32_t* labels,
                    const int32_t* sub_labels, int64_t* Nsub, double* sum, double* S, int64_t* Nc, double* sumc,
                    double* Sc) {
    suff_stats(prior, X, N, D, K, labels, sub_labels, Nsub, sum, S, Nc, sumc, Sc);
}

/* NMI = 2 I(X;Y) / (H(X) + H(Y)) (dp_parallel_sampling.cpp:164-180 with MIToolbox). */
double orc_nmi(const int32_t* a, const int32_t* b, int64_t N) {
    int ma = 0, mb = 0;
    for (int64_t i = 0; i < N; ++i) { ma = std::max(ma, a[i] + 1); mb = std::max(mb, b[i] + 1); }
    std::vector<double> ca(ma, 0), cb(mb, 0), cab((size_t)ma * mb, 0);
    for (int64_t i = 0; i < N; ++i) { ca[a[i]]++; cb[b[i]]++; cab[(size_t)a[i] * mb + b[i]]++; }
    double ha = 0, hb = 0, I = 0;
    const double n = (double)N;
    for (double c : ca) if (c > 0) ha -= c / n * std::log(c / n);
    for (double c : cb) if (c > 0) hb -= c / n * std::log(c / n);
    for (int x = 0; x < ma; ++x)
        for (int y = 0; y < mb; ++y) {
            const double c = cab[(size_t)x * mb + y];
            if (c > 0) I += c / n * std::log(c * n / (ca[x] * cb[y]));
        }
    if (ha + hb <= 0) return 1.0;
    return 2 * I / (ha + hb);
}

}  // extern "C"

namespace {
/* init_model + init_first_clusters (dp_parallel_sampling.cpp:48-87) into M; X is not copied. */
void model_init(Model& M, const orc_run_opts* o, const double* X, int64_t N, int D, double k0, const double* m0,
                double v0, const double* psi0, const double* alpha0, const Hyper* outlier_hyper = nullptr) {
    M.prior = o->prior; M.D = D; M.N = N; M.X = X; M.alpha = o->alpha; M.seed = o->seed;
    M.outlier_mod = outlier_hyper ? o->outlier_mod : 0.0;
    M.burnout = o->burnout; M.full_mu_cov = o->full_mu_cov; M.mult_use_weight = o->mult_use_weight;
    if (M.prior == 0) {
        M.hyper.k = k0; M.hyper.v = v0; M.hyper.m.assign(m0, m0 + D); M.hyper.psi.assign(psi0, psi0 + (size_t)D * D);
    } else {
        M.hyper.alpha.assign(alpha0, alpha0 + D);
    }
    /* init_model (dp_parallel_sampling.cpp:48-62): uniform labels in 0..K0-1 and sub-labels */
    const int K0 = std::max(1, o->init_clusters);
    M.labels.resize(N); M.subs.resize(N);
    for (int64_t i = 0; i < N; ++i) {
        uint32_t w[4];
        philox_at(M.seed, (uint32_t)i, 0, 0, TAG_INIT, w);
        int l = (int)(u01(w[0]) * K0);
        /* with an outlier cluster the initial labels skip it (cudaKernel.cu:158-161) */
        M.labels[i] = std::min(l, K0 - 1) + (M.outlier() ? 1 : 0);
        M.subs[i] = (int32_t)(w[1] & 1u);
    }
    /* init_first_clusters (dp_parallel_sampling.cpp:65-87; create_first_local_cluster 14-46; the outlier
     * cluster first, with its own prior, create_outlier_local_cluster 48-80) */
    const int off = M.outlier() ? 1 : 0, K = K0 + off;
    M.cl.resize(K);
    for (int k = 0; k < K; ++k) {
        Cluster& c = M.cl[k];
        const Hyper& h = (k < off) ? *outlier_hyper : M.hyper;
        c.c.prior = h; c.c.post = h; c.c.st = empty_stats(M.prior, D);
        c.l = c.c; c.r = c.c;
        c.splittable = false;
        reset_hist(M, c);
    }
    if (M.outlier()) {
        /* cluster 1 keeps the parameters create_first_local_cluster drew from the prior (its sub-clusters
         * clones of them): sample_clusters never redraws index 1 in outlier mode */
        Cluster& c1 = M.cl[1];
        draw_dist(M, M.hyper, 3, 0, c1.c.dist);
        c1.l.dist = c1.c.dist;
        c1.r.dist = c1.c.dist;
    }
    std::vector<int> all(K);
    for (int k = 0; k < K; ++k) all[k] = k;
    update_stats(M, all);
    sample_clusters(M, 0);
}

/* One iteration of run_model (dp_parallel_sampling.cpp:139-154) with its final / no_more_splits flags. */
void model_step(Model& M, const orc_run_opts* o, int t) {
    const bool fin = t >= o->iters - o->argmax_sample_stop;
    const bool nms = t >= o->iters - o->split_stop || (double)M.cl.size() >= o->max_clusters;
    group_step(M, (uint32_t)t, nms, fin);
}
}  // namespace

extern "C" {

int orc_run(const orc_run_opts* o, const double* X, int64_t N, int D, double k0, const double* m0, double v0,
            const double* psi0, const double* alpha0, int32_t* labels, int32_t* k_hist, double* weights_out,
            int max_k_out, double* iter_seconds) {
    Model M;
    model_init(M, o, X, N, D, k0, m0, v0, psi0, alpha0);
    for (int t = 1; t <= o->iters; ++t) {
        const auto t0 = std::chrono::steady_clock::now();
        model_step(M, o, t);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (iter_seconds) iter_seconds[t - 1] = dt;
        if (k_hist) k_hist[t - 1] = (int32_t)M.cl.size();
        if (o->verbose) std::printf("oracle iter %d K=%d %.3fs\n", t, (int)M.cl.size(), dt);
    }
    std::copy(M.labels.begin(), M.labels.end(), labels);
    if (weights_out)
        for (int k = 0; k < (int)M.weights.size() && k < max_k_out; ++k) weights_out[k] = M.weights[k];
    return (int)M.cl.size();
}

struct orc_model {
    Model M;
    orc_run_opts o;
    std::vector<double> X;
};

orc_model* orc_model_create(const orc_run_opts* o, const double* X, int64_t N, int D, double k0, const double* m0,
                            double v0, const double* psi0, const double* alpha0) {
    orc_model* m = new orc_model();
    m->o = *o;
    m->X.assign(X, X + (size_t)N * D);
    model_init(m->M, o, m->X.data(), N, D, k0, m0, v0, psi0, alpha0);
    return m;
}

orc_model* orc_model_create_outlier(const orc_run_opts* o, const double* X, int64_t N, int D, double k0,
                                    const double* m0, double v0, const double* psi0, const double* alpha0, double ok0,
                                    const double* om0, double ov0, const double* opsi0, const double* oalpha0) {
    orc_model* m = new orc_model();
    m->o = *o;
    m->X.assign(X, X + (size_t)N * D);
    Hyper oh;
    if (o->prior == 0) {
        oh.k = ok0; oh.v = ov0; oh.m.assign(om0, om0 + D); oh.psi.assign(opsi0, opsi0 + (size_t)D * D);
    } else {
        oh.alpha.assign(oalpha0, oalpha0 + D);
    }
    model_init(m->M, o, m->X.data(), N, D, k0, m0, v0, psi0, alpha0, o->outlier_mod > 0 ? &oh : nullptr);
    return m;
}

void orc_model_destroy(orc_model* m) { delete m; }

void orc_model_trace(orc_model* m, int enable) { m->M.trace = enable != 0; }

int orc_model_step(orc_model* m, int t) {
    model_step(m->M, &m->o, t);
    return (int)m->M.cl.size();
}

int orc_model_num_clusters(const orc_model* m) { return (int)m->M.cl.size(); }

void orc_model_state(const orc_model* m, int32_t* labels, int32_t* subs, int64_t* counts, double* weights) {
    const Model& M = m->M;
    if (labels) std::copy(M.labels.begin(), M.labels.end(), labels);
    if (subs) std::copy(M.subs.begin(), M.subs.end(), subs);
    for (size_t k = 0; k < M.cl.size(); ++k) {
        if (counts) counts[k] = M.cl[k].c.st.N;
        if (weights) weights[k] = k < M.weights.size() ? M.weights[k] : 0.0;
    }
}

int orc_model_sweep_trace(const orc_model* m, int32_t* labels, int32_t* subs, double* margin, double* smargin) {
    const Model& M = m->M;
    if (!M.trace || M.tr_labels.size() != (size_t)M.N) return -1;
    std::copy(M.tr_labels.begin(), M.tr_labels.end(), labels);
    std::copy(M.tr_subs.begin(), M.tr_subs.end(), subs);
    std::copy(M.tr_margin.begin(), M.tr_margin.end(), margin);
    std::copy(M.tr_smargin.begin(), M.tr_smargin.end(), smargin);
    return 0;
}

}  // extern "C"
