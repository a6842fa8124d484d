// Host driver of the sub-cluster split/merge sampler (Chang & Fisher): the per-iteration
// control flow of local_clusters_actions::group_step (local_clusters_actions.cpp:460-490) and
// run_model (dp_parallel_sampling.cpp:175-249), driving the HIP device layer through its C ABI.
// The FP64 conjugate math (posteriors, log marginal likelihoods, Hastings ratios, parameter
// draws) runs here, parallel over clusters and merge pairs.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/dpmm_hip.h"
#include "dpmm_device.h"
#include "dpmm_rng.h"

namespace dpmm {
namespace {

constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------------ random streams -------
// Sequential stream (c0, n, t, tag): words of Philox blocks n = 0, 1, 2, ... in order.
class Stream {
  public:
    Stream(uint64_t seed, uint32_t c0, uint32_t t, uint32_t tag)
        : k0_((uint32_t)seed), k1_((uint32_t)(seed >> 32)), c0_(c0), t_(t), tag_(tag) {}
    uint32_t word() {
        if (pos_ == 4) {
            blk_ = philox4x32_10(c0_, n_++, t_, tag_, k0_, k1_);
            pos_ = 0;
        }
        return blk_.v[pos_++];
    }
    double uniform() { return u01_f64(word()); }
    double normal() {   // Box-Muller, cosine branch
        const double a = uniform();
        const double b = uniform();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * kPi * b);
    }
    double gamma(double a) {   // Marsaglia & Tsang (2000); a < 1 through Gamma(a + 1) U^(1/a)
        if (!(a > 0.0)) return 0.0;
        if (a < 1.0) {
            const double g = gamma(a + 1.0);
            return g * std::pow(uniform(), 1.0 / a);
        }
        const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            const double z = normal();
            double v = 1.0 + c * z;
            if (v <= 0.0) continue;
            v = v * v * v;
            if (std::log(uniform()) < 0.5 * z * z + d - d * v + d * std::log(v)) return d * v;
        }
    }
    void dirichlet(const double* a, int n, double* out) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += (out[i] = gamma(a[i]));
        for (int i = 0; i < n; ++i) out[i] /= s;
    }

  private:
    uint32_t k0_, k1_, c0_, t_, tag_;
    uint32_t n_ = 0;
    Ctr4 blk_{};
    int pos_ = 4;
};

// ------------------------------------------------------------------ dense FP64 ------------
using Mat = std::vector<double>;   // row-major D x D

bool cholesky(int D, Mat& A) {   // in place, lower factor, upper zeroed
    for (int j = 0; j < D; ++j) {
        double* rj = &A[(size_t)j * D];
        double s = rj[j];
        for (int k = 0; k < j; ++k) s -= rj[k] * rj[k];
        if (!(s > 0.0) || !std::isfinite(s)) return false;
        const double d = std::sqrt(s);
        rj[j] = d;
        const double inv = 1.0 / d;
        for (int i = j + 1; i < D; ++i) {
            double* ri = &A[(size_t)i * D];
            double t = ri[j];
            for (int k = 0; k < j; ++k) t -= ri[k] * rj[k];
            ri[j] = t * inv;
        }
        for (int k = j + 1; k < D; ++k) rj[k] = 0.0;
    }
    return true;
}

Mat lower_inverse(int D, const Mat& L) {
    Mat X((size_t)D * D, 0.0);
    for (int j = 0; j < D; ++j) {
        X[(size_t)j * D + j] = 1.0 / L[(size_t)j * D + j];
        for (int i = j + 1; i < D; ++i) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s += L[(size_t)i * D + k] * X[(size_t)k * D + j];
            X[(size_t)i * D + j] = -s / L[(size_t)i * D + i];
        }
    }
    return X;
}

Mat gram_t(int D, const Mat& X) {   // X^T X for lower-triangular X
    Mat G((size_t)D * D);
    for (int a = 0; a < D; ++a)
        for (int b = 0; b <= a; ++b) {
            double s = 0.0;
            for (int k = a; k < D; ++k) s += X[(size_t)k * D + a] * X[(size_t)k * D + b];
            G[(size_t)a * D + b] = s;
            G[(size_t)b * D + a] = s;
        }
    return G;
}

double logdet_spd(int D, const Mat& A) {
    Mat L = A;
    if (!cholesky(D, L)) return std::numeric_limits<double>::quiet_NaN();
    double s = 0.0;
    for (int i = 0; i < D; ++i) s += std::log(L[(size_t)i * D + i]);
    return 2.0 * s;
}

// ------------------------------------------------------------------ model ----------------
struct Hyp {          // niw_hyperparams / multinomial_hyper
    double k = 0, v = 0;
    std::vector<double> m, psi, alpha;
    double logdet_psi = 0;   // cached
    double alpha_sum = 0;    // cached
};
struct Suff {
    int64_t N = 0;
    std::vector<double> sum, S;
};
struct Draw {         // mv_gaussian through the factor R of invSigma, or log theta
    std::vector<double> mu, R, logtheta;
    double logdet_sigma = 0;
};
struct Part {
    Hyp post;
    Suff st;
    Draw dist;
};
struct Cluster {
    Part c, l, r;
    double lrw[2] = {0.5, 0.5};
    bool splittable = false;
    std::vector<std::pair<bool, double>> hist;
};

double lmvgamma(double x, int D) {
    double r = (double)D * (D - 1) / 4.0 * std::log(kPi);
    for (int d = 1; d <= D; ++d) r += std::lgamma(x + (1 - d) / 2.0);
    return r;
}

}  // namespace

struct SamplerImpl {
    dpmm_ctx* ctx = nullptr;
    dpmm_options o{};
    int64_t N = 0;
    int D = 0;
    int prior = 0;
    Hyp h0;
    std::vector<Cluster> cl;
    std::vector<double> weights;
    double host_ms = 0;
    // scratch
    std::vector<int64_t> Nsub;
    std::vector<double> ssum, sS;

    bool gauss() const { return prior == DPMM_PRIOR_GAUSSIAN; }

    // niw::calc_posterior (priors/niw.cpp:23-49); multinomial_prior::calc_posterior (12-22).
    Hyp posterior(const Suff& s) const {
        if (s.N == 0) return h0;
        Hyp p;
        if (!gauss()) {
            p.alpha.resize(D);
            p.alpha_sum = 0;
            for (int d = 0; d < D; ++d) p.alpha_sum += (p.alpha[d] = h0.alpha[d] + s.sum[d]);
            return p;
        }
        const double n = (double)s.N;
        p.k = h0.k + n;
        p.v = h0.v + n;
        p.m.resize(D);
        for (int d = 0; d < D; ++d) p.m[d] = (h0.k * h0.m[d] + s.sum[d]) / p.k;
        p.psi.resize((size_t)D * D);
        const double iv = 1.0 / p.v;
        for (int a = 0; a < D; ++a)
            for (int b = a; b < D; ++b) {
                const double val = (h0.v * h0.psi[(size_t)a * D + b] + h0.k * h0.m[a] * h0.m[b] + s.S[(size_t)a * D + b] -
                                    p.k * p.m[a] * p.m[b]) * iv;
                p.psi[(size_t)a * D + b] = val;
                p.psi[(size_t)b * D + a] = val;
            }
        p.logdet_psi = logdet_spd(D, p.psi);
        return p;
    }

    // niw::log_marginal_likelihood (priors/niw.cpp:122-138) / multinomial (56-67).
    double log_marginal(const Hyp& p, const Suff& s) const {
        if (!gauss()) {
            double v = std::lgamma(h0.alpha_sum) - std::lgamma(p.alpha_sum);
            for (int d = 0; d < D; ++d) v += std::lgamma(p.alpha[d]) - std::lgamma(h0.alpha[d]);
            return v;
        }
        return -(double)(s.N * (int64_t)D) * 0.5 * std::log(kPi) + lmvgamma(p.v / 2.0, D) - lmvgamma(h0.v / 2.0, D) +
               (h0.v / 2.0) * (D * std::log(h0.v) + h0.logdet_psi) - (p.v / 2.0) * (D * std::log(p.v) + p.logdet_psi) +
               (D / 2.0) * std::log(h0.k / p.k);
    }

    // niw::sample_distribution (priors/niw.cpp:82-107): invSigma = R R^T with R = B A, B the
    // Cholesky factor of (v psi)^-1 and A the Bartlett factor (rinvwish.ipp:35-58).  R is
    // exactly the Cholesky factor the sweep needs, so no factorisation of invSigma follows.
    bool draw(const Hyp& p, uint32_t slot, uint32_t t, Draw& out) const {
        Stream rs(o.seed, slot, t, TAG_PARAM);
        if (!gauss()) {
            out.logtheta.resize(D);
            rs.dirichlet(p.alpha.data(), D, out.logtheta.data());
            for (auto& x : out.logtheta) x = std::log(x);
            return true;
        }
        Mat C(p.psi);
        for (auto& x : C) x *= p.v;
        if (!cholesky(D, C)) return false;
        Mat B = gram_t(D, lower_inverse(D, C));   // (v psi)^-1 = C^-T C^-1
        if (!cholesky(D, B)) return false;
        Mat A((size_t)D * D, 0.0);
        for (int i = 1; i < D; ++i)
            for (int j = 0; j < i; ++j) A[(size_t)i * D + j] = rs.normal();
        for (int i = 0; i < D; ++i) A[(size_t)i * D + i] = std::sqrt(2.0 * rs.gamma((p.v - i) / 2.0));
        out.R.assign((size_t)D * D, 0.0);
        double ld = 0.0;
        for (int i = 0; i < D; ++i) {
            for (int j = 0; j <= i; ++j) {
                double s = 0.0;
                for (int k = j; k <= i; ++k) s += B[(size_t)i * D + k] * A[(size_t)k * D + j];
                out.R[(size_t)i * D + j] = s;
            }
            ld += std::log(out.R[(size_t)i * D + i]);
        }
        out.logdet_sigma = -2.0 * ld;
        Mat Sig = gram_t(D, lower_inverse(D, out.R));   // Sigma = R^-T R^-1
        for (auto& x : Sig) x /= p.k;
        if (!cholesky(D, Sig)) return false;
        out.mu.resize(D);
        std::vector<double> z(D);
        for (int i = 0; i < D; ++i) z[i] = rs.normal();
        for (int i = 0; i < D; ++i) {
            double x = p.m[i];
            if (o.full_mu_cov)
                for (int j = 0; j <= i; ++j) x += Sig[(size_t)i * D + j] * z[j];
            else
                x += Sig[(size_t)i * D + i] * z[i];   // the reference reads only the diagonal (pdflib.cpp:4007-4014)
            out.mu[i] = x;
        }
        return true;
    }

    void reset_hist(Cluster& c) const { c.hist.assign(o.burnout + 5, std::make_pair(false, 0.0)); }

    Suff empty_suff() const {
        Suff s;
        s.sum.assign(D, 0.0);
        if (gauss()) s.S.assign((size_t)D * D, 0.0);
        return s;
    }

    static void add_suff(Suff& a, const Suff& b) {
        a.N += b.N;
        for (size_t i = 0; i < a.sum.size(); ++i) a.sum[i] += b.sum[i];
        for (size_t i = 0; i < a.S.size(); ++i) a.S[i] += b.S[i];
    }

    // create_sufficient_statistics + update_suff_stats_posterior (local_clusters_actions.cpp:
    // 109-168) for the clusters flagged in `which` (empty: all).
    int update_stats(const std::vector<char>& which) {
        const int K = (int)cl.size();
        Nsub.resize(2 * K);
        ssum.resize((size_t)2 * K * D);
        if (gauss()) sS.resize((size_t)2 * K * D * D);
        int rc = dpmm_sufficient_statistics(ctx, K, Nsub.data(), ssum.data(), gauss() ? sS.data() : nullptr);
        if (rc) return rc;
        const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic)
        for (int k = 0; k < K; ++k) {
            if (!which.empty() && !which[k]) continue;
            Cluster& c = cl[k];
            for (int s = 0; s < 2; ++s) {
                Suff& st = s ? c.r.st : c.l.st;
                const int q = 2 * k + s;
                st.N = Nsub[q];
                st.sum.assign(ssum.begin() + (size_t)q * D, ssum.begin() + (size_t)(q + 1) * D);
                if (gauss()) st.S.assign(sS.begin() + (size_t)q * D * D, sS.begin() + (size_t)(q + 1) * D * D);
            }
            c.c.st = c.l.st;   // the cluster's statistics aggregate its two sub-clusters (niw.cpp:140-148)
            add_suff(c.c.st, c.r.st);
            c.c.post = posterior(c.c.st);
            c.l.post = posterior(c.l.st);
            c.r.post = posterior(c.r.st);
        }
        host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return DPMM_OK;
    }

    // sample_clusters (local_clusters_actions.cpp:341-371) + sample_cluster_params
    // (shared_actions.cpp:9-55).
    int sample_clusters(uint32_t t) {
        const auto t0 = std::chrono::steady_clock::now();
        const int K = (int)cl.size();
        int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : bad)
        for (int k = 0; k < K; ++k) {
            Cluster& c = cl[k];
            bad += !draw(c.c.post, 3 * k + 0, t, c.c.dist);
            bad += !draw(c.l.post, 3 * k + 1, t, c.l.dist);
            bad += !draw(c.r.post, 3 * k + 2, t, c.r.dist);
            const double a[2] = {(double)c.l.st.N + o.alpha / 2, (double)c.r.st.N + o.alpha / 2};
            Stream(o.seed, k, t, TAG_LRW).dirichlet(a, 2, c.lrw);
            const double L = log_marginal(c.l.post, c.l.st) + log_marginal(c.r.post, c.r.st);
            const int B = o.burnout;
            std::rotate(c.hist.begin(), c.hist.begin() + 1, c.hist.begin() + B);
            c.hist[B - 1] = std::make_pair(true, L);
            double now = 0.0;
            bool ready = true;
            for (int i = 0; i < B; ++i) {
                if (c.hist[i].first) now += c.hist[i].second * (1 / (B - 0.1));
                else ready = false;
            }
            if (ready && now - c.hist[B - 1].second < 1e-2) c.splittable = true;
        }
        // mixture weights ~ Dir(N_1, ..., N_K, alpha), last entry dropped (357-365)
        std::vector<double> a(K + 1), w(K + 1);
        for (int k = 0; k < K; ++k) a[k] = (double)cl[k].c.st.N;
        a[K] = o.alpha;
        if (K == 1 && a[0] == 0.0) { w[0] = 0.0; w[1] = 1.0; }   // get_dirichlet_distribution (323-339)
        else Stream(o.seed, 0, t, TAG_MIXW).dirichlet(a.data(), K + 1, w.data());
        weights.assign(w.begin(), w.begin() + K);
        host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (bad) {
            return -4;
        }
        return DPMM_OK;
    }

    int upload_params() {
        const int K = (int)cl.size();
        std::vector<double> lrw(2 * K);
        for (int k = 0; k < K; ++k) { lrw[2 * k] = cl[k].lrw[0]; lrw[2 * k + 1] = cl[k].lrw[1]; }
        if (gauss()) {
            std::vector<double> mu((size_t)K * D), R((size_t)K * D * D), ld(K);
            std::vector<double> smu((size_t)2 * K * D), sR((size_t)2 * K * D * D), sld(2 * K);
            for (int k = 0; k < K; ++k) {
                const Draw* ds[3] = {&cl[k].c.dist, &cl[k].l.dist, &cl[k].r.dist};
                std::copy(ds[0]->mu.begin(), ds[0]->mu.end(), mu.begin() + (size_t)k * D);
                std::copy(ds[0]->R.begin(), ds[0]->R.end(), R.begin() + (size_t)k * D * D);
                ld[k] = ds[0]->logdet_sigma;
                for (int s = 0; s < 2; ++s) {
                    const int q = 2 * k + s;
                    std::copy(ds[1 + s]->mu.begin(), ds[1 + s]->mu.end(), smu.begin() + (size_t)q * D);
                    std::copy(ds[1 + s]->R.begin(), ds[1 + s]->R.end(), sR.begin() + (size_t)q * D * D);
                    sld[q] = ds[1 + s]->logdet_sigma;
                }
            }
            return dpmm_set_gaussian_factors(ctx, K, mu.data(), R.data(), ld.data(), weights.data(), smu.data(),
                                             sR.data(), sld.data(), lrw.data());
        }
        std::vector<double> lt((size_t)K * D), slt((size_t)2 * K * D);
        for (int k = 0; k < K; ++k) {
            std::copy(cl[k].c.dist.logtheta.begin(), cl[k].c.dist.logtheta.end(), lt.begin() + (size_t)k * D);
            std::copy(cl[k].l.dist.logtheta.begin(), cl[k].l.dist.logtheta.end(), slt.begin() + (size_t)(2 * k) * D);
            std::copy(cl[k].r.dist.logtheta.begin(), cl[k].r.dist.logtheta.end(), slt.begin() + (size_t)(2 * k + 1) * D);
        }
        return dpmm_set_multinomial_params(ctx, K, lt.data(), weights.data(), slt.data(), lrw.data(), o.mult_use_weight);
    }

    int relabel(uint32_t t, const std::vector<int32_t>& nl, const std::vector<int32_t>& act) {
        return dpmm_relabel(ctx, t, (int)act.size(), nl.data(), act.data());
    }

    int init() {
        int rc = dpmm_sample_initial_labels(ctx, std::max(1, o.init_clusters));
        if (rc) return rc;
        const int K0 = std::max(1, o.init_clusters);
        cl.assign(K0, Cluster());
        for (auto& c : cl) {
            c.c.post = h0;
            c.c.st = empty_suff();
            c.l = c.c;
            c.r = c.c;
            reset_hist(c);
        }
        rc = update_stats({});
        if (rc) return rc;
        return sample_clusters(0);
    }

    int step(int t) {
        const bool fin = t >= o.iters - o.argmax_sample_stop;
        const bool no_more_splits = t >= o.iters - o.split_stop || (double)cl.size() >= o.max_clusters;
        const uint32_t it = (uint32_t)t;
        int rc;
        host_ms = 0;
        if (t > 0) {   // group_step resamples parameters first (local_clusters_actions.cpp:462)
            rc = sample_clusters(it);
            if (rc) return rc;
        }
        rc = upload_params();
        if (rc) return rc;
        rc = dpmm_sample_cluster_labels(ctx, it, fin ? 1 : 0);
        if (rc) return rc;
        rc = dpmm_sample_sub_labels(ctx, it);
        if (rc) return rc;
        rc = update_stats({});
        if (rc) return rc;

        // reset_bad_clusters (local_clusters_actions.cpp:429-447)
        int K = (int)cl.size();
        {
            std::vector<char> bad(K, 0);
            bool any = false;
            for (int k = 0; k < K; ++k)
                if (cl[k].l.st.N == 0 || cl[k].r.st.N == 0) {
                    bad[k] = 1;
                    any = true;
                    reset_hist(cl[k]);
                    cl[k].splittable = false;
                }
            if (any) {
                std::vector<int32_t> nl(2 * K), act(K, 0);
                for (int q = 0; q < 2 * K; ++q) nl[q] = q / 2;
                for (int k = 0; k < K; ++k) act[k] = bad[k] ? 1 : 0;
                if ((rc = relabel(it, nl, act))) return rc;
                if ((rc = update_stats(bad))) return rc;
            }
        }

        std::vector<int> merge_into;   // per cluster: absorbing cluster or -1
        if (!no_more_splits) {
            // check_and_split (local_clusters_actions.cpp:244-285), should_split_local (215-242)
            const auto t0 = std::chrono::steady_clock::now();
            std::vector<char> acc(K, 0);
#pragma omp parallel for schedule(dynamic)
            for (int k = 0; k < K; ++k) {
                const Cluster& c = cl[k];
                if (!(c.splittable && c.c.st.N > 1) || fin || c.l.st.N == 0 || c.r.st.N == 0) continue;
                const double L = log_marginal(posterior(c.c.st), c.c.st);
                const double Ll = log_marginal(posterior(c.l.st), c.l.st);
                const double Lr = log_marginal(posterior(c.r.st), c.r.st);
                const double lhr = std::log(o.alpha) + std::lgamma((double)c.l.st.N) + Ll +
                                   std::lgamma((double)c.r.st.N) + Lr - (std::lgamma((double)c.c.st.N) + L);
                acc[k] = lhr > std::log(Stream(o.seed, k, it, TAG_SPLIT_U).uniform());
            }
            host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::vector<int32_t> nl(2 * K), act(K, 0);
            for (int q = 0; q < 2 * K; ++q) nl[q] = q / 2;
            std::vector<char> touched;
            int nsplit = 0;
            for (int k = 0; k < K; ++k) {
                if (!acc[k]) continue;
                const int nk = (int)cl.size();
                // split_cluster_local (183-194): k keeps the left half, nk gets the right half
                Cluster left, right;
                left.c = cl[k].l; right.c = cl[k].r;
                left.l = left.r = left.c;
                right.l = right.r = right.c;
                reset_hist(left);
                reset_hist(right);
                cl[k] = left;
                cl.push_back(right);
                nl[2 * k + 1] = nk;
                act[k] = 2;
                ++nsplit;
            }
            if (nsplit) {
                if ((rc = relabel(it, nl, act))) return rc;
                touched.assign(cl.size(), 0);
                for (int k = 0; k < K; ++k) touched[k] = acc[k];
                for (size_t k = K; k < cl.size(); ++k) touched[k] = 1;
                if ((rc = update_stats(touched))) return rc;
            }

            // check_and_merge (287-321) + should_merge (shared_actions.cpp:88-111): every
            // eligible pair is tested on pre-merge statistics, then accepted greedily in
            // lexicographic order; a merged cluster leaves the eligible set, which makes the
            // outcome identical to the reference's sequential loop.
            const auto t1 = std::chrono::steady_clock::now();
            K = (int)cl.size();
            std::vector<int> elig;
            for (int k = 0; k < K; ++k)
                if (cl[k].splittable && cl[k].c.st.N > 0) elig.push_back(k);
            const int ne = (int)elig.size();
            std::vector<double> Lc(K, 0.0);
            for (int k : elig) Lc[k] = log_marginal(cl[k].c.post, cl[k].c.st);
            std::vector<std::pair<int, int>> pairs;
            for (int a = 0; a < ne; ++a)
                for (int b = a + 1; b < ne; ++b) pairs.emplace_back(elig[a], elig[b]);
            std::vector<char> ok(pairs.size(), 0);
            const double al = o.alpha;
            const double cst = -std::log(al) + std::lgamma(al) - 2 * std::lgamma(0.5 * al);
#pragma omp parallel for schedule(dynamic)
            for (int64_t p = 0; p < (int64_t)pairs.size(); ++p) {
                const int i = pairs[p].first, j = pairs[p].second;
                Suff m = cl[i].c.st;
                add_suff(m, cl[j].c.st);
                const double L = log_marginal(posterior(m), m);
                const double Ni = (double)cl[i].c.st.N, Nj = (double)cl[j].c.st.N, Nm = Ni + Nj;
                const double lhr = cst + std::lgamma(Nm) - std::lgamma(Nm + al) + std::lgamma(Ni + 0.5 * al) -
                                   std::lgamma(Ni) - std::lgamma(Nj) + std::lgamma(Nj + 0.5 * al) + L - Lc[i] - Lc[j];
                const double u = Stream(o.seed, ((uint32_t)i << 16) | (uint32_t)j, it, TAG_MERGE_U).uniform();
                ok[p] = (lhr > std::log(u)) || (fin && lhr > std::log(0.1));
            }
            merge_into.assign(K, -1);
            std::vector<char> used(K, 0);
            for (size_t p = 0; p < pairs.size(); ++p) {
                const int i = pairs[p].first, j = pairs[p].second;
                if (!ok[p] || used[i] || used[j]) continue;
                used[i] = used[j] = 1;
                merge_into[j] = i;
                // merge_clusters (204-213) + merge_clusters_to_splittable (shared_actions.cpp:72-86)
                Cluster& a = cl[i];
                Cluster& b = cl[j];
                add_suff(a.c.st, b.c.st);
                a.c.post = posterior(a.c.st);
                a.l = a.c;   // the reference clones the merged parameters into the left sub-cluster
                a.r = b.c;
                a.splittable = false;
                reset_hist(a);
                b.c.st.N = 0;
                b.splittable = false;
            }
            host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        }

        // merges + remove_empty_clusters (local_clusters_actions.cpp:384-419) in one relabel pass
        K = (int)cl.size();
        std::vector<int32_t> remap(K, -1);
        int nk = 0;
        for (int k = 0; k < K; ++k)
            if (cl[k].c.st.N > 0) remap[k] = nk++;
        bool any_merge = false;
        for (int k = 0; k < (int)merge_into.size(); ++k) any_merge |= merge_into[k] >= 0;
        if (any_merge || nk != K) {
            std::vector<int32_t> nl(2 * K), act(K, 0);
            for (int k = 0; k < K; ++k) {
                int target = k;
                if (k < (int)merge_into.size() && merge_into[k] >= 0) { target = merge_into[k]; act[k] = 4; }
                const int32_t lab = remap[target] >= 0 ? remap[target] : 0;
                nl[2 * k] = nl[2 * k + 1] = lab;
            }
            for (int k = 0; k < (int)merge_into.size(); ++k)
                if (merge_into[k] >= 0) act[merge_into[k]] = 3;
            if ((rc = relabel(it, nl, act))) return rc;
            std::vector<Cluster> kept;
            kept.reserve(nk);
            for (int k = 0; k < K; ++k)
                if (remap[k] >= 0) kept.push_back(std::move(cl[k]));
            cl.swap(kept);
        }
        return DPMM_OK;
    }
};

}  // namespace dpmm

struct dpmm_sampler {
    dpmm::SamplerImpl s;
};

extern "C" {

int dpmm_sampler_create(dpmm_ctx* ctx, const dpmm_options* opt, double k0, const double* m0, double v0,
                        const double* psi0, const double* alpha0, dpmm_sampler** out) {
    if (!ctx || !opt || !out) return DPMM_EINVAL;
    *out = nullptr;
    int64_t N;
    int D, prior, prec;
    dpmm::ctx_dims(ctx, &N, &D, &prior, &prec);
    if (opt->prior != prior || opt->iters < 0 || opt->burnout < 1 || !(opt->alpha > 0)) return DPMM_EINVAL;
    auto* S = new dpmm_sampler();
    dpmm::SamplerImpl& s = S->s;
    s.ctx = ctx;
    s.o = *opt;
    s.N = N;
    s.D = D;
    s.prior = prior;
    if (prior == DPMM_PRIOR_GAUSSIAN) {
        if (!m0 || !psi0 || !(k0 > 0) || !(v0 > D - 1)) { delete S; return DPMM_EINVAL; }
        s.h0.k = k0;
        s.h0.v = v0;
        s.h0.m.assign(m0, m0 + D);
        s.h0.psi.assign(psi0, psi0 + (size_t)D * D);
        s.h0.logdet_psi = dpmm::logdet_spd(D, s.h0.psi);
        if (!std::isfinite(s.h0.logdet_psi)) { delete S; return DPMM_ENOTSPD; }
    } else {
        if (!alpha0) { delete S; return DPMM_EINVAL; }
        s.h0.alpha.assign(alpha0, alpha0 + D);
        s.h0.alpha_sum = 0;
        for (double a : s.h0.alpha) {
            if (!(a > 0)) { delete S; return DPMM_EINVAL; }
            s.h0.alpha_sum += a;
        }
    }
    if (opt->precision != prec) {
        int rc = dpmm_set_precision(ctx, opt->precision);
        if (rc) { delete S; return rc; }
    }
    int rc = s.init();
    if (rc) { delete S; return rc; }
    *out = S;
    return DPMM_OK;
}

int dpmm_sampler_step(dpmm_sampler* S, int iter) {
    if (!S || iter < 1) return DPMM_EINVAL;
    return S->s.step(iter);
}

int dpmm_sampler_num_clusters(dpmm_sampler* S) { return S ? (int)S->s.cl.size() : DPMM_EINVAL; }

int dpmm_sampler_weights(dpmm_sampler* S, double* w, int64_t* counts, int max_k) {
    if (!S) return DPMM_EINVAL;
    const int K = (int)S->s.cl.size();
    for (int k = 0; k < K && k < max_k; ++k) {
        if (w) w[k] = k < (int)S->s.weights.size() ? S->s.weights[k] : 0.0;
        if (counts) counts[k] = S->s.cl[k].c.st.N;
    }
    return DPMM_OK;
}

int dpmm_sampler_host_ms(dpmm_sampler* S, double* ms) {
    if (!S || !ms) return DPMM_EINVAL;
    *ms = S->s.host_ms;
    return DPMM_OK;
}

int dpmm_sampler_destroy(dpmm_sampler* S) {
    delete S;
    return DPMM_OK;
}

int dpmm_run(dpmm_ctx* ctx, const dpmm_options* opt, double k0, const double* m0, double v0, const double* psi0,
             const double* alpha0, int32_t* labels, int32_t* k_hist, double* iter_seconds, double* weights, int max_k,
             int* final_k) {
    dpmm_sampler* S = nullptr;
    int rc = dpmm_sampler_create(ctx, opt, k0, m0, v0, psi0, alpha0, &S);
    if (rc) return rc;
    for (int t = 1; t <= opt->iters; ++t) {
        const auto t0 = std::chrono::steady_clock::now();
        rc = dpmm_sampler_step(S, t);
        if (rc) break;
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (iter_seconds) iter_seconds[t - 1] = dt;
        if (k_hist) k_hist[t - 1] = dpmm_sampler_num_clusters(S);
        if (opt->verbose)
            std::printf("Iteration: %d || Clusters count: %d || Iter Time: %f\n", t, dpmm_sampler_num_clusters(S), dt);
    }
    if (!rc) {
        if (final_k) *final_k = dpmm_sampler_num_clusters(S);
        if (weights) dpmm_sampler_weights(S, weights, nullptr, max_k);
        if (labels) rc = dpmm_get_labels(ctx, labels, nullptr);
    }
    dpmm_sampler_destroy(S);
    return rc;
}

}  // extern "C"
