// Host driver of the sub-cluster split/merge sampler (Chang & Fisher): the per-iteration
// control flow of local_clusters_actions::group_step (local_clusters_actions.cpp:460-490) and
// run_model (dp_parallel_sampling.cpp:175-249), driving the HIP device layer through its C ABI.
//
// The host holds, per cluster part (cluster, left, right), only its point count, the list of
// sub-cluster statistics entries whose sum it is, and that sum's log marginal likelihood.  All
// FP64 matrix work runs on the device in batched launches: one log-marginal launch per statistics
// pass covers every cluster, sub-cluster and candidate merge pair, and one draw launch samples
// every part's parameters and writes the sweeps' parameter slots.  The host does the scalar work:
// Dirichlet weights, burn-out history, Hastings ratios and the greedy merge resolution.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dpmm_hip.h"
#include "dpmm_device.h"
#include "dpmm_rng.h"

namespace dpmm {
namespace {

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
        return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * 3.14159265358979323846 * b);
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

// Mixture weights ~ Dir(a_0, ..., a_{n-1}): component i draws its gamma from stream (i, t, TAG_MIXW),
// so that the device draws the components in parallel; the sum runs in index order.
void mixture_dirichlet(uint64_t seed, uint32_t t, const double* a, int n, double* out) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (out[i] = Stream(seed, (uint32_t)i, t, TAG_MIXW).gamma(a[i]));
    for (int i = 0; i < n; ++i) out[i] /= s;
}

// ------------------------------------------------------------------ model ----------------
// A cluster part: the sum of up to four sub-cluster statistics entries of the last statistics
// pass (entry q = 2k + s), its point count and the log marginal likelihood of that sum.
struct Part {
    int64_t N = 0;
    int32_t mem[4] = {-1, -1, -1, -1};
    double lml = 0.0;
    double ld = 0.0;     // log det psi_n of the posterior
    int prior = 0;       // 0: the model's prior, 1: the outlier cluster's (outlier mode)
    double lml_split = 0.0;   // log ML with the posterior formed from the cluster's prior (the split test)
};
// Where a part's parameters come from when the draw call does not redraw them (outlier mode leaves
// cluster index 1 undrawn): a row of the last draw call, or a draw on the Philox stream (slot, iter) from
// the posterior of the statistics entries `mem` (parameters the reference drew at a split).
struct ParamRef {
    int32_t row = -1;
    int32_t slot = -1, iter = -1;
    int32_t mem[4] = {-1, -1, -1, -1};
    int prior = 0;
};
struct Cluster {
    Part c, l, r;
    double lrw[2] = {0.5, 0.5};
    bool splittable = false;
    std::vector<std::pair<bool, double>> hist;
    ParamRef pr[3];
};

// A part's statistics entries and count after a statistics pass (its prior stays).
void set_entries(Part& p, int64_t N, int q0, int q1) {
    p.N = N;
    p.mem[0] = q0;
    p.mem[1] = q1;
    p.mem[2] = p.mem[3] = -1;
}
ParamRef keep_row(int row) {
    ParamRef r;
    r.row = row;
    return r;
}
ParamRef draw_from(const int32_t* mem, int prior, int slot, int iter) {
    ParamRef r;
    r.slot = slot;
    r.iter = iter;
    std::copy(mem, mem + 4, r.mem);
    r.prior = prior;
    return r;
}

}  // namespace

struct SamplerImpl {
    dpmm_ctx* ctx = nullptr;
    dpmm_options o{};
    int64_t N = 0;
    int D = 0;
    int prior = 0;
    int kcap = 0;   // the device kernels' largest cluster count (dpmm_max_clusters)
    bool outlier = false;   // outlier mode (cluster 0 is the outlier cluster)
    std::vector<Cluster> cl;
    std::vector<double> weights;
    double host_ms = 0;
    std::vector<int64_t> Nsub;
    // merge candidates of the current iteration: clusters and log ML of every pair's union (NaN
    // where the screener skipped the pair)
    std::vector<int> cand, cand_pos;
    std::vector<double> cand_L;
    int ncand = 0;
    // merge screener
    dpmm_screener* scr = nullptr;
    int scr_every = 10;
    float scr_thr = 0.f;
    bool log_pairs = false;
    bool full_scan = true;
    std::vector<float> pair_feat, pair_logit;
    std::vector<float> log_feat;
    std::vector<double> log_h, log_u;
    int64_t n_exact = 0, n_skipped = 0, n_missed = 0, n_full = 0;
    // merge bound of the device iteration (NIW): pairs rejected without their log ML
    bool merge_bound = true;
    int64_t n_bound = 0;
    // totals since creation: splits, merges, resets, clusters removed (merged away or emptied)
    int64_t ev_split = 0, ev_merge = 0, ev_reset = 0, ev_removed = 0;
    bool counts_fresh = false;   // Nsub holds the counts of the current labels (from the sweep call)
    int32_t* trace_labels = nullptr;   // dpmm_sampler_trace_sweep: host copies of each step's sweep output
    int32_t* trace_subs = nullptr;
    // Device-resident iterations: while dev_on, the cluster list lives in the device table (Kdev
    // records, the last draw's Kw_dev mixture weights); cl and weights are refreshed from it when host
    // code reads them (host_fresh).
    bool dev_on = false;
    bool host_fresh = true;
    int Kdev = 0, Kw_dev = 0;

    bool device_step_eligible() const { return !o.host_decisions && !outlier && !scr && !log_pairs; }

    int num_clusters() const { return dev_on ? Kdev : (int)cl.size(); }

    // The host cluster list into the device table (history: the valid entries, oldest first).
    int push_to_device() {
        const int K = (int)cl.size(), B = o.burnout;
        std::vector<StepCluster> v(K);
        std::vector<double> h((size_t)K * B, 0.0);
        auto put = [](const Part& p, StepPart& q) {
            q.N = p.N;
            std::copy(p.mem, p.mem + 4, q.mem);
            q.lml = p.lml;
            q.ld = p.ld;
        };
        for (int k = 0; k < K; ++k) {
            const Cluster& c = cl[k];
            StepCluster& d = v[k];
            std::memset(&d, 0, sizeof d);
            put(c.c, d.c);
            put(c.l, d.l);
            put(c.r, d.r);
            d.lrw[0] = c.lrw[0];
            d.lrw[1] = c.lrw[1];
            d.splittable = c.splittable ? 1 : 0;
            int n = 0;
            for (int i = 0; i < B; ++i)
                if (c.hist[i].first) h[(size_t)k * B + n++] = c.hist[i].second;
            d.hist_n = n;
            d.hist_head = n % B;
        }
        int rc = step_upload(ctx, B, K, v.data(), h.data());
        if (rc) return rc;
        dev_on = true;
        host_fresh = true;
        Kdev = K;
        Kw_dev = 0;
        return DPMM_OK;
    }

    // The device table into the host cluster list and weights, when they are stale.
    int pull_from_device() {
        if (!dev_on || host_fresh) return DPMM_OK;
        const int K = Kdev, B = o.burnout;
        std::vector<StepCluster> v(K);
        std::vector<double> h((size_t)K * B), w(Kw_dev);
        int rc = step_download(ctx, K, Kw_dev, v.data(), h.data(), w.data());
        if (rc) return rc;
        auto get = [](const StepPart& q, Part& p) {
            p.N = q.N;
            std::copy(q.mem, q.mem + 4, p.mem);
            p.lml = p.lml_split = q.lml;
            p.ld = q.ld;
            p.prior = 0;
        };
        cl.assign(K, Cluster());
        for (int k = 0; k < K; ++k) {
            Cluster& c = cl[k];
            const StepCluster& d = v[k];
            get(d.c, c.c);
            get(d.l, c.l);
            get(d.r, c.r);
            c.lrw[0] = d.lrw[0];
            c.lrw[1] = d.lrw[1];
            c.splittable = d.splittable != 0;
            reset_hist(c);
            for (int j = 0; j < d.hist_n; ++j) {
                int q = d.hist_head - d.hist_n + j;
                if (q < 0) q += B;
                c.hist[B - d.hist_n + j] = std::make_pair(true, h[(size_t)k * B + q]);
            }
        }
        weights = w;
        host_fresh = true;
        return DPMM_OK;
    }

    // group_step on the device (step_iteration): the same flags as step() below, one synchronisation.
    int step_device(int t) {
        int rc;
        if (!dev_on && (rc = push_to_device())) return rc;
        const int K = Kdev;
        const bool fin = t >= o.iters - o.argmax_sample_stop;
        const bool no_more_splits = t >= o.iters - o.split_stop || (double)K >= o.max_clusters || K >= kcap;
        StepArgs a{};
        a.key0 = (uint32_t)o.seed;
        a.key1 = (uint32_t)(o.seed >> 32);
        a.t = (uint32_t)t;
        a.K = K;
        a.fin = fin ? 1 : 0;
        a.no_more_splits = no_more_splits ? 1 : 0;
        a.kcap = kcap;
        a.alpha = o.alpha;
        a.merge_bound = merge_bound && prior == DPMM_PRIOR_GAUSSIAN ? 1 : 0;
        StepSummary sm{};
        host_ms = 0;
        rc = step_iteration(ctx, a, (fin || o.hard_clustering) ? 1 : 0, o.full_mu_cov, o.mult_use_weight, trace_labels,
                            trace_subs, &sm);
        if (rc) return rc;
        n_exact += sm.npairs;
        n_bound += sm.nbound;
        ev_split += sm.nsplit;
        ev_merge += sm.nmerge;
        ev_reset += sm.nreset;
        ev_removed += sm.K_split - sm.K_new;
        Kw_dev = K;
        Kdev = sm.K_new;
        host_fresh = false;
        return DPMM_OK;
    }

    void reset_hist(Cluster& c) const { c.hist.assign(o.burnout + 5, std::make_pair(false, 0.0)); }

    // Log-ML jobs of one device launch: up to four statistics entries each, a prior selector (bit 0 the
    // posterior's prior, bit 1 the prior of the marginal's terms) and where the results go.
    struct Jobs {
        std::vector<int32_t> mem, sel;
        std::vector<double*> lml, ld;
        void add(const int32_t* m, int s, double* l, double* d) {
            mem.insert(mem.end(), m, m + 4);
            sel.push_back(s);
            lml.push_back(l);
            ld.push_back(d);
        }
        int size() const { return (int)sel.size(); }
        void clear() { mem.clear(); sel.clear(); lml.clear(); ld.clear(); }
    };

    int run_jobs(Jobs& jb) {
        const int n = jb.size();
        if (n == 0) return DPMM_OK;
        bool any_sel = false;
        for (int s : jb.sel) any_sel |= s != 0;
        std::vector<double> L(n), ld(n);
        int rc = dpmm_log_marginal_batch_ex(ctx, n, jb.mem.data(), any_sel ? jb.sel.data() : nullptr, L.data(),
                                            ld.data());
        if (rc) return rc;
        for (int j = 0; j < n; ++j) {
            if (jb.ld[j]) {   // a part: its log ML must exist (pairs keep NaN for a failed factorisation)
                if (!std::isfinite(L[j])) return DPMM_ENOTSPD;
                *jb.ld[j] = ld[j];
            }
            *jb.lml[j] = L[j];
        }
        jb.clear();
        return DPMM_OK;
    }

    // create_sufficient_statistics + update_suff_stats_posterior (local_clusters_actions.cpp:
    // 109-168) for the clusters flagged in `which` (empty: all), with the log marginal likelihoods
    // of their parts from one device launch.  With `pairs`, the same launch also evaluates the
    // union of every pair of merge candidates (splittable, non-empty clusters; in outlier mode not
    // the outlier cluster, local_clusters_actions.cpp:295) that the screener, when set, passes.
    int update_stats(const std::vector<char>& which, bool pairs, uint32_t t = 0) {
        const int K = (int)cl.size();
        int rc;
        if (!counts_fresh) {
            Nsub.resize(2 * K);
            if ((rc = dpmm_sufficient_statistics(ctx, K, Nsub.data(), nullptr, nullptr))) return rc;
        }
        counts_fresh = false;
        const auto t0 = std::chrono::steady_clock::now();
        Jobs jb;
        for (int k = 0; k < K; ++k) {
            if (!which.empty() && !which[k]) continue;
            Cluster& c = cl[k];
            set_entries(c.l, Nsub[2 * k], 2 * k, -1);
            set_entries(c.r, Nsub[2 * k + 1], 2 * k + 1, -1);
            set_entries(c.c, Nsub[2 * k] + Nsub[2 * k + 1], 2 * k, 2 * k + 1);   // cluster = l + r (niw.cpp:140-148)
            for (Part* pp : {&c.c, &c.l, &c.r}) jb.add(pp->mem, pp->prior * 3, &pp->lml, &pp->ld);
            // the split test forms the sub-clusters' posteriors from the cluster's prior
            // (local_clusters_actions.cpp:221-228): a job of its own where a sub-cluster has another prior
            for (Part* sp : {&c.l, &c.r}) {
                if (sp->prior == c.c.prior) continue;
                jb.add(sp->mem, c.c.prior | (sp->prior << 1), &sp->lml_split, nullptr);
            }
        }
        std::vector<size_t> tested;   // pair indices whose exact test joins this launch
        if (pairs) {
            cand_pos.assign(K, -1);
            cand.clear();
            for (int k = outlier ? 1 : 0; k < K; ++k)
                if (cl[k].splittable && cl[k].c.N > 0) {
                    cand_pos[k] = (int)cand.size();
                    cand.push_back(k);
                }
            ncand = (int)cand.size();
            const size_t np = (size_t)ncand * (ncand - 1) / 2;
            cand_L.assign(np, std::numeric_limits<double>::quiet_NaN());
            full_scan = !scr || (scr_every > 0 && t % (uint32_t)scr_every == 0);
            pair_feat.clear();
            pair_logit.clear();
            if ((scr || log_pairs) && np > 0) {
                // the screener's features need this launch's cluster log MLs: evaluate the parts first
                if ((rc = run_jobs(jb))) return rc;
                if ((rc = screen_pairs())) return rc;
            }
            size_t p = 0;
            for (int a = 0; a < ncand; ++a)
                for (int b = a + 1; b < ncand; ++b, ++p) {
                    if (!full_scan && !(pair_logit[p] > scr_thr)) continue;
                    const int i = cand[a], j = cand[b];
                    const int32_t m[4] = {2 * i, 2 * i + 1, 2 * j, 2 * j + 1};
                    // the merged cluster keeps cluster i's prior (shared_actions.cpp:90-92)
                    jb.add(m, cl[i].c.prior * 3, &cand_L[p], nullptr);
                    tested.push_back(p);
                }
            n_exact += (int64_t)tested.size();
            n_skipped += (int64_t)(np - tested.size());
            if (scr && full_scan) ++n_full;
        }
        if ((rc = run_jobs(jb))) return rc;
        for (int k = 0; k < K; ++k) {
            if (!which.empty() && !which[k]) continue;
            for (Part* sp : {&cl[k].l, &cl[k].r})
                if (sp->prior == cl[k].c.prior) sp->lml_split = sp->lml;
        }
        host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return DPMM_OK;
    }

    // Features of every candidate pair and, with a screener, their logits.
    int screen_pairs() {
        const int K = (int)cl.size();
        std::vector<double> scal((size_t)K * 7);
        for (int k = 0; k < K; ++k) {
            const Cluster& c = cl[k];
            const double v[7] = {(double)c.c.N, (double)c.l.N, (double)c.r.N, c.c.lml, c.l.lml, c.r.lml, c.c.ld};
            std::copy(v, v + 7, scal.begin() + (size_t)k * 7);
        }
        const size_t np = cand_L.size();
        std::vector<int32_t> pi(np), pj(np);
        size_t p = 0;
        for (int a = 0; a < ncand; ++a)
            for (int b = a + 1; b < ncand; ++b, ++p) {
                pi[p] = cand[a];
                pj[p] = cand[b];
            }
        pair_feat.resize(np * 16);
        int rc = dpmm_pair_features(ctx, K, scal.data(), (int)np, pi.data(), pj.data(), pair_feat.data());
        if (rc) return rc;
        if (scr) {
            pair_logit.resize(np);
            // the forward pass takes at most 2^20 rows per call
            const size_t chunk = (size_t)1 << 20;
            for (size_t p0 = 0; p0 < np && !rc; p0 += chunk)
                rc = dpmm_screener_forward(scr, (int)std::min(chunk, np - p0), pair_feat.data() + p0 * 16,
                                           pair_logit.data() + p0);
        }
        return rc;
    }

    // sample_clusters (local_clusters_actions.cpp:341-371) + sample_cluster_params
    // (shared_actions.cpp:9-55): weights and burn-out history here, the 3K draws on the device.  In
    // outlier mode cluster index 1 is skipped (348-351): its parameters, sub-cluster weights and history
    // stay, it is left out of the mixture's Dirichlet draw, and the weights are shifted behind the
    // outlier cluster's fixed weight outlier_mod (362-370).
    int sample_clusters(uint32_t t) {
        const auto t0 = std::chrono::steady_clock::now();
        const int K = (int)cl.size();
        std::vector<int32_t> mem((size_t)12 * K), ctl((size_t)12 * K);
        std::vector<double> lrw(2 * K);
        bool keep_any = false;
        for (int k = 0; k < K; ++k) {
            Cluster& c = cl[k];
            const Part* ps[3] = {&c.c, &c.l, &c.r};
            if (outlier && k == 1) {
                for (int p = 0; p < 3; ++p) {
                    const ParamRef& r = c.pr[p];
                    int32_t* q = ctl.data() + (size_t)4 * (3 * k + p);
                    q[0] = r.row; q[1] = r.slot; q[2] = r.iter; q[3] = r.prior;
                    std::copy(r.mem, r.mem + 4, mem.begin() + (size_t)4 * (3 * k + p));
                }
                keep_any = true;
            } else {
                const double a[2] = {(double)c.l.N + o.alpha / 2, (double)c.r.N + o.alpha / 2};
                Stream(o.seed, k, t, TAG_LRW).dirichlet(a, 2, c.lrw);
                const double L = c.l.lml + c.r.lml;
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
                for (int p = 0; p < 3; ++p) {
                    std::copy(ps[p]->mem, ps[p]->mem + 4, mem.begin() + (size_t)4 * (3 * k + p));
                    int32_t* q = ctl.data() + (size_t)4 * (3 * k + p);
                    q[0] = q[1] = q[2] = -1;
                    q[3] = ps[p]->prior;
                }
            }
            lrw[2 * k] = c.lrw[0];
            lrw[2 * k + 1] = c.lrw[1];
        }
        // mixture weights ~ Dir(N_1, ..., N_K, alpha), last entry dropped (357-365)
        std::vector<double> a, w;
        for (int k = 0; k < K; ++k)
            if (!(outlier && k == 1)) a.push_back((double)cl[k].c.N);
        a.push_back(o.alpha);
        w.resize(a.size());
        if (a.size() == 2 && a[0] == 0.0) { w[0] = 0.0; w[1] = 1.0; }   // get_dirichlet_distribution (323-339)
        else mixture_dirichlet(o.seed, t, a.data(), (int)a.size(), w.data());
        weights.clear();
        if (outlier) weights.push_back(o.outlier_mod);
        for (size_t i = 0; i + 1 < w.size(); ++i) weights.push_back(w[i] * (1.0 - o.outlier_mod));
        bool any_prior = false;
        for (size_t p = 0; p < ctl.size(); p += 4) any_prior |= ctl[p + 3] != 0;
        const int rc = dpmm_sample_params_ex(ctx, t, K, mem.data(), (keep_any || any_prior) ? ctl.data() : nullptr,
                                             weights.data(), lrw.data(), o.full_mu_cov, o.mult_use_weight);
        // every part's parameters are now row 3k + p of this call's draws
        for (int k = 0; k < K; ++k)
            for (int p = 0; p < 3; ++p) cl[k].pr[p] = keep_row(3 * k + p);
        host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    }

    int relabel(uint32_t t, const std::vector<int32_t>& nl, const std::vector<int32_t>& act) {
        return dpmm_relabel(ctx, t, (int)act.size(), nl.data(), act.data());
    }

    int init() {
        const int K0 = std::max(1, o.init_clusters);
        int rc;
        outlier = o.outlier_mod > 0;
        if (outlier && o.init_from_labels)
            return set_error(DPMM_EINVAL, "outlier mode starts from sampled labels");
        if (o.init_from_labels) {
            std::vector<int32_t> lab(N);
            if ((rc = dpmm_get_labels(ctx, lab.data(), nullptr))) return rc;
            for (int64_t i = 0; i < N; ++i)
                if (lab[i] < 0 || lab[i] >= K0) return set_error(DPMM_EINVAL, "initial label outside 0..init_clusters-1");
            // uniform sub-labels (sample_sub_labels, cudaKernel.cu:1100-1111): the reset relabel of every cluster
            std::vector<int32_t> nl(2 * K0), act(K0, 1);
            for (int q = 0; q < 2 * K0; ++q) nl[q] = q / 2;
            if ((rc = relabel(0, nl, act))) return rc;
        } else if ((rc = dpmm_sample_initial_labels(ctx, K0))) {
            return rc;
        }
        if (outlier) {
            // sample_labels with outlier_mod > 0 (cudaKernel.cu:158-161): the initial labels skip the outlier
            // cluster, 1 .. K0; the outlier cluster (dp_parallel_sampling.cpp:68-72) takes index 0 with its
            // own prior (create_outlier_local_cluster, local_clusters_actions.cpp:48-80)
            std::vector<int32_t> nl(2 * K0), act(K0, 0);
            for (int q = 0; q < 2 * K0; ++q) nl[q] = q / 2 + 1;
            if ((rc = relabel(0, nl, act))) return rc;
            cl.assign(K0 + 1, Cluster());
            cl[0].c.prior = cl[0].l.prior = cl[0].r.prior = 1;
            // cluster 1 is never redrawn by sample_clusters: it keeps the parameters drawn from the prior when
            // it was created, the sub-clusters' the same ones (create_first_local_cluster, 14-46)
            const int32_t none[4] = {-1, -1, -1, -1};
            for (int p = 0; p < 3; ++p) cl[1].pr[p] = draw_from(none, 0, 3, 0);
        } else {
            cl.assign(K0, Cluster());
        }
        for (auto& c : cl) reset_hist(c);
        rc = update_stats({}, false);
        if (rc) return rc;
        return sample_clusters(0);
    }

    int step(int t) {
        if (device_step_eligible()) return step_device(t);
        if (dev_on) {   // back to host decisions (a screener was set): the device table into the host list
            int rc = pull_from_device();
            if (rc) return rc;
            dev_on = false;
        }
        const bool fin = t >= o.iters - o.argmax_sample_stop;
        // run_model's flags (dp_parallel_sampling.cpp:143-150); the device's cluster cap ends splitting too
        const bool no_more_splits =
            t >= o.iters - o.split_stop || (double)cl.size() >= o.max_clusters || (int)cl.size() >= kcap;
        const uint32_t it = (uint32_t)t;
        int rc;
        host_ms = 0;
        if (t > 0) {   // group_step resamples parameters first (local_clusters_actions.cpp:462)
            rc = sample_clusters(it);
            if (rc) return rc;
        }
        // label sweep, sub-label sweep and statistics in one device call (one hipGraph in graph mode)
        Nsub.resize(2 * cl.size());
        rc = dpmm_sweep_iteration(ctx, it, (fin || o.hard_clustering) ? 1 : 0, Nsub.data());
        if (rc) return rc;
        counts_fresh = true;
        if ((trace_labels || trace_subs) && (rc = dpmm_get_labels(ctx, trace_labels, trace_subs))) return rc;
        // one log-marginal launch: every part, and every pair of merge candidates (a cluster that is
        // reset or split below leaves the candidates, and the others keep these statistics)
        rc = update_stats({}, !no_more_splits, it);
        if (rc) return rc;

        // reset_bad_clusters (local_clusters_actions.cpp:429-447)
        int K = (int)cl.size();
        {
            std::vector<char> bad(K, 0);
            bool any = false;
            for (int k = 0; k < K; ++k)
                if (cl[k].l.N == 0 || cl[k].r.N == 0) {
                    bad[k] = 1;
                    any = true;
                    reset_hist(cl[k]);
                    cl[k].splittable = false;
                }
            if (any) {
                for (int k = 0; k < K; ++k) ev_reset += bad[k];
                std::vector<int32_t> nl(2 * K), act(K, 0);
                for (int q = 0; q < 2 * K; ++q) nl[q] = q / 2;
                for (int k = 0; k < K; ++k) act[k] = bad[k] ? 1 : 0;
                if ((rc = relabel(it, nl, act))) return rc;
                if ((rc = update_stats(bad, false))) return rc;
            }
        }

        std::vector<int> merge_into;   // per cluster: absorbing cluster or -1
        if (!no_more_splits) {
            // check_and_split (local_clusters_actions.cpp:244-285), should_split_local (215-242)
            const auto t0 = std::chrono::steady_clock::now();
            std::vector<char> acc(K, 0);
            for (int k = 0; k < K; ++k) {
                const Cluster& c = cl[k];
                if (outlier && k == 1) continue;   // local_clusters_actions.cpp:250-253
                if (!(c.splittable && c.c.N > 1) || fin || c.l.N == 0 || c.r.N == 0) continue;
                const double lhr = std::log(o.alpha) + std::lgamma((double)c.l.N) + c.l.lml_split +
                                   std::lgamma((double)c.r.N) + c.r.lml_split -
                                   (std::lgamma((double)c.c.N) + c.c.lml);
                acc[k] = lhr > std::log(Stream(o.seed, k, it, TAG_SPLIT_U).uniform());
            }
            host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::vector<int32_t> nl(2 * K), act(K, 0);
            for (int q = 0; q < 2 * K; ++q) nl[q] = q / 2;
            int nsplit = 0;
            // accepted splits beyond the device's cluster cap are dropped, the highest indices first
            int room = kcap - K;
            for (int k = 0; k < K; ++k) {
                if (acc[k] && room <= 0) acc[k] = 0;
                else if (acc[k]) --room;
            }
            for (int k = 0; k < K; ++k) {
                if (!acc[k]) continue;
                const int nk = (int)cl.size();
                // split_cluster_local (183-194): k keeps the left half, nk gets the right half; each half is
                // create_splittable_from_params of the old sub-cluster (shared_actions.cpp:57-70): its
                // parameters, fresh sub-cluster parameters drawn from its posterior (after the relabel below,
                // the statistics of the half's own entries), sub-cluster weights ~ Dir(alpha / 2, alpha / 2).
                // These are what the sampler uses only where sample_clusters does not redraw them (cluster
                // index 1 in outlier mode).
                Cluster left, right;
                left.c = cl[k].l;
                right.c = cl[k].r;
                left.l = left.r = left.c;
                right.l = right.r = right.c;
                reset_hist(left);
                reset_hist(right);
                left.pr[0] = cl[k].pr[1];
                right.pr[0] = cl[k].pr[2];
                const int32_t ml[4] = {2 * k, 2 * k + 1, -1, -1}, mr[4] = {2 * nk, 2 * nk + 1, -1, -1};
                for (int p = 1; p < 3; ++p) {
                    left.pr[p] = draw_from(ml, left.c.prior, 3 * k + p, (int)it);
                    right.pr[p] = draw_from(mr, right.c.prior, 3 * nk + p, (int)it);
                }
                const double half[2] = {o.alpha / 2, o.alpha / 2};
                Stream(o.seed, k, it, TAG_SPLIT_LRW).dirichlet(half, 2, left.lrw);
                Stream(o.seed, nk, it, TAG_SPLIT_LRW).dirichlet(half, 2, right.lrw);
                cl[k] = left;
                cl.push_back(right);
                nl[2 * k + 1] = nk;
                act[k] = 2;
                ++nsplit;
            }
            ev_split += nsplit;
            if (nsplit) {
                if ((rc = relabel(it, nl, act))) return rc;
                std::vector<char> touched(cl.size(), 0);
                for (int k = 0; k < K; ++k) touched[k] = acc[k];
                for (size_t k = K; k < cl.size(); ++k) touched[k] = 1;
                if ((rc = update_stats(touched, false))) return rc;
            }

            // check_and_merge (287-321) + should_merge (shared_actions.cpp:88-111): every
            // eligible pair is tested on pre-merge statistics, then accepted greedily in
            // lexicographic order; a merged cluster leaves the eligible set, which makes the
            // outcome identical to the reference's sequential loop.
            const auto t1 = std::chrono::steady_clock::now();
            K = (int)cl.size();
            const double al = o.alpha;
            const double cst = -std::log(al) + std::lgamma(al) - 2 * std::lgamma(0.5 * al);
            merge_into.assign(K, -1);
            std::vector<char> used(K, 0);
            const bool logging = log_pairs && !pair_feat.empty();
            log_feat.clear();
            log_h.clear();
            log_u.clear();
            // per-cluster terms of the ratio, computed once (the same values as inline calls)
            std::vector<double> lg_n(K), lg_nh(K);
            for (int k = 0; k < K; ++k) {
                const double Nk = (double)cl[k].c.N;
                lg_n[k] = cl[k].c.N > 0 ? std::lgamma(Nk) : 0.0;
                lg_nh[k] = std::lgamma(Nk + 0.5 * al);
            }
            size_t p = 0;
            for (int a = 0; a < ncand; ++a)
                for (int b = a + 1; b < ncand; ++b, ++p) {
                    const int i = cand[a], j = cand[b];
                    const double L = cand_L[p];
                    if (!std::isfinite(L)) continue;   // skipped by the screener
                    // a cluster split or reset in this iteration is no longer a candidate
                    const bool elig = cl[i].splittable && cl[j].splittable && cl[i].c.N > 0 && cl[j].c.N > 0;
                    if (!elig || ((used[i] || used[j]) && !logging)) continue;
                    if (!logging) {
                        // lgamma(Nm) - lgamma(Nm + alpha) < 0 for Nm >= 2, so without it the ratio is an
                        // upper bound; far below log(2^-32) (the smallest uniform) and log(0.1) the pair
                        // is rejected whatever its uniform, which is then not drawn (keyed streams)
                        const double ub = cst + lg_nh[i] - lg_n[i] - lg_n[j] + lg_nh[j] + L - cl[i].c.lml - cl[j].c.lml;
                        if (ub < -100.0) continue;
                    }
                    const double Ni = (double)cl[i].c.N, Nj = (double)cl[j].c.N, Nm = Ni + Nj;
                    const double lhr = cst + std::lgamma(Nm) - std::lgamma(Nm + al) + lg_nh[i] - lg_n[i] - lg_n[j] +
                                       lg_nh[j] + L - cl[i].c.lml - cl[j].c.lml;
                    const double u = Stream(o.seed, ((uint32_t)i << 16) | (uint32_t)j, it, TAG_MERGE_U).uniform();
                    if (logging) {
                        log_feat.insert(log_feat.end(), pair_feat.begin() + p * 16, pair_feat.begin() + p * 16 + 16);
                        log_h.push_back(lhr);
                        log_u.push_back(std::log(u));
                    }
                    if (used[i] || used[j]) continue;
                    if (!((lhr > std::log(u)) || (fin && lhr > std::log(0.1)))) continue;
                    used[i] = used[j] = 1;
                    merge_into[j] = i;
                    if (scr && full_scan && !(pair_logit[p] > scr_thr)) ++n_missed;
                    // merge_clusters (204-213) + merge_clusters_to_splittable (shared_actions.cpp:72-86)
                    Cluster& A = cl[i];
                    Cluster& Bc = cl[j];
                    Part merged;
                    merged.N = A.c.N + Bc.c.N;
                    merged.mem[0] = 2 * i; merged.mem[1] = 2 * i + 1;
                    merged.mem[2] = 2 * j; merged.mem[3] = 2 * j + 1;
                    merged.lml = merged.lml_split = L;
                    merged.prior = A.c.prior;
                    // sub-cluster weights ~ Dir(N_i + alpha / 2, N_j + alpha / 2), the left sub-cluster a clone
                    // of the merged cluster (its parameters: cluster i's), the right one cluster j's
                    const double mw[2] = {(double)A.c.N + 0.5 * al, (double)Bc.c.N + 0.5 * al};
                    Stream(o.seed, i, it, TAG_MERGE_LRW).dirichlet(mw, 2, A.lrw);
                    A.pr[1] = A.pr[0];
                    A.pr[2] = Bc.pr[0];
                    A.r = Bc.c;
                    A.c = merged;
                    A.l = merged;   // the reference clones the merged parameters into the left sub-cluster
                    A.splittable = false;
                    reset_hist(A);
                    Bc.c.N = 0;
                    Bc.splittable = false;
                }
            host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        }

        // merges + remove_empty_clusters (local_clusters_actions.cpp:384-419) in one relabel pass
        K = (int)cl.size();
        std::vector<int32_t> remap(K, -1);
        int nk = 0;
        // outlier mode keeps the outlier cluster, and cluster 1 while there are two (408-411)
        for (int k = 0; k < K; ++k)
            if (cl[k].c.N > 0 || (outlier && (k == 0 || (k == 1 && K == 2)))) remap[k] = nk++;
        bool any_merge = false;
        for (int k = 0; k < (int)merge_into.size(); ++k) {
            any_merge |= merge_into[k] >= 0;
            ev_merge += merge_into[k] >= 0;
        }
        ev_removed += K - nk;
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
    if (opt->prior != prior || opt->iters < 0 || opt->burnout < 1 || !(opt->alpha > 0))
        return dpmm::set_error(DPMM_EINVAL, "bad sampler options");
    if (!(opt->outlier_mod >= 0.0 && opt->outlier_mod < 1.0))
        return dpmm::set_error(DPMM_EINVAL, "outlier_mod must lie in [0, 1)");
    if (opt->outlier_mod > 0.0 && !dpmm::ctx_has_outlier_prior(ctx))
        return dpmm::set_error(DPMM_ESTATE, "outlier mode needs dpmm_set_outlier_prior first");
    int rc = dpmm_set_prior(ctx, k0, m0, v0, psi0, alpha0);
    if (rc) return rc;
    if (opt->precision != prec) {
        rc = dpmm_set_precision(ctx, opt->precision);
        if (rc) return rc;
    }
    auto* S = new dpmm_sampler();
    dpmm::SamplerImpl& s = S->s;
    s.ctx = ctx;
    s.o = *opt;
    s.N = N;
    s.D = D;
    s.prior = prior;
    if ((rc = dpmm_max_clusters(ctx, &s.kcap))) { delete S; return rc; }
    rc = s.init();
    if (rc) { delete S; return rc; }
    *out = S;
    return DPMM_OK;
}

int dpmm_sampler_step(dpmm_sampler* S, int iter) {
    if (!S || iter < 1) return DPMM_EINVAL;
    return S->s.step(iter);
}

int dpmm_sampler_num_clusters(dpmm_sampler* S) { return S ? S->s.num_clusters() : DPMM_EINVAL; }

int dpmm_sampler_weights(dpmm_sampler* S, double* w, int64_t* counts, int max_k) {
    if (!S) return DPMM_EINVAL;
    int rc = S->s.pull_from_device();
    if (rc) return rc;
    const int K = (int)S->s.cl.size();
    for (int k = 0; k < K && k < max_k; ++k) {
        if (w) w[k] = k < (int)S->s.weights.size() ? S->s.weights[k] : 0.0;
        if (counts) counts[k] = S->s.cl[k].c.N;
    }
    return DPMM_OK;
}

// dp_parallel_sampling_class::calculate_posterior (dp_parallel_sampling.cpp:218-236):
// lgamma(alpha) - lgamma(N + alpha) + sum over non-empty clusters of
// log ML_k + log alpha + lgamma(N_k), from the cluster scalars the last step left (the log ML of a
// cluster split or merged in the step is its part's, as the reference keeps it).
int dpmm_sampler_log_posterior(dpmm_sampler* S, double* lp) {
    if (!S || !lp) return DPMM_EINVAL;
    int rc = S->s.pull_from_device();
    if (rc) return rc;
    const dpmm::SamplerImpl& s = S->s;
    const double a = s.o.alpha;
    double v = std::lgamma(a) - std::lgamma((double)s.N + a);
    for (const auto& c : s.cl) {
        if (c.c.N == 0) continue;
        v += c.c.lml + std::log(a) + std::lgamma((double)c.c.N);
    }
    *lp = v;
    return DPMM_OK;
}

int dpmm_sampler_trace_sweep(dpmm_sampler* S, int32_t* labels, int32_t* sub_labels) {
    if (!S) return DPMM_EINVAL;
    S->s.trace_labels = labels;
    S->s.trace_subs = sub_labels;
    return DPMM_OK;
}

int dpmm_sampler_host_ms(dpmm_sampler* S, double* ms) {
    if (!S || !ms) return DPMM_EINVAL;
    *ms = S->s.host_ms;
    return DPMM_OK;
}

int dpmm_sampler_set_screener(dpmm_sampler* S, dpmm_screener* scr, int every, float threshold, int log_pairs) {
    if (!S || every < 0) return dpmm::set_error(DPMM_EINVAL, "bad args");
    if ((scr || log_pairs) && S->s.prior != DPMM_PRIOR_GAUSSIAN)
        return dpmm::set_error(DPMM_EINVAL, "merge screening needs the Gaussian prior");
    S->s.scr = scr;
    S->s.scr_every = every;
    S->s.scr_thr = threshold;
    S->s.log_pairs = log_pairs != 0;
    return DPMM_OK;
}

int dpmm_sampler_screen_stats(dpmm_sampler* S, int64_t* exact, int64_t* skipped, int64_t* missed, int64_t* full) {
    if (!S) return DPMM_EINVAL;
    if (exact) *exact = S->s.n_exact;
    if (skipped) *skipped = S->s.n_skipped;
    if (missed) *missed = S->s.n_missed;
    if (full) *full = S->s.n_full;
    return DPMM_OK;
}

int dpmm_sampler_set_merge_bound(dpmm_sampler* S, int enable) {
    if (!S) return DPMM_EINVAL;
    S->s.merge_bound = enable != 0;
    return DPMM_OK;
}

int dpmm_sampler_merge_bound_rejects(dpmm_sampler* S, int64_t* rejected) {
    if (!S || !rejected) return DPMM_EINVAL;
    *rejected = S->s.n_bound;
    return DPMM_OK;
}

int dpmm_sampler_pair_log(dpmm_sampler* S, int max_rows, float* feat, double* log_h, double* log_u, int* n_rows) {
    if (!S || !n_rows || max_rows < 0) return DPMM_EINVAL;
    const int n = (int)std::min<size_t>(S->s.log_h.size(), (size_t)max_rows);
    for (int r = 0; r < n; ++r) {
        if (feat) std::copy(S->s.log_feat.begin() + (size_t)r * 16, S->s.log_feat.begin() + (size_t)r * 16 + 16,
                            feat + (size_t)r * 16);
        if (log_h) log_h[r] = S->s.log_h[r];
        if (log_u) log_u[r] = S->s.log_u[r];
    }
    *n_rows = (int)S->s.log_h.size();
    return DPMM_OK;
}

int dpmm_sampler_events(dpmm_sampler* S, int64_t* splits, int64_t* merges, int64_t* resets, int64_t* removed) {
    if (!S) return DPMM_EINVAL;
    if (splits) *splits = S->s.ev_split;
    if (merges) *merges = S->s.ev_merge;
    if (resets) *resets = S->s.ev_reset;
    if (removed) *removed = S->s.ev_removed;
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
