// C++ drop-in for the reference's public surface over the C ABI of dpmm_hip.h (header only, C++17):
//   prior_type                      global_params.h:14-18
//   hyperparams, niw_hyperparams    hyperparams.h, priors/niw_hyperparams.h:13, 40-43 (psi is Psi / v)
//   multinomial_hyper               priors/multinomial_hyper.h:13, 29
//   ModelInfo                       modelInfo.h:44-56 (labels are 1-based, as in the reference)
//   dp_parallel_sampling_class      dp_parallel_sampling.h:11-44: the (numLabels, all_data, seed,
//                                   prior) constructor and dp_parallel() with the reference's
//                                   parameters and defaults.
// all_data is D x N column-major (each point contiguous): any matrix type with data(), rows() (= D)
// and cols() (= N), e.g. Eigen::MatrixXd, or a raw pointer with D.  Gaussian data of D <= 64 is
// swept in FP64 like the reference; larger D, or set_precision(DPMM_PREC_FP32 / DPMM_PREC_MIXED),
// sweeps an FP32 copy.  Multinomial data are counts, passed to the device as int32.
//
// The types are declared in namespace dpmm and, unless DPMM_NO_GLOBAL_NAMES is defined, also in the
// global namespace, where the reference declares them, so its callers compile unchanged.
//
// Differences from the reference, each reported by an exception rather than ignored: save_model is not
// built; draw_labels is accepted and ignored (the reference plots with OpenCV).  ModelInfo::dp_model
// carries the model's mixture weights (dp_model->group.weights, as main.cpp:76-80 reads them), not the
// whole host-side model.
#pragma once
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../dpmm_hip.h"
#include "io.hpp"

namespace dpmm {

enum prior_type { Gaussian = DPMM_PRIOR_GAUSSIAN, Multinomial = DPMM_PRIOR_MULTINOMIAL };
typedef int LabelType;
typedef std::vector<LabelType> LabelsType;

struct hyperparams {
    virtual ~hyperparams() = default;
};

struct niw_hyperparams : hyperparams {
    double k = 0, v = 0;
    std::vector<double> m, psi;   // psi row-major D x D (symmetric)
    niw_hyperparams(double k_, std::vector<double> m_, double v_, std::vector<double> psi_)
        : k(k_), v(v_), m(std::move(m_)), psi(std::move(psi_)) {
        if (psi.size() != m.size() * m.size()) throw std::invalid_argument("niw_hyperparams: psi must be D x D");
    }
    // Eigen-style vector / matrix (data(), size(); data(), rows(), cols())
    template <class V, class M, class = decltype(std::declval<const V&>().data()),
              class = decltype(std::declval<const M&>().rows())>
    niw_hyperparams(double k_, const V& m_, double v_, const M& psi_) : k(k_), v(v_) {
        m.assign(m_.data(), m_.data() + m_.size());
        psi.assign(psi_.data(), psi_.data() + (size_t)psi_.rows() * psi_.cols());
        if (psi.size() != m.size() * m.size()) throw std::invalid_argument("niw_hyperparams: psi must be D x D");
    }
};

struct multinomial_hyper : hyperparams {
    std::vector<double> alpha;
    explicit multinomial_hyper(std::vector<double> a) : alpha(std::move(a)) {}
    template <class V, class = decltype(std::declval<const V&>().data())>
    explicit multinomial_hyper(const V& a) : alpha(a.data(), a.data() + a.size()) {}
};

// The parts of the reference's model (ds.h:330-378) that its callers read after a run.
struct local_group {
    std::vector<double> weights;   // mixture weights of the final clusters (ds.h:339)
};
struct dp_parallel_sampling {
    local_group group;
};

struct ModelInfo {
    std::shared_ptr<dp_parallel_sampling> dp_model;   // dp_model->group.weights (main.cpp:76-80)
    std::vector<double> iter_count;           // seconds per iteration
    std::vector<double> nmi_score_history;    // per iteration in verbose mode when gt is given
    std::vector<double> likelihood_history;   // log posterior per iteration in verbose mode, else 1
    std::vector<int> cluster_count_history;
    std::shared_ptr<LabelsType> labels;       // 1-based
    std::vector<double> weights;              // the same mixture weights as dp_model->group.weights
};

class dpmm_error : public std::runtime_error {
  public:
    dpmm_error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc) {
    if (rc != DPMM_OK) throw dpmm_error(rc, std::string("dpmm: ") + dpmm_last_error());
}

// Normalised mutual information 2 I(a;b) / (H(a) + H(b)), the reference's verbose NMI
// (dp_parallel_sampling.cpp:170-178).
inline double nmi(const LabelsType& a, const LabelsType& b) {
    const size_t n = a.size();
    if (n == 0 || b.size() != n) return 0.0;
    std::unordered_map<long long, double> joint;
    std::unordered_map<int, double> pa, pb;
    for (size_t i = 0; i < n; ++i) {
        joint[((long long)a[i] << 32) ^ (unsigned)b[i]] += 1.0;
        pa[a[i]] += 1.0;
        pb[b[i]] += 1.0;
    }
    double ha = 0, hb = 0, mi = 0;
    for (auto& e : pa) ha -= e.second / n * std::log(e.second / n);
    for (auto& e : pb) hb -= e.second / n * std::log(e.second / n);
    for (auto& e : joint) {
        const int x = (int)(e.first >> 32), y = (int)(unsigned)(e.first & 0xffffffffu);
        const double p = e.second / n;
        mi += p * std::log(p / (pa[x] / n * pb[y] / n));
    }
    return (ha + hb > 0) ? 2.0 * mi / (ha + hb) : 1.0;
}

class dp_parallel_sampling_class {
  public:
    template <class Matrix, class = decltype(std::declval<Matrix&>().rows())>
    dp_parallel_sampling_class(int numLabels, Matrix& all_data, unsigned long long randomSeed, prior_type priorType)
        : dp_parallel_sampling_class(numLabels, all_data.data(), (int)all_data.rows(), randomSeed, priorType) {
        if ((long long)all_data.cols() != (long long)numLabels)
            throw std::invalid_argument("all_data must be D x numLabels");
    }

    dp_parallel_sampling_class(int numLabels, const double* all_data, int D, unsigned long long randomSeed,
                               prior_type priorType)
        : N_(numLabels), D_(D), seed_(randomSeed), prior_(priorType) {
        if (!all_data || numLabels <= 0 || D <= 0) throw std::invalid_argument("empty data");
        const size_t n = (size_t)numLabels * D;
        if (prior_ == Multinomial) {
            counts_.resize(n);
            for (size_t i = 0; i < n; ++i) counts_[i] = (int32_t)std::lround(all_data[i]);
        } else {
            data64_.assign(all_data, all_data + n);
        }
        precision_ = (prior_ == Gaussian && D_ <= 64) ? DPMM_PREC_FP64 : DPMM_PREC_FP32;
    }

    // The model-file constructor (dp_parallel_sampling.h:15; dp_parallel_sampling.cpp:25-31): points from a
    // .npy file of N x D values (".npy" appended when missing; read as D x N, utils.cpp:12-57) and the run's
    // settings from a JSON params file with global_params::init's keys and defaults (global_params.cpp:8-70,
    // global_params.h:36-78): alpha 10, iterations 100, initial_clusters 1, use_verbose false,
    // burnout_period 20, max_num_of_clusters unbounded, outlier_mod 0.05, hyper_params and
    // outlier_hyper_params {"k", "m", "v", "psi"} or {"alpha"} (default niw(1, 0, 5, I), or alpha = 0 for
    // the multinomial prior, which the sampler rejects), randomSeed (0: seeded from the clock), gt.
    dp_parallel_sampling_class(const std::string& modelDataFileName, const std::string& modelParamsFileName,
                               prior_type priorType)
        : prior_(priorType) {
        std::string path = modelDataFileName;
        if (path.find(".npy") == std::string::npos) path += ".npy";
        const NpyArray a = load_npy(path);
        N_ = (int)a.rows;
        D_ = (int)a.cols;
        if (N_ <= 0 || D_ <= 0) throw std::invalid_argument("dpmm: empty data file");
        if (prior_ == Multinomial) {
            counts_.resize(a.data.size());
            for (size_t i = 0; i < a.data.size(); ++i) counts_[i] = (int32_t)std::lround(a.data[i]);
        } else {
            data64_ = a.data;   // point i at i * D: the reference's D x N column-major layout
        }
        precision_ = (prior_ == Gaussian && D_ <= 64) ? DPMM_PREC_FP64 : DPMM_PREC_FP32;
        const json::Value root = json::parse_file(modelParamsFileName);
        file_.alpha = root.get("alpha", 10.0);
        file_.iterations = (int)root.get("iterations", 100.0);
        file_.initial_clusters = (int)root.get("initial_clusters", 1.0);
        file_.verbose = root.get("use_verbose", false);
        file_.draw_labels = root.get("draw_labels", false);
        file_.save_model = root.get("should_save_model", false);
        file_.burnout = (int)root.get("burnout_period", 20.0);
        file_.max_clusters = root.get("max_num_of_clusters", DBL_MAX);
        file_.outlier_mod = root.get("outlier_mod", 0.05);
        file_.hyper = hyper_from_json(root["hyper_params"]);
        file_.outlier_hyper = hyper_from_json(root["outlier_hyper_params"]);
        const double s = root.get("randomSeed", 0.0);
        seed_ = s != 0.0 ? (unsigned long long)s
                         : (unsigned long long)std::chrono::system_clock::now().time_since_epoch().count();
        if (root.has("gt")) {
            file_.gt = std::make_shared<LabelsType>();
            for (double v : root["gt"].numbers()) file_.gt->push_back((LabelType)v);
        }
        from_file_ = true;
    }

    // dp_parallel_from_file (dp_parallel_sampling.h:31): dp_parallel with the params file's settings.
    ModelInfo dp_parallel_from_file() {
        if (!from_file_) throw std::logic_error("dpmm: dp_parallel_from_file needs the model-file constructor");
        // checkpoints are written every 1000 iterations (global_params.h:58); shorter runs write none
        const bool save = file_.save_model && file_.iterations >= 1000;
        return dp_parallel(file_.hyper, file_.alpha, file_.iterations, file_.initial_clusters, file_.verbose,
                           file_.draw_labels, save, file_.burnout, file_.gt, file_.max_clusters, file_.outlier_mod,
                           file_.outlier_hyper);
    }
    int num_points() const { return N_; }
    int dimension() const { return D_; }

    dp_parallel_sampling_class(const dp_parallel_sampling_class&) = delete;
    dp_parallel_sampling_class& operator=(const dp_parallel_sampling_class&) = delete;
    ~dp_parallel_sampling_class() = default;

    // DPMM_PREC_FP64 (Gaussian, D <= 64), DPMM_PREC_FP32 or DPMM_PREC_MIXED for the label sweep.
    void set_precision(int p) { precision_ = p; }
    // global_params::hard_clustering (global_params.h:38): argmax labels in every sweep.
    void set_hard_clustering(bool h) { hard_clustering_ = h; }

    ModelInfo dp_parallel(std::shared_ptr<hyperparams>& local_hyper_params, double alpha_param, int iters = 100,
                          int init_clusters = 1, bool verbose = true, bool draw_labels = false,
                          bool save_model = false, int burnout = 15, std::shared_ptr<LabelsType> gt = nullptr,
                          double max_clusters = DBL_MAX, double outlier_weight = 0,
                          std::shared_ptr<hyperparams> outlier_params = nullptr) {
        (void)draw_labels;
        if (outlier_weight > 0 && !outlier_params)
            throw std::invalid_argument("dpmm: outlier_weight > 0 needs outlier_params");
        if (save_model) throw std::invalid_argument("dpmm: save_model is not built");
        dpmm_ctx* ctx = nullptr;
        if (prior_ == Multinomial) {
            check(dpmm_create(counts_.data(), N_, D_, DPMM_DTYPE_I32, DPMM_PRIOR_MULTINOMIAL, seed_, &ctx));
        } else if (precision_ == DPMM_PREC_FP64) {
            check(dpmm_create(data64_.data(), N_, D_, DPMM_DTYPE_F64, DPMM_PRIOR_GAUSSIAN, seed_, &ctx));
        } else {
            std::vector<float> f(data64_.begin(), data64_.end());
            check(dpmm_create(f.data(), N_, D_, DPMM_DTYPE_F32, DPMM_PRIOR_GAUSSIAN, seed_, &ctx));
        }
        std::unique_ptr<dpmm_ctx, int (*)(dpmm_ctx*)> ctx_guard(ctx, dpmm_destroy);
        dpmm_options o{};
        o.prior = prior_;
        o.alpha = alpha_param;
        o.iters = iters;
        o.init_clusters = init_clusters;
        o.burnout = burnout;
        o.argmax_sample_stop = 5;
        o.split_stop = 5;
        o.max_clusters = max_clusters;
        o.seed = seed_;
        o.full_mu_cov = 0;
        o.mult_use_weight = 0;
        o.precision = precision_;
        o.verbose = 0;
        o.hard_clustering = hard_clustering_ ? 1 : 0;
        o.outlier_mod = outlier_weight > 0 ? outlier_weight : 0.0;
        if (outlier_weight > 0) {   // the outlier cluster's own prior (local_clusters_actions.cpp:48-61)
            if (prior_ == Gaussian) {
                auto* h = dynamic_cast<niw_hyperparams*>(outlier_params.get());
                if (!h || (int)h->m.size() != D_)
                    throw std::invalid_argument("dpmm: outlier_params must be niw_hyperparams of dimension D");
                check(dpmm_set_outlier_prior(ctx, h->k, h->m.data(), h->v, h->psi.data(), nullptr));
            } else {
                auto* h = dynamic_cast<multinomial_hyper*>(outlier_params.get());
                if (!h || (int)h->alpha.size() != D_)
                    throw std::invalid_argument("dpmm: outlier_params must be multinomial_hyper of dimension D");
                check(dpmm_set_outlier_prior(ctx, 0.0, nullptr, 0.0, nullptr, h->alpha.data()));
            }
        }
        dpmm_sampler* s = nullptr;
        if (prior_ == Gaussian) {
            auto* h = dynamic_cast<niw_hyperparams*>(local_hyper_params.get());
            if (!h || (int)h->m.size() != D_) throw std::invalid_argument("dpmm: Gaussian prior needs niw_hyperparams of dimension D");
            check(dpmm_sampler_create(ctx, &o, h->k, h->m.data(), h->v, h->psi.data(), nullptr, &s));
        } else {
            auto* h = dynamic_cast<multinomial_hyper*>(local_hyper_params.get());
            if (!h || (int)h->alpha.size() != D_) throw std::invalid_argument("dpmm: multinomial prior needs multinomial_hyper of dimension D");
            check(dpmm_sampler_create(ctx, &o, 0.0, nullptr, 0.0, nullptr, h->alpha.data(), &s));
        }
        std::unique_ptr<dpmm_sampler, int (*)(dpmm_sampler*)> s_guard(s, dpmm_sampler_destroy);
        ModelInfo info;
        LabelsType lab(N_);
        if (gt) {   // classes 0 .. C-1 in order of first appearance, for the device contingency table
            if ((int64_t)gt->size() != N_) throw std::invalid_argument("dpmm: gt needs one label per point");
            std::unordered_map<int, int> cls;
            std::vector<int32_t> c(N_);
            for (int64_t i = 0; i < N_; ++i) {
                auto it = cls.emplace((*gt)[i], (int)cls.size()).first;
                c[i] = it->second;
            }
            check(dpmm_set_ground_truth(ctx, c.data(), (int)cls.size()));
        }
        double total = 0.0;
        for (int t = 1; t <= iters; ++t) {
            const auto t0 = std::chrono::steady_clock::now();
            check(dpmm_sampler_step(s, t));
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const int K = dpmm_sampler_num_clusters(s);
            total += dt;
            info.iter_count.push_back(dt);
            info.cluster_count_history.push_back(K);
            double lp = 1.0;   // the reference records 1 when not verbose (dp_parallel_sampling.cpp:192)
            if (verbose) check(dpmm_sampler_log_posterior(s, &lp));
            info.likelihood_history.push_back(lp);
            double score = -1.0;
            if (verbose && gt) {   // the label / class table is built on the device; only K x C counts come back
                check(dpmm_label_nmi(ctx, K, &score));
                info.nmi_score_history.push_back(score);
            }
            if (verbose)
                std::printf("Iteration: %d || Clusters count: %d || Log posterior: %f || NMI score: %f || "
                            "Iter Time: %f || Total time: %f\n", t, K, lp, score, dt, total);
        }
        check(dpmm_get_labels(ctx, lab.data(), nullptr));
        for (auto& l : lab) l += 1;   // the reference's labels are 1-based
        info.labels = std::make_shared<LabelsType>(std::move(lab));
        const int K = dpmm_sampler_num_clusters(s);
        info.weights.resize(K);
        check(dpmm_sampler_weights(s, info.weights.data(), nullptr, K));
        info.dp_model = std::make_shared<dp_parallel_sampling>();
        info.dp_model->group.weights = info.weights;
        return info;
    }

  private:
    // niw::create_hyperparams / multinomial_prior::create_hyperparams (niw.cpp:155-198,
    // multinomial_prior.cpp:80-102): from the JSON object, or the defaults when it is absent.
    std::shared_ptr<hyperparams> hyper_from_json(const json::Value& v) const {
        if (prior_ == Gaussian) {
            if (v.is_null())
                return std::make_shared<niw_hyperparams>(1.0, std::vector<double>(D_, 0.0), 5.0, identity(D_));
            std::vector<double> m = v["m"].numbers(), psi;
            for (const auto& row : v["psi"].arr)
                for (double x : row.numbers()) psi.push_back(x);
            return std::make_shared<niw_hyperparams>(v["k"].as_double(), m, v["v"].as_double(), psi);
        }
        if (v.is_null()) return std::make_shared<multinomial_hyper>(std::vector<double>(D_, 0.0));
        return std::make_shared<multinomial_hyper>(v["alpha"].numbers());
    }
    static std::vector<double> identity(int D) {
        std::vector<double> I((size_t)D * D, 0.0);
        for (int i = 0; i < D; ++i) I[(size_t)i * D + i] = 1.0;
        return I;
    }
    struct FileParams {
        double alpha = 10.0, max_clusters = DBL_MAX, outlier_mod = 0.05;
        int iterations = 100, initial_clusters = 1, burnout = 20;
        bool verbose = false, draw_labels = false, save_model = false;
        std::shared_ptr<hyperparams> hyper, outlier_hyper;
        std::shared_ptr<LabelsType> gt;
    } file_;
    bool from_file_ = false;
    int N_, D_;
    unsigned long long seed_;
    prior_type prior_;
    int precision_;
    bool hard_clustering_ = false;
    std::vector<double> data64_;
    std::vector<int32_t> counts_;
};

}  // namespace dpmm

#ifndef DPMM_NO_GLOBAL_NAMES
using dpmm::dp_parallel_sampling;
using dpmm::dp_parallel_sampling_class;
using dpmm::hyperparams;
using dpmm::LabelsType;
using dpmm::LabelType;
using dpmm::local_group;
using dpmm::ModelInfo;
using dpmm::multinomial_hyper;
using dpmm::niw_hyperparams;
using dpmm::prior_type;
#endif
