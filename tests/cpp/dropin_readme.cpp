// The reference's README example through the C++ drop-in header (include/dpmm/dp_parallel_sampling.hpp):
// a 2-D mixture of two Gaussians, niw_hyperparams(1, 0, 5, I), alpha = 1, 400 iterations (the split of
// the single initial cluster waits for its sub-clusters to drift apart: 50 to 150 iterations or more,
// depending on the data); then a
// small multinomial mixture.  Prints one line per run: prior, K, NMI against the truth, label range,
// weight sum, first iteration with the final K.
#include <cstdio>
#include <random>
#include <vector>

#include "dpmm/dp_parallel_sampling.hpp"

// The reference's names, in the global namespace as its callers use them (no dpmm:: prefix).

static void gaussian_run() {
    const int N = 20000, D = 2, K = 2;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-100, 100);
    std::normal_distribution<double> Z(0, 1);
    double mu[K][D], L[K][D][D];
    for (int k = 0; k < K; ++k) {
        double A[D][D], S[D][D];
        for (int i = 0; i < D; ++i) { mu[k][i] = U(rng); for (int j = 0; j < D; ++j) A[i][j] = Z(rng); }
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) {
                S[i][j] = (i == j) ? 1.0 : 0.0;
                for (int t = 0; t < D; ++t) S[i][j] += A[i][t] * A[j][t];
            }
        L[k][0][0] = std::sqrt(S[0][0]); L[k][0][1] = 0;
        L[k][1][0] = S[1][0] / L[k][0][0]; L[k][1][1] = std::sqrt(S[1][1] - L[k][1][0] * L[k][1][0]);
    }
    std::vector<double> X((size_t)N * D);   // D x N column-major: point i at X[i * D]
    auto gt = std::make_shared<LabelsType>(N);
    for (int i = 0; i < N; ++i) {
        const int k = i % K;
        (*gt)[i] = k + 1;
        const double z0 = Z(rng), z1 = Z(rng);
        X[(size_t)i * D] = mu[k][0] + L[k][0][0] * z0;
        X[(size_t)i * D + 1] = mu[k][1] + L[k][1][0] * z0 + L[k][1][1] * z1;
    }
    dp_parallel_sampling_class model(N, X.data(), D, 5, prior_type::Gaussian);
    std::shared_ptr<hyperparams> hyper =
        std::make_shared<niw_hyperparams>(1.0, std::vector<double>{0.0, 0.0}, 5.0, std::vector<double>{1, 0, 0, 1});
    ModelInfo info = model.dp_parallel(hyper, 1.0, 400, 1, false, false, false, 15, gt);
    int lo = 1 << 30, hi = -1;
    for (int l : *info.labels) { lo = std::min(lo, l); hi = std::max(hi, l); }
    double ws = 0;
    for (double w : info.dp_model->group.weights) ws += w;   // as main.cpp:76-80 reads them
    size_t first = info.cluster_count_history.size();
    while (first > 0 && info.cluster_count_history[first - 1] == info.cluster_count_history.back()) --first;
    std::printf("gaussian K=%d nmi=%.6f labels=%d..%d weights=%.6f iters=%zu reached_at=%zu nmi_history=%zu\n",
                info.cluster_count_history.back(), dpmm::nmi(*info.labels, *gt), lo, hi, ws, info.iter_count.size(),
                first + 1, info.nmi_score_history.size());
}

static void multinomial_run() {
    const int N = 6000, D = 30, K = 3, W = 60;
    std::mt19937_64 rng(11);
    std::vector<double> X((size_t)N * D, 0.0);
    auto gt = std::make_shared<LabelsType>(N);
    for (int i = 0; i < N; ++i) {
        const int k = i % K;
        (*gt)[i] = k + 1;
        std::uniform_int_distribution<int> word(10 * k, 10 * k + 9);   // disjoint vocabularies
        for (int w = 0; w < W; ++w) X[(size_t)i * D + word(rng)] += 1.0;
    }
    dp_parallel_sampling_class model(N, X.data(), D, 3, prior_type::Multinomial);
    std::shared_ptr<hyperparams> hyper = std::make_shared<multinomial_hyper>(std::vector<double>(D, 1.0));
    ModelInfo info = model.dp_parallel(hyper, 10.0, 80, 1, false, false, false, 15, gt);
    int lo = 1 << 30, hi = -1;
    for (int l : *info.labels) { lo = std::min(lo, l); hi = std::max(hi, l); }
    std::printf("multinomial K=%d nmi=%.6f labels=%d..%d iters=%zu\n", info.cluster_count_history.back(),
                dpmm::nmi(*info.labels, *gt), lo, hi, info.iter_count.size());
}

int main() {
    try {
        gaussian_run();
        multinomial_run();
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}
