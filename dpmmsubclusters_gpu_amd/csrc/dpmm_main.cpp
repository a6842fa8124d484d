// Command-line front end with the reference's flags and files (main.cpp:6-129):
//
//   dpmm_main --model_path=<points .npy, N x D> --params_path=<params .json>
//             [--prior_type=Gaussian|Multinomial] [--result_path=result.json] [--labels_path=...]
//
// runs dp_parallel_from_file() through the C++ drop-in (include/dpmm/dp_parallel_sampling.hpp) and
// writes result.json with the run's labels (1-based), the model's mixture weights and the seconds of each
// iteration (main.cpp:59-101).  An exception is written as {"error": [message]} with exit status 1
// (main.cpp:109-118).  Without --model_path it runs a small demonstration: the README example's shape,
// a mixture of two 2-D Gaussians, in place of the reference's module_tests::RandomMess.
#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/dpmm/dp_parallel_sampling.hpp"

namespace {

void write_result(const std::string& path, const ModelInfo& info) {
    std::ofstream out(path);
    out << "{\n\t\"labels\" : [";
    for (size_t i = 0; i < info.labels->size(); ++i) out << (i ? "," : "") << (*info.labels)[i];
    out << "],\n\t\"weights\" : [";
    const auto& w = info.dp_model->group.weights;
    for (size_t i = 0; i < w.size(); ++i) out << (i ? "," : "") << dpmm::json::number(w[i]);
    out << "],\n\t\"iter_count\" : [";
    for (size_t i = 0; i < info.iter_count.size(); ++i) out << (i ? "," : "") << dpmm::json::number(info.iter_count[i]);
    out << "]\n}\n";
}

int demo() {
    const int N = 20000, D = 2;
    std::mt19937_64 rng(1);
    std::uniform_real_distribution<double> U(-100, 100);
    std::normal_distribution<double> Z(0, 1);
    const double mu[2][2] = {{U(rng), U(rng)}, {U(rng), U(rng)}};
    std::vector<double> X((size_t)N * D);
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < D; ++d) X[(size_t)i * D + d] = mu[i % 2][d] + Z(rng);
    dp_parallel_sampling_class model(N, X.data(), D, 1, prior_type::Gaussian);
    std::shared_ptr<hyperparams> hyper =
        std::make_shared<niw_hyperparams>(1.0, std::vector<double>{0.0, 0.0}, 5.0, std::vector<double>{1, 0, 0, 1});
    ModelInfo info = model.dp_parallel(hyper, 1.0, 200, 1, true);
    std::printf("demo: %d clusters\n", info.cluster_count_history.back());
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string ParamsPath = "--params_path=", ModelPath = "--model_path=", LabelsPath = "--labels_path=",
                      PriorType = "--prior_type=", ResultPath = "--result_path=";
    std::printf("main was run with %d arguments\n", argc - 1);
    std::string params_path, model_path, labels_path, priortype, result_path = "result.json";
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        auto take = [&](const std::string& flag, std::string& dst) {
            if (s.rfind(flag, 0) != 0) return false;
            dst = s.substr(flag.size());
            return true;
        };
        if (take(ParamsPath, params_path)) std::printf("Will look for params_path in:%s\n", params_path.c_str());
        else if (take(ModelPath, model_path)) std::printf("Will look for model_path in:%s\n", model_path.c_str());
        else if (take(LabelsPath, labels_path)) std::printf("Will look for labels_path in:%s\n", labels_path.c_str());
        else if (take(PriorType, priortype)) std::printf("Prior is:%s\n", priortype.c_str());
        else if (take(ResultPath, result_path)) std::printf("result_path is:%s\n", result_path.c_str());
    }
    if (model_path.empty()) {
        try {
            return demo();
        } catch (const std::exception& e) {
            std::printf("exception:%s\n", e.what());
            return 1;
        }
    }
    try {
        const prior_type pt = priortype.rfind("Multinomial", 0) == 0 ? prior_type::Multinomial : prior_type::Gaussian;
        dp_parallel_sampling_class dps(model_path, params_path, pt);
        ModelInfo info = dps.dp_parallel_from_file();
        write_result(result_path, info);
    } catch (const std::exception& e) {
        std::printf("exception:%s\n", e.what());
        std::ofstream out(result_path);
        out << "{\n\t\"error\" : [" << dpmm::json::quote(e.what()) << "]\n}\n";
        return 1;
    }
    return 0;
}
