/*
 * FP64 CPU oracle of the merge-pair screener (TEST INFRASTRUCTURE ONLY, like dpmm_oracle.cpp):
 * the pair features and the MLP 16 -> 128 -> 128 -> 1 whose forward and backward passes the HIP
 * kernels must reproduce.  The screener has no counterpart in the reference; it only chooses which
 * merge pairs get the reference's exact Hastings test (shared_actions.cpp:88-111).
 *
 * Parameter vector: [W1 (128 x 16) | b1 (128) | W2 (128 x 128) | b2 (128) | w3 (128) | b3],
 * h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), z = w3 . h2 + b3 (the logit of "merge accepted").
 * Loss: mean over rows of binary cross-entropy with logits, rows with y = 1 weighted pos_weight.
 */
#include <cmath>
#include <cstring>
#include <vector>

#include "dpmm_oracle.h"

namespace {
constexpr int F = 16, H = 128;
constexpr int OW1 = 0, OB1 = OW1 + H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
constexpr int NP = OB3 + 1;

double softplus(double z) { return z > 0 ? z + std::log1p(std::exp(-z)) : std::log1p(std::exp(z)); }
double sigmoid(double z) { return z >= 0 ? 1.0 / (1.0 + std::exp(-z)) : std::exp(z) / (1.0 + std::exp(z)); }

void forward_row(const double* p, const double* x, double* h1, double* h2, double* z) {
    for (int u = 0; u < H; ++u) {
        double s = p[OB1 + u];
        for (int i = 0; i < F; ++i) s += p[OW1 + u * F + i] * x[i];
        h1[u] = s > 0 ? s : 0.0;
    }
    for (int u = 0; u < H; ++u) {
        double s = p[OB2 + u];
        for (int k = 0; k < H; ++k) s += p[OW2 + u * H + k] * h1[k];
        h2[u] = s > 0 ? s : 0.0;
    }
    double s = p[OB3];
    for (int u = 0; u < H; ++u) s += p[OW3 + u] * h2[u];
    *z = s;
}
}  // namespace

extern "C" {

int orc_mlp_param_count(void) { return NP; }

void orc_mlp_forward(const double* params, const double* X, int n, double* z) {
    std::vector<double> h1(H), h2(H);
    for (int r = 0; r < n; ++r) forward_row(params, X + (size_t)r * F, h1.data(), h2.data(), z + r);
}

double orc_mlp_loss_grad(const double* params, const double* X, const double* y, int n, double pos_weight,
                         double* grad) {
    std::memset(grad, 0, sizeof(double) * NP);
    std::vector<double> h1(H), h2(H), d2(H), d1(H);
    double loss = 0.0;
    for (int r = 0; r < n; ++r) {
        const double* x = X + (size_t)r * F;
        double z;
        forward_row(params, x, h1.data(), h2.data(), &z);
        const double w = y[r] > 0.5 ? pos_weight : 1.0;
        loss += w * (y[r] > 0.5 ? softplus(-z) : softplus(z));
        const double dz = w * (sigmoid(z) - (y[r] > 0.5 ? 1.0 : 0.0)) / n;
        grad[OB3] += dz;
        for (int u = 0; u < H; ++u) {
            grad[OW3 + u] += dz * h2[u];
            d2[u] = h2[u] > 0 ? dz * params[OW3 + u] : 0.0;
        }
        for (int u = 0; u < H; ++u) {
            grad[OB2 + u] += d2[u];
            for (int k = 0; k < H; ++k) grad[OW2 + u * H + k] += d2[u] * h1[k];
        }
        for (int k = 0; k < H; ++k) {
            double s = 0.0;
            for (int u = 0; u < H; ++u) s += d2[u] * params[OW2 + u * H + k];
            d1[k] = h1[k] > 0 ? s : 0.0;
        }
        for (int u = 0; u < H; ++u) {
            grad[OB1 + u] += d1[u];
            for (int i = 0; i < F; ++i) grad[OW1 + u * F + i] += d1[u] * x[i];
        }
    }
    return loss / n;
}

/* Pair features of the screener from cluster summaries: counts N, Nl, Nr, log marginal
 * likelihoods L, Ll, Lr, log det psi_n (ld), posterior means m (K x D) and diagonals of psi_n
 * (dpsi, K x D).  Row p of out (16 values) describes pair (pi[p], pj[p]). */
void orc_pair_features(int K, int D, const double* N, const double* Nl, const double* Nr, const double* L,
                       const double* Ll, const double* Lr, const double* ld, const double* m, const double* dpsi,
                       int npairs, const int32_t* pi, const int32_t* pj, double* out) {
    for (int p = 0; p < npairs; ++p) {
        const int i = pi[p], j = pj[p];
        double di = 0.0, dj = 0.0, dpool = 0.0;
        for (int a = 0; a < D; ++a) {
            const double d = m[(size_t)i * D + a] - m[(size_t)j * D + a];
            const double si = dpsi[(size_t)i * D + a], sj = dpsi[(size_t)j * D + a];
            di += d * d / si;
            dj += d * d / sj;
            dpool += d * d / (si + sj);
        }
        double* f = out + (size_t)p * 16;
        f[0] = std::log(N[i]);
        f[1] = std::log(N[j]);
        f[2] = std::log(N[i] + N[j]);
        f[3] = L[i] / (N[i] * D);
        f[4] = L[j] / (N[j] * D);
        f[5] = ld[i] / D;
        f[6] = ld[j] / D;
        f[7] = std::log1p(di / D);
        f[8] = std::log1p(dj / D);
        f[9] = std::log1p(dpool / D);
        f[10] = (Ll[i] + Lr[i] - L[i]) / (N[i] * D);
        f[11] = (Ll[j] + Lr[j] - L[j]) / (N[j] * D);
        f[12] = std::fabs(std::log(Nl[i] + 1.0) - std::log(Nr[i] + 1.0));
        f[13] = std::fabs(std::log(Nl[j] + 1.0) - std::log(Nr[j] + 1.0));
        f[14] = std::fabs(f[5] - f[6]);
        f[15] = std::fabs(f[0] - f[1]);
    }
}

}  // extern "C"
