// Learned merge-pair screener: MLP 16 -> 128 -> 128 -> 1 over pair features (no counterpart in
// the reference; it only decides which pairs reach the exact merge test of shared_actions.cpp:
// 88-111).  Parameters [W1 (128 x 16) | b1 | W2 (128 x 128) | b2 | w3 | b3] in FP32 on the device;
// inputs are standardised per feature (x - mean) / scale before the first layer.
//
// Training (FP32): k_fwd keeps both hidden layers of a batch, k_bwd_rows forms each row's
// gradients of the hidden layers, k_param_grad reduces the parameter gradients over the batch in
// a fixed order (no atomics), k_adam updates the parameters.
// Inference (BF16 MFMA, FP32 accumulation): k_infer scores 32 rows per wave with
// v_mfma_f32_32x32x16_bf16 for both layers (weights packed once into the MFMA lane order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dpmm_hip.h"
#include "dpmm_device.h"
#include "dpmm_rng.h"

namespace {

constexpr int F = 16, H = 128;
constexpr int OW1 = 0, OB1 = OW1 + H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
constexpr int NP = OB3 + 1;
constexpr int kRows = 32;   // rows per training block
constexpr int kMaxRows = 1 << 20;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(256) k_fwd(const float* __restrict__ P, const float* __restrict__ X,
                                             const float* __restrict__ norm, int n, float* __restrict__ xh,
                                             float* __restrict__ h1, float* __restrict__ h2,
                                             float* __restrict__ z) {
    __shared__ float xs[kRows][F + 1];
    __shared__ float a1[kRows][H + 1];
    __shared__ float a2[kRows][H + 1];
    const int r0 = blockIdx.x * kRows, t = threadIdx.x;
    for (int e = t; e < kRows * F; e += 256) {
        const int r = e / F, i = e % F;
        float v = 0.f;
        if (r0 + r < n) v = (X[(size_t)(r0 + r) * F + i] - norm[i]) / norm[F + i];
        xs[r][i] = v;
        if (r0 + r < n) xh[(size_t)(r0 + r) * F + i] = v;
    }
    __syncthreads();
    for (int e = t; e < kRows * H; e += 256) {
        const int r = e / H, u = e % H;
        float s = P[OB1 + u];
        for (int i = 0; i < F; ++i) s = fmaf(P[OW1 + u * F + i], xs[r][i], s);
        a1[r][u] = s > 0.f ? s : 0.f;
    }
    __syncthreads();
    for (int e = t; e < kRows * H; e += 256) {
        const int r = e / H, u = e % H;
        float s = P[OB2 + u];
        const float* w = P + OW2 + u * H;
        for (int k = 0; k < H; ++k) s = fmaf(w[k], a1[r][k], s);
        a2[r][u] = s > 0.f ? s : 0.f;
    }
    __syncthreads();
    for (int e = t; e < kRows * H; e += 256) {
        const int r = e / H, u = e % H;
        if (r0 + r < n) {
            h1[(size_t)(r0 + r) * H + u] = a1[r][u];
            h2[(size_t)(r0 + r) * H + u] = a2[r][u];
        }
    }
    if (t < kRows && r0 + t < n) {
        float s = P[OB3];
        for (int u = 0; u < H; ++u) s = fmaf(P[OW3 + u], a2[t][u], s);
        z[r0 + t] = s;
    }
}

// dz = w (sigmoid(z) - y) / n; d2 = dz w3 [h2 > 0]; d1 = (W2^T d2) [h1 > 0].
__global__ void __launch_bounds__(256) k_bwd_rows(const float* __restrict__ P, const float* __restrict__ y, int n,
                                                  float pos_weight, const float* __restrict__ z,
                                                  const float* __restrict__ h1, const float* __restrict__ h2,
                                                  float* __restrict__ dz, float* __restrict__ d1,
                                                  float* __restrict__ d2) {
    __shared__ float g2[kRows][H + 1];
    __shared__ float gz[kRows];
    const int r0 = blockIdx.x * kRows, t = threadIdx.x;
    if (t < kRows) {
        float v = 0.f;
        if (r0 + t < n) {
            const float zz = z[r0 + t];
            const float sg = zz >= 0.f ? 1.f / (1.f + expf(-zz)) : expf(zz) / (1.f + expf(zz));
            const bool pos = y[r0 + t] > 0.5f;
            v = (pos ? pos_weight : 1.f) * (sg - (pos ? 1.f : 0.f)) / (float)n;
            dz[r0 + t] = v;
        }
        gz[t] = v;
    }
    __syncthreads();
    for (int e = t; e < kRows * H; e += 256) {
        const int r = e / H, u = e % H;
        float v = 0.f;
        if (r0 + r < n && h2[(size_t)(r0 + r) * H + u] > 0.f) v = gz[r] * P[OW3 + u];
        g2[r][u] = v;
        if (r0 + r < n) d2[(size_t)(r0 + r) * H + u] = v;
    }
    __syncthreads();
    for (int e = t; e < kRows * H; e += 256) {
        const int r = e / H, k = e % H;
        if (r0 + r >= n) continue;
        float s = 0.f;
        for (int u = 0; u < H; ++u) s = fmaf(g2[r][u], P[OW2 + u * H + k], s);
        d1[(size_t)(r0 + r) * H + k] = h1[(size_t)(r0 + r) * H + k] > 0.f ? s : 0.f;
    }
}

// One thread per parameter: its gradient summed over the batch rows in order.
__global__ void __launch_bounds__(256) k_param_grad(int n, const float* __restrict__ xh, const float* __restrict__ h1,
                                                    const float* __restrict__ h2, const float* __restrict__ dz,
                                                    const float* __restrict__ d1, const float* __restrict__ d2,
                                                    float* __restrict__ G) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NP) return;
    float s = 0.f;
    if (p < OB1) {
        const int u = p / F, i = p % F;
        for (int r = 0; r < n; ++r) s = fmaf(d1[(size_t)r * H + u], xh[(size_t)r * F + i], s);
    } else if (p < OW2) {
        const int u = p - OB1;
        for (int r = 0; r < n; ++r) s += d1[(size_t)r * H + u];
    } else if (p < OB2) {
        const int u = (p - OW2) / H, k = (p - OW2) % H;
        for (int r = 0; r < n; ++r) s = fmaf(d2[(size_t)r * H + u], h1[(size_t)r * H + k], s);
    } else if (p < OW3) {
        const int u = p - OB2;
        for (int r = 0; r < n; ++r) s += d2[(size_t)r * H + u];
    } else if (p < OB3) {
        const int u = p - OW3;
        for (int r = 0; r < n; ++r) s = fmaf(dz[r], h2[(size_t)r * H + u], s);
    } else {
        for (int r = 0; r < n; ++r) s += dz[r];
    }
    G[p] = s;
}

// Adam (Kingma & Ba 2015) with bias correction; step t >= 1.
__global__ void k_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V, const float* __restrict__ G,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= NP) return;
    const float g = G[p];
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// W1 and W2 in MFMA lane order: tile (unit block ub, k chunk kc), lane L: unit 32 ub + L % 32,
// k = 16 kc + 8 (L / 32) + e.  W1: 4 x 1 tiles, W2: 4 x 8 tiles; 512 bf16 per tile.
__global__ void k_pack_bf16(const float* __restrict__ P, unsigned short* __restrict__ W1b,
                            unsigned short* __restrict__ W2b) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int n1 = 4 * 512, n2 = 32 * 512;
    if (e < n1) {
        const int tile = e / 512, w = e % 512, L = w / 8, el = w % 8;
        const int u = 32 * tile + L % 32, k = 8 * (L / 32) + el;
        W1b[e] = bf16_rne(P[OW1 + u * F + k]);
    } else if (e < n1 + n2) {
        const int f = e - n1, tile = f / 512, w = f % 512, L = w / 8, el = w % 8;
        const int ub = tile / 8, kc = tile % 8;
        const int u = 32 * ub + L % 32, k = 16 * kc + 8 * (L / 32) + el;
        W2b[f] = bf16_rne(P[OW2 + u * H + k]);
    }
}

// One wave per 32 rows.  Output element v of lane L of a 32 x 32 x 16 MFMA tile: row (unit)
// 8 (v / 4) + 4 (L / 32) + v % 4, column (input row) L % 32.
__global__ void __launch_bounds__(256) k_infer(const float* __restrict__ P, const unsigned short* __restrict__ W1b,
                                               const unsigned short* __restrict__ W2b,
                                               const float* __restrict__ norm, const float* __restrict__ X, int n,
                                               float* __restrict__ z) {
    __shared__ __align__(16) unsigned short hs[4][kRows][H + 8];   // per wave: h1 as BF16, row-major
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, half = lane >> 5, pr = lane & 31;
    const int row = (blockIdx.x * 4 + wv) * kRows + pr;
    const bool valid = row < n;
    bf16x8 xb;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int i = 8 * half + e;
        const float v = valid ? (X[(size_t)row * F + i] - norm[i]) / norm[F + i] : 0.f;
        xb[e] = __builtin_bit_cast(__bf16, bf16_rne(v));
    }
    const bf16x8* A1 = reinterpret_cast<const bf16x8*>(W1b);
#pragma unroll
    for (int ub = 0; ub < 4; ++ub) {
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A1[ub * 64 + lane], xb, acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int u = 32 * ub + 8 * (v / 4) + 4 * half + v % 4;
            const float s = acc[v] + P[OB1 + u];
            hs[wv][pr][u] = bf16_rne(s > 0.f ? s : 0.f);
        }
    }
    __syncthreads();
    const bf16x8* A2 = reinterpret_cast<const bf16x8*>(W2b);
    float zp = 0.f;
#pragma unroll
    for (int ub = 0; ub < 4; ++ub) {
        f32x16 acc = {};
#pragma unroll
        for (int kc = 0; kc < 8; ++kc) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&hs[wv][pr][16 * kc + 8 * half]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A2[(ub * 8 + kc) * 64 + lane], b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int u = 32 * ub + 8 * (v / 4) + 4 * half + v % 4;
            const float s = acc[v] + P[OB2 + u];
            zp = fmaf(P[OW3 + u], s > 0.f ? s : 0.f, zp);
        }
    }
    zp += __shfl_xor(zp, 32);
    if (half == 0 && valid) z[row] = zp + P[OB3];
}

int fail(int code, const std::string& msg) { return dpmm::set_error(code, msg.c_str()); }
#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(DPMM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct dpmm_screener {
    hipStream_t stream = nullptr;
    float *dP = nullptr, *dM = nullptr, *dV = nullptr, *dG = nullptr, *dNorm = nullptr;
    unsigned short *dW1b = nullptr, *dW2b = nullptr;
    bool packed = false;
    int step = 0;
    // batch buffers
    int cap = 0;
    float *dX = nullptr, *dY = nullptr, *dXh = nullptr, *dH1 = nullptr, *dH2 = nullptr, *dZ = nullptr, *dDz = nullptr,
          *dD1 = nullptr, *dD2 = nullptr;
};

namespace {
int ensure_rows(dpmm_screener* s, int n) {
    if (n <= s->cap) return DPMM_OK;
    for (float* p : {s->dX, s->dY, s->dXh, s->dH1, s->dH2, s->dZ, s->dDz, s->dD1, s->dD2})
        if (p) (void)hipFree(p);
    const size_t r = (size_t)n;
    if (hipMalloc((void**)&s->dX, r * F * 4) != hipSuccess || hipMalloc((void**)&s->dY, r * 4) != hipSuccess ||
        hipMalloc((void**)&s->dXh, r * F * 4) != hipSuccess || hipMalloc((void**)&s->dH1, r * H * 4) != hipSuccess ||
        hipMalloc((void**)&s->dH2, r * H * 4) != hipSuccess || hipMalloc((void**)&s->dZ, r * 4) != hipSuccess ||
        hipMalloc((void**)&s->dDz, r * 4) != hipSuccess || hipMalloc((void**)&s->dD1, r * H * 4) != hipSuccess ||
        hipMalloc((void**)&s->dD2, r * H * 4) != hipSuccess) {
        s->cap = 0;
        return fail(DPMM_ENOMEM, "screener batch buffers");
    }
    s->cap = n;
    return DPMM_OK;
}

int loss_grad(dpmm_screener* s, int n, const float* X, const float* y, float pos_weight, float* loss) {
    if (n <= 0 || n > kMaxRows || !X || !y) return fail(DPMM_EINVAL, "bad batch");
    int rc = ensure_rows(s, n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->dX, X, (size_t)n * F * 4, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->dY, y, (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
    const int nb = (n + kRows - 1) / kRows;
    hipLaunchKernelGGL(k_fwd, dim3(nb), dim3(256), 0, s->stream, s->dP, s->dX, s->dNorm, n, s->dXh, s->dH1, s->dH2,
                       s->dZ);
    hipLaunchKernelGGL(k_bwd_rows, dim3(nb), dim3(256), 0, s->stream, s->dP, s->dY, n, pos_weight, s->dZ, s->dH1,
                       s->dH2, s->dDz, s->dD1, s->dD2);
    hipLaunchKernelGGL(k_param_grad, dim3((NP + 255) / 256), dim3(256), 0, s->stream, n, s->dXh, s->dH1, s->dH2,
                       s->dDz, s->dD1, s->dD2, s->dG);
    HIPCHK(hipGetLastError());
    if (loss) {
        std::vector<float> zz(n);
        HIPCHK(hipMemcpyAsync(zz.data(), s->dZ, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        double L = 0.0;
        for (int r = 0; r < n; ++r) {
            const double v = zz[r];
            const bool pos = y[r] > 0.5f;
            const double sp = pos ? (v < 0 ? -v + std::log1p(std::exp(v)) : std::log1p(std::exp(-v)))
                                  : (v > 0 ? v + std::log1p(std::exp(-v)) : std::log1p(std::exp(v)));
            L += (pos ? pos_weight : 1.0) * sp;
        }
        *loss = (float)(L / n);
    }
    s->packed = false;
    return DPMM_OK;
}
}  // namespace

extern "C" {

int dpmm_screener_create(uint64_t seed, dpmm_screener** out) {
    if (!out) return fail(DPMM_EINVAL, "out is null");
    *out = nullptr;
    auto* s = new dpmm_screener();
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&s->dP, NP * 4) != hipSuccess || hipMalloc((void**)&s->dM, NP * 4) != hipSuccess ||
        hipMalloc((void**)&s->dV, NP * 4) != hipSuccess || hipMalloc((void**)&s->dG, NP * 4) != hipSuccess ||
        hipMalloc((void**)&s->dNorm, 2 * F * 4) != hipSuccess ||
        hipMalloc((void**)&s->dW1b, 4 * 512 * 2) != hipSuccess || hipMalloc((void**)&s->dW2b, 32 * 512 * 2) != hipSuccess) {
        dpmm_screener_destroy(s);
        return fail(DPMM_ENOMEM, "screener alloc");
    }
    // He initialisation from the Philox stream (seed, 0, 0, 11): normal * sqrt(2 / fan_in), zero biases
    std::vector<float> P(NP, 0.f);
    uint32_t ctr = 0;
    auto normal = [&]() {
        const dpmm::Ctr4 w = dpmm::philox4x32_10(ctr++, 0, 0, 11, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double a = dpmm::u01_f64(w.v[0]), b = dpmm::u01_f64(w.v[1]);
        return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * 3.14159265358979323846 * b);
    };
    for (int i = 0; i < H * F; ++i) P[OW1 + i] = (float)(normal() * std::sqrt(2.0 / F));
    for (int i = 0; i < H * H; ++i) P[OW2 + i] = (float)(normal() * std::sqrt(2.0 / H));
    for (int i = 0; i < H; ++i) P[OW3 + i] = (float)(normal() * std::sqrt(1.0 / H));
    std::vector<float> norm(2 * F);
    for (int i = 0; i < F; ++i) { norm[i] = 0.f; norm[F + i] = 1.f; }
    if (hipMemcpy(s->dP, P.data(), NP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->dNorm, norm.data(), 2 * F * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->dM, 0, NP * 4) != hipSuccess || hipMemset(s->dV, 0, NP * 4) != hipSuccess) {
        dpmm_screener_destroy(s);
        return fail(DPMM_EHIP, "screener init copy");
    }
    *out = s;
    return DPMM_OK;
}

int dpmm_screener_destroy(dpmm_screener* s) {
    if (!s) return DPMM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void* p : {(void*)s->dP, (void*)s->dM, (void*)s->dV, (void*)s->dG, (void*)s->dNorm, (void*)s->dW1b,
                    (void*)s->dW2b, (void*)s->dX, (void*)s->dY, (void*)s->dXh, (void*)s->dH1, (void*)s->dH2,
                    (void*)s->dZ, (void*)s->dDz, (void*)s->dD1, (void*)s->dD2})
        if (p) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return DPMM_OK;
}

int dpmm_screener_param_count(void) { return NP; }

int dpmm_screener_set_params(dpmm_screener* s, const float* params, const float* mean, const float* scale) {
    if (!s || !params) return fail(DPMM_EINVAL, "bad args");
    HIPCHK(hipMemcpyAsync(s->dP, params, NP * 4, hipMemcpyHostToDevice, s->stream));
    if (mean && scale) {
        std::vector<float> norm(2 * F);
        for (int i = 0; i < F; ++i) {
            if (!(scale[i] > 0.f)) return fail(DPMM_EINVAL, "feature scales must be positive");
            norm[i] = mean[i];
            norm[F + i] = scale[i];
        }
        HIPCHK(hipMemcpyAsync(s->dNorm, norm.data(), 2 * F * 4, hipMemcpyHostToDevice, s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->packed = false;
    return DPMM_OK;
}

int dpmm_screener_get_params(dpmm_screener* s, float* params, float* mean, float* scale) {
    if (!s || !params) return fail(DPMM_EINVAL, "bad args");
    HIPCHK(hipMemcpyAsync(params, s->dP, NP * 4, hipMemcpyDeviceToHost, s->stream));
    std::vector<float> norm(2 * F);
    HIPCHK(hipMemcpyAsync(norm.data(), s->dNorm, 2 * F * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int i = 0; i < F; ++i) {
        if (mean) mean[i] = norm[i];
        if (scale) scale[i] = norm[F + i];
    }
    return DPMM_OK;
}

int dpmm_screener_loss_grad(dpmm_screener* s, int n, const float* X, const float* y, float pos_weight, float* loss,
                            float* grad) {
    if (!s || !grad) return fail(DPMM_EINVAL, "bad args");
    int rc = loss_grad(s, n, X, y, pos_weight, loss);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(grad, s->dG, NP * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return DPMM_OK;
}

int dpmm_screener_train_step(dpmm_screener* s, int n, const float* X, const float* y, float pos_weight, float lr,
                             float* loss) {
    if (!s || !(lr > 0.f)) return fail(DPMM_EINVAL, "bad args");
    int rc = loss_grad(s, n, X, y, pos_weight, loss);
    if (rc) return rc;
    ++s->step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)s->step), c2 = 1.f - std::pow(b2, (float)s->step);
    hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, s->stream, s->dP, s->dM, s->dV, s->dG, lr, b1,
                       b2, 1e-8f, c1, c2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return DPMM_OK;
}

int dpmm_screener_forward(dpmm_screener* s, int n, const float* X, float* z) {
    if (!s || n < 0 || n > kMaxRows || (n > 0 && (!X || !z))) return fail(DPMM_EINVAL, "bad args");
    if (n == 0) return DPMM_OK;
    int rc = ensure_rows(s, n);
    if (rc) return rc;
    if (!s->packed) {
        hipLaunchKernelGGL(k_pack_bf16, dim3((36 * 512 + 255) / 256), dim3(256), 0, s->stream, s->dP, s->dW1b,
                           s->dW2b);
        s->packed = true;
    }
    HIPCHK(hipMemcpyAsync(s->dX, X, (size_t)n * F * 4, hipMemcpyHostToDevice, s->stream));
    const int nb = (n + 4 * kRows - 1) / (4 * kRows);
    hipLaunchKernelGGL(k_infer, dim3(nb), dim3(256), 0, s->stream, s->dP, s->dW1b, s->dW2b, s->dNorm, s->dX, n, s->dZ);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(z, s->dZ, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return DPMM_OK;
}

}  // extern "C"
