// Mixed-precision Gaussian label sweep: BF16 MFMA screen with an error bound, FP32 recompute of
// the points the screen cannot decide.
//
// For cluster k and a block of 32 points, y = W_k (x - mu_k) is one 32 x 32 output tile per row
// block of W_k: A = W_k (rows j, columns l) in BF16, B = the centred residual x - mu_k rounded to
// BF16, FP32 accumulation (v_mfma_f32_32x32x16_bf16).  W_k is upper triangular, so tiles with
// every column below every row (kc < 2 rb) are skipped.  Lane L of the wave holds point
// L % 32 of the block and, for every 16-wide column chunk, the 8 coordinates
// 16 kc + 8 (L / 32) + e; the tile's 32 output rows of that point are split between lanes L and
// L ^ 32, so q = sum of squares is one in-lane sum plus one cross-half add.
//
// Each point keeps its best screened score s_b (with error bound e_b = tau q_b) and the largest
// s_k + e_k of every other cluster.  The label is certain when s_b - e_b exceeds that maximum;
// other points are appended to a recheck list that the FP32 sweep processes exactly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpmm_kernels.h"

namespace dpmm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Slot of one cluster for the screen: [tiles RB x KC x 64 lanes x 8 bf16 | mu (DP f32) | c | pad].
__host__ __device__ constexpr int mx_tiles_bytes(int DP) { return (DP / 32) * (DP / 16) * 64 * 16; }
__host__ __device__ constexpr int mx_slot_bytes(int DP) { return (mx_tiles_bytes(DP) + DP * 4 + 4 + 15) & ~15; }
// Lane-order index of W[j][l] inside the tile array (row j, column l).
__host__ __device__ constexpr int mx_tile_index(int DP, int j, int l) {
    // tile (rb, kc); lane = (j % 32) + 32 * ((l % 16) / 8); element e = l % 8
    return (((j / 32) * (DP / 16) + (l / 16)) * 64 + (j % 32) + 32 * ((l % 16) / 8)) * 8 + (l % 8);
}

// Screen's Gumbel noise: hardware log2, and -log(u) = v + v^2 / 2 for v = 1 - u < 2^-10.  Its
// absolute error stays far below kNoiseSlack, which the error bound adds to every cluster.
constexpr float kNoiseSlack = 1e-3f;
__device__ __forceinline__ float gumbel_screen(uint32_t w) {
    const float v = ((float)(0xFFFFFFFFu - w) + 0.5f) * 0x1p-32f;
    float e;
    if (v < 0x1p-10f) e = v * (1.0f + 0.5f * v);
    else e = -__logf(((float)w + 0.5f) * 0x1p-32f);
    return -__logf(e);
}

template <int DP>
__global__ void __launch_bounds__(kThreads, 2)
k_label_gauss_mx(const float* __restrict__ X, int64_t N, int D, int K, const unsigned char* __restrict__ PB,
                 float tau, uint32_t key0, uint32_t key1, uint32_t iter, int final_iter,
                 int32_t* __restrict__ labels, int32_t* __restrict__ recheck, int* __restrict__ recheck_count) {
    constexpr int RB = DP / 32, KC = DP / 16;
    // 32-point blocks per wave: with two, half h of the wave scores block h; with one (wide points,
    // where two blocks of coordinates would not fit in registers) both halves score the same block.
    constexpr int NB = DP <= 64 ? 2 : 1;
    constexpr int PTS = 4 * 32 * NB;
    constexpr int SLOT = mx_slot_bytes(DP);
    __shared__ __align__(16) unsigned char sp[2][SLOT];
    const int lane = threadIdx.x & 63, h = lane >> 5, pl = lane & 31, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * PTS + wv * 32 * NB + pl;   // point of block nb: base + 32 nb
    const int64_t own = NB == 2 ? base + 32 * h : base;                  // the point this lane scores

    float x[NB][KC][8];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int64_t i = base + 32 * nb;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int l = 16 * kc + 8 * h + e;
                x[nb][kc][e] = (i < N && l < D) ? X[i * D + l] : 0.f;
            }
    }
    float sb = neg_inf<float>(), eb = 0.f, U = neg_inf<float>();
    int bi = 0;
    Ctr4 w = {};

    auto stage = [&](int k, int buf) {
        const uint4* src = reinterpret_cast<const uint4*>(PB + (size_t)k * SLOT);
        uint4* dst = reinterpret_cast<uint4*>(sp[buf]);
        for (int e = threadIdx.x; e < SLOT / 16; e += kThreads) dst[e] = src[e];
    };
    stage(0, 0);
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        if (k + 1 < K) stage(k + 1, (k + 1) & 1);
        const unsigned char* slot = sp[k & 1];
        const bf16x8* tiles = reinterpret_cast<const bf16x8*>(slot);
        const float* mu = reinterpret_cast<const float*>(slot + mx_tiles_bytes(DP));
        const float c = mu[DP];
        float qp[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            bf16x8 Bv[KC];
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                for (int e = 0; e < 8; ++e) Bv[kc][e] = (__bf16)(x[nb][kc][e] - mu[16 * kc + 8 * h + e]);
            float q = 0.f;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                f32x16 acc = {};
#pragma unroll
                for (int kc = 2 * rb; kc < KC; ++kc)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(tiles[(rb * KC + kc) * 64 + lane], Bv[kc], acc, 0,
                                                                  0, 0);
#pragma unroll
                for (int v = 0; v < 16; ++v) q = fmaf(acc[v], acc[v], q);
            }
            qp[nb] = q;   // this half's rows of block nb
        }
        // half h keeps block h: it sends its rows of the other block and receives the other half's rows of its own
        float q;
        if constexpr (NB == 2) q = (h ? qp[1] : qp[0]) + __shfl_xor(h ? qp[0] : qp[1], 32);
        else q = qp[0] + __shfl_xor(qp[0], 32);
        float s = c - 0.5f * q;
        const float err = tau * q + kNoiseSlack;
        if (!final_iter) {
            if ((k & 3) == 0) w = philox4x32_10((uint32_t)own, (uint32_t)(k >> 2), iter, TAG_LABEL, key0, key1);
            s += gumbel_screen(w.v[0]);
            rotate_words(w);
        }
        if (s > sb || k == 0) {
            U = fmaxf(U, sb + eb);
            sb = s; eb = err; bi = k;
        } else {
            U = fmaxf(U, s + err);
        }
        __syncthreads();
    }
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const bool valid = own < N && (NB == 2 || h == 0);
    const bool sure = sb - eb > U;
    if (valid && sure) labels[own] = bi;
    const bool need = valid && !sure;
    const uint64_t m = __ballot(need);
    if (m) {
        int pos0 = 0;
        if (lane == 0) pos0 = atomicAdd(recheck_count, __popcll(m));   // one atomic per wave
        pos0 = __shfl(pos0, 0);
        if (need) recheck[pos0 + __popcll(m & lt)] = (int32_t)own;
    }
}

template <int DP>
void launch_label_gauss_mx(hipStream_t st, int64_t N, int D, int K, const float* X, const unsigned char* PB, float tau,
                           uint32_t key0, uint32_t key1, uint32_t iter, int fin, int32_t* labels, int32_t* recheck,
                           int* recheck_count) {
    constexpr int PTS = 4 * 32 * (DP <= 64 ? 2 : 1);
    const int grid = (int)((N + PTS - 1) / PTS);
    hipLaunchKernelGGL((k_label_gauss_mx<DP>), dim3(grid), dim3(kThreads), 0, st, X, N, D, K, PB, tau, key0, key1,
                       iter, fin, labels, recheck, recheck_count);
}

#define DPMM_MX_INSTANTIATE(DP)                                                                               \
    template void launch_label_gauss_mx<DP>(hipStream_t, int64_t, int, int, const float*, const unsigned char*, \
                                            float, uint32_t, uint32_t, uint32_t, int, int32_t*, int32_t*, int*);
#define DPMM_MX_EXTERN(DP)                                                                                    \
    extern template void launch_label_gauss_mx<DP>(hipStream_t, int64_t, int, int, const float*,               \
                                                   const unsigned char*, float, uint32_t, uint32_t, uint32_t,  \
                                                   int, int32_t*, int32_t*, int*);

}  // namespace dpmm
