// Mixed-precision multinomial label sweep: x . log theta as a split-BF16 MFMA GEMM with FP32
// accumulation (v_mfma_f32_32x32x16_bf16), counts exact in BF16 (<= 256), log theta = hi + lo with
// hi = bf16(log theta), lo = bf16(log theta - hi).  Every score carries a rigorous bound of its
// distance to the FP32 sweep's score; a document whose best perturbed score does not beat the
// runner-up by twice that bound is re-scored by the FP32 sweep (k_label_mult_list), so labels equal
// the FP32 sweep's.
//
// Parameter block (device): hi tiles | lo tiles | c (NCB x 32 floats, -inf past K) | max |log theta|.
// Tile (part hl, chunk c of 16 dimensions, cluster block cb): lane L holds cluster 32 cb + L % 32 and
// dimensions 16 c + 8 (L / 32) + e, e < 8; tiles ordered (hl, c, cb) so a run of chunks is contiguous.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpmm_kernels.h"
#include "dpmm_mixed.h"
#include "dpmm_rng.h"

namespace dpmm {

__host__ __device__ constexpr int mm_chunks(int D) { return (D + 15) / 16; }
__host__ __device__ constexpr size_t mm_tiles_bytes(int D, int NCB) { return (size_t)mm_chunks(D) * NCB * 1024; }
__host__ __device__ constexpr size_t mm_c_offset(int D, int NCB) { return 2 * mm_tiles_bytes(D, NCB); }
__host__ __device__ constexpr size_t mm_lmax_offset(int D, int NCB) { return mm_c_offset(D, NCB) + (size_t)NCB * 32 * 4; }
__host__ __device__ constexpr size_t mm_block_bytes(int D, int NCB) { return mm_lmax_offset(D, NCB) + 16; }

__device__ __forceinline__ uint16_t mm_bf16_rne(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float mm_bf16_f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Zero tiles, c = -inf, max |log theta| = 0.
__global__ void k_mult_mx_clear(unsigned char* __restrict__ blk, int D, int NCB) {
    const size_t nt = 2 * mm_tiles_bytes(D, NCB) / 4;
    uint32_t* w = reinterpret_cast<uint32_t*>(blk);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nt; e += (size_t)gridDim.x * blockDim.x) w[e] = 0;
    if (blockIdx.x == 0) {
        float* c = reinterpret_cast<float*>(blk + mm_c_offset(D, NCB));
        for (int k = threadIdx.x; k < NCB * 32; k += blockDim.x) c[k] = -__builtin_inff();
        if (threadIdx.x == 0) *reinterpret_cast<uint32_t*>(blk + mm_lmax_offset(D, NCB)) = 0;
    }
}

// Cluster k's log theta (FP64, D values at src + k * row_stride): its hi / lo tiles, its score
// constant (log weight or 0, as the FP32 slot holds it) and the running max |log theta|.
__global__ void __launch_bounds__(kThreads) k_pack_mult_mx(const double* __restrict__ src, size_t row_stride, int D,
                                                           int NCB, const double* __restrict__ logw, int use_weight,
                                                           unsigned char* __restrict__ blk) {
    const int k = blockIdx.x;
    const double* lt = src + (size_t)k * row_stride;
    uint16_t* hi = reinterpret_cast<uint16_t*>(blk);
    uint16_t* lo = reinterpret_cast<uint16_t*>(blk + mm_tiles_bytes(D, NCB));
    const int cb = k >> 5, kl = k & 31;
    float lmax = 0.f;
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
        double v = lt[d];
        if (v < -1e30) v = -1e30;   // keeps 0 x log theta finite
        const uint16_t h = mm_bf16_rne((float)v);
        const uint16_t l = mm_bf16_rne((float)(v - (double)mm_bf16_f(h)));
        const int c = d >> 4, L = kl + 32 * ((d & 15) >> 3), e = d & 7;
        const size_t idx = (((size_t)c * NCB + cb) * 64 + L) * 8 + e;
        hi[idx] = h;
        lo[idx] = l;
        lmax = fmaxf(lmax, fabsf((float)v));
    }
    for (int o = 32; o > 0; o >>= 1) lmax = fmaxf(lmax, __shfl_xor(lmax, o));
    if ((threadIdx.x & 63) == 0)
        atomicMax(reinterpret_cast<unsigned int*>(blk + mm_lmax_offset(D, NCB)), __float_as_uint(lmax));
    if (threadIdx.x == 0)
        reinterpret_cast<float*>(blk + mm_c_offset(D, NCB))[k] = use_weight ? (float)logw[k] : 0.f;
}

// 32 documents per wave, 4 waves per workgroup sharing the LDS-staged tiles of G chunks at a time.
template <int NCB>
__global__ void __launch_bounds__(kThreads, NCB >= 4 && NCB <= 6 ? 3 : 2) k_label_mult_mx(const int32_t* __restrict__ X, int64_t N, int D,
                                                            int vec, const unsigned char* __restrict__ blk,
                                                            uint32_t key0,
                                                            uint32_t key1, uint32_t iter, int final_iter,
                                                            int32_t* __restrict__ labels,
                                                            int32_t* __restrict__ recheck,
                                                            int* __restrict__ recheck_count) {
    // groups of G chunks (16 G words): the tiles of group g + 1 stream into the other LDS buffer by
    // LDS DMA and its counts into registers while group g runs on MFMA
    constexpr int G = NCB >= 4 ? 2 : 4;
    __shared__ __align__(16) bf16x8 sA[2][2][G][NCB][64];   // [buffer][hi / lo][chunk][cluster block][lane]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, half = lane >> 5;
    const int64_t doc = ((int64_t)blockIdx.x * 4 + wv) * 32 + (lane & 31);
    const bool valid = doc < N;
    const int NC = mm_chunks(D);
    const unsigned char* hiB = blk;
    const unsigned char* loB = blk + mm_tiles_bytes(D, NCB);
    f32x16 acc[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) acc[cb] = f32x16{};
    int words = 0;
    const int32_t* xr = X + (valid ? doc : 0) * (int64_t)D;
    auto load_counts = [&](int g0, int (&xg)[G][8]) {
#pragma unroll
        for (int gc = 0; gc < G; ++gc) {
            const int d0 = 16 * (g0 + gc) + 8 * half;
            if (vec && valid && d0 + 8 <= D) {   // rows 16-byte aligned: two int4 loads
                const int4 a = *reinterpret_cast<const int4*>(xr + d0);
                const int4 b = *reinterpret_cast<const int4*>(xr + d0 + 4);
                xg[gc][0] = a.x; xg[gc][1] = a.y; xg[gc][2] = a.z; xg[gc][3] = a.w;
                xg[gc][4] = b.x; xg[gc][5] = b.y; xg[gc][6] = b.z; xg[gc][7] = b.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) xg[gc][e] = (valid && d0 + e < D) ? xr[d0 + e] : 0;
            }
        }
    };
    // tiles of chunks g0 .. g0 + G - 1 (those that exist): contiguous in the parameter block
    auto stage = [&](int g0, int buf) {
        const int bytes = min(G, NC - g0) * NCB * 1024;
        lds_dma_copy(reinterpret_cast<unsigned char*>(&sA[buf][0][0][0][0]), hiB + (size_t)g0 * NCB * 1024, bytes, wv,
                     lane);
        lds_dma_copy(reinterpret_cast<unsigned char*>(&sA[buf][1][0][0][0]), loB + (size_t)g0 * NCB * 1024, bytes, wv,
                     lane);
    };
    int xg[G][8], xn[G][8];
    load_counts(0, xg);
    stage(0, 0);
    lds_dma_wait();
    __syncthreads();
    for (int g0 = 0, buf = 0; g0 < NC; g0 += G, buf ^= 1) {
        const bool more = g0 + G < NC;
        if (more) {   // the previous group's readers of sA[buf ^ 1] passed the barrier below
            load_counts(g0 + G, xn);
            stage(g0 + G, buf ^ 1);
        }
#pragma unroll
        for (int gc = 0; gc < G; ++gc) {
            if (g0 + gc >= NC) break;
            bf16x8 xb;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int v = xg[gc][e];
                words += v;
                xb[e] = (__bf16)(float)v;
            }
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sA[buf][0][gc][cb][lane], xb, acc[cb], 0, 0, 0);
                acc[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sA[buf][1][gc][cb][lane], xb, acc[cb], 0, 0, 0);
            }
        }
        lds_dma_wait();   // the next group's tiles (this wave's share) and counts have landed
        __syncthreads();  // every wave's share is in, and every wave is done with sA[buf]
#pragma unroll
        for (int gc = 0; gc < G; ++gc)
#pragma unroll
            for (int e = 0; e < 8; ++e) xg[gc][e] = xn[gc][e];
    }
    words += __shfl_xor(words, 32);
    const float lmax = *reinterpret_cast<const float*>(blk + mm_lmax_offset(D, NCB));
    const float* cst = reinterpret_cast<const float*>(blk + mm_c_offset(D, NCB));
    // |score - FP32 sweep score| <= words * lmax * (split residual 2^-17 + accumulation of both sums)
    const float err = (float)words * lmax * (0x1p-17f + 4.0f * (float)(D + 16) * 0x1p-24f) + kNoiseSlack;
    float s1 = neg_inf<float>(), s2 = neg_inf<float>();
    int k1 = 0x7fffffff;
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Ctr4 w = {};
            const int kq = 32 * cb + 8 * q + 4 * half;
            if (!final_iter) w = philox4x32_10((uint32_t)doc, (uint32_t)(kq >> 2), iter, TAG_LABEL, key0, key1);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = kq + r;
                float s = acc[cb][4 * q + r] + cst[k];
                if (!final_iter) s += gumbel_screen(w.v[r]);
                if (s > s1 || (s == s1 && k < k1)) { s2 = s1; s1 = s; k1 = k; }
                else if (s > s2) s2 = s;
            }
        }
    }
    // combine with the other half of the document's clusters
    const float o1 = __shfl_xor(s1, 32), o2 = __shfl_xor(s2, 32);
    const int ok1 = __shfl_xor(k1, 32);
    float b1 = s1, b2 = fmaxf(s2, o1);
    int bk = k1;
    if (o1 > s1 || (o1 == s1 && ok1 < k1)) { b1 = o1; bk = ok1; b2 = fmaxf(s1, o2); }
    const bool sure = b1 - b2 > 2.0f * err;
    const bool own = valid && half == 0;
    if (own && sure) labels[doc] = bk;
    const bool need = own && !sure;
    const uint64_t m = __ballot(need);
    if (m) {
        const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        int pos0 = 0;
        if (lane == 0) pos0 = atomicAdd(recheck_count, __popcll(m));   // one atomic per wave
        pos0 = __shfl(pos0, 0);
        if (need) recheck[pos0 + __popcll(m & lt)] = (int32_t)doc;
    }
}

// FP32 re-scoring of listed documents, one wave per document, with exactly the FP32 sweep's
// arithmetic (k_label_mult): per cluster the fmaf chain over the dimensions in order, the slot's
// constant, then the Gumbel noise; the lowest cluster index wins ties.
__global__ void __launch_bounds__(kThreads) k_label_mult_list(const int32_t* __restrict__ X, int D, int K,
                                                              const float* __restrict__ P, int stride, uint32_t key0,
                                                              uint32_t key1, uint32_t iter, int final_iter,
                                                              const int32_t* __restrict__ list,
                                                              const int* __restrict__ count,
                                                              int32_t* __restrict__ labels) {
    const int lane = threadIdx.x & 63;
    const int n = *count;
    for (int w = blockIdx.x * 4 + (threadIdx.x >> 6); w < n; w += gridDim.x * 4) {
        const int64_t i = list[w];
        const int32_t* x = X + i * (int64_t)D;
        float best = neg_inf<float>();
        int bk = 0x7fffffff;
        for (int k = lane; k < K; k += 64) {
            const float* th = P + (size_t)k * stride;
            float a = 0.f;
            for (int d = 0; d < D; ++d) a = fmaf((float)x[d], th[d], a);
            float s = a + th[D];
            if (!final_iter) s += gumbel_f32(philox4x32_10((uint32_t)i, (uint32_t)(k >> 2), iter, TAG_LABEL, key0, key1).v[k & 3]);
            if (s > best) { best = s; bk = k; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const float ob = __shfl_xor(best, o);
            const int ok = __shfl_xor(bk, o);
            if (ok != 0x7fffffff && (bk == 0x7fffffff || ob > best || (ob == best && ok < bk))) { best = ob; bk = ok; }
        }
        if (lane == 0) labels[i] = bk;
    }
}

}  // namespace dpmm
