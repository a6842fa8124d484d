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
// Error bound against the FP32 sweep (k_label_gauss, which forms y_j = sum_l W_jl d_l by FMAs in
// FP32 and q = sum_j y_j^2 the same way, with the same FP32 W, mu and c): BF16 rounding (8 significant
// bits, to nearest) of W and of d = x - mu is relative 2^-8 each, so every product W_jl d_l is off by at
// most (2^-7 + 2^-16) |W_jl| |d_l|, and the screen's and the sweep's FP32 sums add at most (2 DP + 16) 2^-24 of
// a_j = sum_l |W_jl| |d_l|.  A second MFMA per tile, on the magnitudes |bf16(W)| |bf16(d)|, gives a_j
// (inflated by 2^-6 for its own rounding), so |y_screen,j - y_fp32,j| <= eps a_j and, with
// A = sum_j a_j^2 and sum_j |y_j| a_j <= sqrt(q A) (Cauchy-Schwarz),
//   |q_screen - q_fp32| <= sum_j (2 |y_j| + eps a_j) eps a_j + (2 DP + 4) 2^-24 q
//                       <= 2 eps sqrt(q A) + eps^2 A + (2 DP + 4) 2^-24 q
// (the last term: both sums of squares).  Half of it, the rounding of c - q / 2 in both and the
// screen's fast Gumbel noise (kNoiseSlack) bound |s_screen - s_fp32|; this is e_k.  The bound scales
// with |W| |d|, not with |W d|, so it holds whatever the conditioning of the covariance.
//
// Each point keeps its best screened score s_b (with error bound e_b) and the largest s_k + e_k of
// every other cluster.  The label is the FP32 sweep's when s_b - e_b exceeds that maximum; other
// points are appended to a recheck list that the FP32 sweep processes exactly.  `tau` scales the
// bound: 1 is the derived bound; 0 keeps every screened label (a diagnostic of the raw screen).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpmm_kernels.h"

namespace dpmm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void glob_cvoid;

// Copy of `bytes` (a multiple of 16) from global memory into LDS by the wave-level LDS DMA
// (global_load_lds_dwordx4): wave w of the workgroup moves 1 KB chunks w, w + 4, ..., lane L the 16
// bytes at L * 16 of its chunk, straight into LDS with no register round trip.  The copy is
// asynchronous: the caller waits for it with lds_dma_wait() before the barrier that publishes it.
__device__ __forceinline__ void lds_dma_copy(unsigned char* dst, const unsigned char* src, int bytes, int wv,
                                             int lane) {
    const int units = bytes >> 4;
    for (int ch = wv; ch * 64 < units; ch += kThreads / 64) {
        const int e = ch * 64 + lane;
        if (e < units)
            __builtin_amdgcn_global_load_lds((glob_cvoid*)(src + (size_t)e * 16), (lds_void*)(dst + ch * 1024), 16, 0,
                                             0);
    }
}
// vmcnt(0), expcnt and lgkmcnt left at their maxima: every outstanding vector memory load of the wave,
// the LDS DMA copies included, has completed.
__device__ __forceinline__ void lds_dma_wait() { __builtin_amdgcn_s_waitcnt(0x0f70); }

// Wave-private staging of a recheck list: entries are appended by ballot into 64 LDS slots owned by
// the wave and moved to the global list with one atomicAdd per 64 entries (and once at the end), so
// waves that loop over many point blocks do not each add to the single global counter per block,
// where the atomics serialise.  n is wave-uniform; every lane of the wave calls push / flush.
struct WaveList {
    int32_t* buf;   // 64 entries of LDS, private to the wave
    int n;
    __device__ __forceinline__ void sync_wave() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // entry i of the list goes to list[dir * i] (dir -1: a list that grows down from `list`)
    __device__ __forceinline__ void flush(int32_t* list, int* count, int lane, int dir = 1) {
        if (n == 0) return;
        sync_wave();   // the lanes' appends are visible to every lane
        int pos0 = 0;
        if (lane == 0) pos0 = atomicAdd(count, n);
        pos0 = __shfl(pos0, 0);
        if (lane < n) list[dir * (pos0 + lane)] = buf[lane];
        sync_wave();   // read before the next appends overwrite
        n = 0;
    }
    __device__ __forceinline__ void push(bool need, int32_t v, int32_t* list, int* count, int lane, uint64_t lt,
                                         int dir = 1) {
        const uint64_t m = __ballot(need);
        if (!m) return;
        const int c = __popcll(m);
        if (n + c > 64) flush(list, count, lane, dir);
        if (need) buf[n + __popcll(m & lt)] = v;
        n += c;
    }
};

// Slot of one cluster for the screen: [tiles RB x KC x 64 lanes x 8 bf16 | mu (DP f32) | c | pad].
__host__ __device__ constexpr int mx_tiles_bytes(int DP) { return (DP / 32) * (DP / 16) * 64 * 16; }
__host__ __device__ constexpr int mx_slot_bytes(int DP) { return (mx_tiles_bytes(DP) + DP * 4 + 4 + 15) & ~15; }
// A cluster slot also holds, after c, F >= ||W||_F^2 of the FP32 W (the fused sweep's anchor bound).
// Sub-cluster slot of the sub-label screen: W split as hi + lo tiles, then mu (DP f32), c, pad, and for
// the screen relative to the parent cluster's mean mu_a (the fused sweep): -c_j (DP f32) with
// c = W (mu - mu_a) from the FP32 slot values, and g_j (DP f32) = kappa (sum_l |W_jl| |mu_l - mu_a,l| + |c_j|)
// rounded up, kappa = sub_rel_eps_g / sub_rel_eps_a.
__host__ __device__ constexpr int mx2_slot_bytes(int DP) { return 2 * mx_tiles_bytes(DP) + DP * 4 + 16 + 2 * DP * 4; }
__host__ __device__ constexpr int mx2_rel_offset(int DP) { return 2 * mx_tiles_bytes(DP) + DP * 4 + 16; }
// BF16 has 8 significant bits: rounding to nearest is relative 2^-8, |v - bf16(v)| <= 2^-8 |v|.  An FP32
// v split as v_hi = bf16(v), v_lo = bf16(v - v_hi) has |v_lo| <= 2^-8 (1 + 2^-8) |v| and leaves
// |v - v_hi - v_lo| <= 2^-17 |v| (v - v_hi is exact, below half an ulp of v_hi, with at most 16 significant
// bits).  Per-row factor of the split screens, |y_screen,j - y_fp32,j| <= eps a_j with a_j = sum_l |W_jl| |d_l|
// from the BF16 magnitudes |W_hi| |d_hi| inflated by 2^-6 (which covers |W| <= |W_hi| / (1 - 2^-8), the same
// for d, and the magnitude MFMA's rounding): the three products W_hi d_hi + W_hi d_lo + W_lo d_hi leave
// W_lo d_lo + (W_hi + W_lo) delta_d + delta_W d, at most (2^-16 (1 + 2^-8)^2 + 2^-17 (2 + 2^-7)) |W| |d|
// < (2^-15 + 2^-22) |W| |d| (attained to within 2^-6); the screen's FP32 sum of 3 DP products (their
// magnitudes sum to at most (1 + 2^-7) |W| |d|), the sweep's sum of DP and the FP32 roundings of the
// residuals add at most (4.03 DP + 2) 2^-24 of a_j.  (A fourth product, W_lo d_lo, would halve the first
// term at 12 more MFMAs per 32 points of the fused sweep, which cost more than the extra FP32 draws.)
__host__ __device__ constexpr float split_screen_eps(int DP) {
    return (0x1p-15f + 0x1p-22f + (5.0f * (float)DP + 24.0f) * 0x1p-24f) * (1.0f + 0x1p-6f);
}
// A sub-cluster slot also holds, after c, gn >= ||g|| and fh >= ||W_hi||_F (the fused sweep's bound on
// sum_j a'_j^2, see fused_screens_split).
// Per-row error factors of the relative sub-cluster screen (fused_screens_split): |y_screen,j - y_fp32,j| <=
// eps_a a_j + eps_g g_j, eps_a as above; eps_g covers the screen's FP32 sum started from -c_j (3 DP + 1
// additions), the rounding of c_j and the sweep's residual and sum over |W| |mu_h - mu_a|.
__host__ __device__ constexpr float sub_rel_eps_a(int DP) { return split_screen_eps(DP); }
__host__ __device__ constexpr float sub_rel_eps_g(int DP) { return (3.0f * (float)DP + 16.0f) * 0x1p-24f * (1.0f + 0x1p-6f); }
// Lane-order index of W[j][l] inside the tile array (row j, column l).
__host__ __device__ constexpr int mx_tile_index(int DP, int j, int l) {
    // tile (rb, kc); lane = (j % 32) + 32 * ((l % 16) / 8); element e = l % 8
    return (((j / 32) * (DP / 16) + (l / 16)) * 64 + (j % 32) + 32 * ((l % 16) / 8)) * 8 + (l % 8);
}

// Screen's Gumbel noise: hardware log2, and -log(u) = v + v^2 / 2 for v = 1 - u < 2^-10.  Its
// absolute error stays far below kNoiseSlack, which the error bound adds to every cluster.
constexpr float kNoiseSlack = 1e-3f;

// eps of the label screen's per-row bound |y_screen,j - y_fp32,j| <= eps a_j (see the header).
__host__ __device__ constexpr float label_screen_eps_of(int DP) {
    return (0x1p-7f + 0x1p-16f + (2.0f * (float)DP + 16.0f) * 0x1p-24f) * (1.0f + 0x1p-6f);
}
template <int DP> __device__ __forceinline__ constexpr float label_screen_eps() { return label_screen_eps_of(DP); }
// e_k from q = sum y_j^2, A = sum a_j^2 and the score constant c (sqrt rounding covered by 2^-10).
template <int DP> __device__ __forceinline__ float label_screen_err(float tau, float q, float A, float c) {
    constexpr float eps = label_screen_eps<DP>();
    const float dq = (eps * (sqrtf(q * A) + 0.5f * eps * A) + ((float)DP + 2.0f) * 0x1p-24f * q) * (1.0f + 0x1p-10f);
    return tau * (dq + 0x1p-22f * (fabsf(c) + 0.5f * q)) + kNoiseSlack;
}
// |v| of eight packed BF16 values
__device__ __forceinline__ bf16x8 bf16x8_abs(bf16x8 v) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    return __builtin_bit_cast(bf16x8, __builtin_bit_cast(u32x4, v) & 0x7fff7fffu);
}
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
        const float* row = X + (i < N ? i : 0) * D;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) load_coords8(row, 16 * kc + 8 * h, D, i < N, x[nb][kc]);
    }
    float sb = neg_inf<float>(), eb = 0.f, U = neg_inf<float>();
    int bi = 0;
    Ctr4 w = {};

    // cluster k + 1's slot streams into the other LDS buffer while cluster k is scored
    auto stage = [&](int k, int buf) { lds_dma_copy(sp[buf], PB + (size_t)k * SLOT, SLOT, wv, lane); };
    stage(0, 0);
    lds_dma_wait();
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        if (k + 1 < K) stage(k + 1, (k + 1) & 1);
        const unsigned char* slot = sp[k & 1];
        const bf16x8* tiles = reinterpret_cast<const bf16x8*>(slot);
        const float* mu = reinterpret_cast<const float*>(slot + mx_tiles_bytes(DP));
        const float c = mu[DP];
        float qp[NB], ep[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            bf16x8 Bv[KC];
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                for (int e = 0; e < 8; ++e) Bv[kc][e] = (__bf16)(x[nb][kc][e] - mu[16 * kc + 8 * h + e]);
            float q = 0.f, ee = 0.f;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                f32x16 acc = {}, aa = {};
#pragma unroll
                for (int kc = 2 * rb; kc < KC; ++kc) {
                    const bf16x8 wt = tiles[(rb * KC + kc) * 64 + lane];
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wt, Bv[kc], acc, 0, 0, 0);
                    aa = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bf16x8_abs(wt), bf16x8_abs(Bv[kc]), aa, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    q = fmaf(acc[v], acc[v], q);
                    ee = fmaf(aa[v], aa[v], ee);
                }
            }
            qp[nb] = q;   // this half's rows of block nb
            ep[nb] = ee;
        }
        // half h keeps block h: it sends its rows of the other block and receives the other half's rows of its own
        float q, ee;
        if constexpr (NB == 2) {
            q = (h ? qp[1] : qp[0]) + __shfl_xor(h ? qp[0] : qp[1], 32);
            ee = (h ? ep[1] : ep[0]) + __shfl_xor(h ? ep[0] : ep[1], 32);
        } else {
            q = qp[0] + __shfl_xor(qp[0], 32);
            ee = ep[0] + __shfl_xor(ep[0], 32);
        }
        float s = c - 0.5f * q;
        const float err = label_screen_err<DP>(tau, q, ee, c);
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
        lds_dma_wait();
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

// Split-BF16 score of one cluster for the 32 points of a wave (lane L: point L % 32, coordinates
// 16 kc + 8 (L / 32) .. + 7 in x): with d = x - mu in FP32 split as d_hi + d_lo and W = W_hi + W_lo,
// y = W_hi d_hi + W_hi d_lo + W_lo d_hi on MFMA (FP32 accumulation), and a fourth MFMA gives
// a_j = sum_l |W_hi,jl| |d_hi,l|.  Each y_j is then within eps a_j of the FP32 sweep's (eps covers the
// split residuals and both FP32 accumulations), so q is within sum_j (2 |y_j| + eps a_j) eps a_j of the
// FP32 sweep's q.  Returns s = c - q / 2 and sets dq, the bound on |q - q_fp32| (both sums of squares
// included).  Every lane of the wave calls it.
template <int DP>
__device__ __forceinline__ float split_score(const float (&x)[DP / 16][8], const bf16x8* thi, const bf16x8* tlo,
                                             const float* mu, float c, int lane, float& dq) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    constexpr int RB = DP / 32, KC = DP / 16;
    const int h = lane >> 5;
    // |y_screen - y_fp32| <= eps a_j (split_screen_eps)
    constexpr float eps = split_screen_eps(DP);
    const float heps = 0.5f * eps;
    // d = x - mu split as d_hi + d_lo: one packed conversion per coordinate pair gives both BF16 highs,
    // read back as FP32 by shifting / masking the packed word
    u32x4 dh[KC], dl[KC];
#pragma unroll
    for (int kc = 0; kc < KC; ++kc)
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            const float d0 = x[kc][e] - mu[16 * kc + 8 * h + e];
            const float d1 = x[kc][e + 1] - mu[16 * kc + 8 * h + e + 1];
            bf16x2 hp;
            hp[0] = (__bf16)d0;
            hp[1] = (__bf16)d1;
            const uint32_t hb = __builtin_bit_cast(uint32_t, hp);
            bf16x2 lp;
            lp[0] = (__bf16)(d0 - __uint_as_float(hb << 16));
            lp[1] = (__bf16)(d1 - __uint_as_float(hb & 0xffff0000u));
            dh[kc][e / 2] = hb;
            dl[kc][e / 2] = __builtin_bit_cast(uint32_t, lp);
        }
    // qq = sum y_j^2; ee = sum a_j (|y_j| + eps a_j / 2), so that the bound
    // sum (2 |y_j| + eps a_j) eps a_j is 2 eps ee
    float qq = 0.f, ee = 0.f;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        f32x16 y = {}, a = {};
#pragma unroll
        for (int kc = 2 * rb; kc < KC; ++kc) {
            const bf16x8 wh = thi[(rb * KC + kc) * 64 + lane];
            const bf16x8 wl = tlo[(rb * KC + kc) * 64 + lane];
            const bf16x8 bh = __builtin_bit_cast(bf16x8, dh[kc]), bl = __builtin_bit_cast(bf16x8, dl[kc]);
            const u32x4 wabs = __builtin_bit_cast(u32x4, wh) & 0x7fff7fffu;
            const u32x4 dabs = dh[kc] & 0x7fff7fffu;
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh, bh, y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh, bl, y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl, bh, y, 0, 0, 0);
            a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wabs),
                                                        __builtin_bit_cast(bf16x8, dabs), a, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            qq = fmaf(y[v], y[v], qq);
            ee = fmaf(a[v], fmaf(heps, a[v], fabsf(y[v])), ee);
        }
    }
    qq += __shfl_xor(qq, 32);
    ee = 2.0f * eps * (ee + __shfl_xor(ee, 32));
    dq = ee + (2.0f * (float)DP + 4.0f) * 0x1p-24f * qq;   // the two sums of squares
    return c - 0.5f * qq;
}

// Split-BF16 screen of the two sub-clusters whose slots start at sp (mx2 slots, one after the other):
// returns s_0 - s_1 of the lane's point and sets err, the bound on its distance from the FP32 sweep's
// S_0 - S_1 (see k_sub_gauss_mx).  Every lane of the wave calls it.
template <int DP>
__device__ __forceinline__ float sub_screen(const float (&x)[DP / 16][8], const unsigned char* sp, int lane,
                                            float& err) {
    constexpr int SLOT = mx2_slot_bytes(DP), TB = mx_tiles_bytes(DP);
    float s0 = 0.f, s1 = 0.f, dq0 = 0.f, dq1 = 0.f;
#pragma unroll 1
    for (int hh = 0; hh < 2; ++hh) {
        const unsigned char* slot = sp + hh * SLOT;
        const float* mu = reinterpret_cast<const float*>(slot + 2 * TB);
        float dv;
        const float sv = split_score<DP>(x, reinterpret_cast<const bf16x8*>(slot),
                                         reinterpret_cast<const bf16x8*>(slot + TB), mu, mu[DP], lane, dv);
        if (hh == 0) { s0 = sv; dq0 = dv; } else { s1 = sv; dq1 = dv; }
    }
    err = 0.5f * (dq0 + dq1) * (1.0f + 0x1p-10f) + 0x1p-21f * (fabsf(s0) + fabsf(s1)) + 1e-6f;
    return s0 - s1;
}

// Sub-label screen of one chunk (the points of cluster k): both sub-clusters' split-BF16 slots in
// LDS, 32 gathered points per wave, scored by split_score.  The exact draw sets sub-label 0 when
// logit(u) < S_0 - S_1 (k_sub_gauss); a point whose threshold is farther than the bound from s_0 - s_1
// gets its sub-label here, the others go to the recheck list.
template <int DP>
__global__ void __launch_bounds__(kThreads) k_sub_gauss_mx(const float* __restrict__ X, int D, int K, int CH,
                                                           const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ label_start,
                                                           const int32_t* __restrict__ chunk_start,
                                                           const unsigned char* __restrict__ SPB, float tau,
                                                           uint32_t key0, uint32_t key1, uint32_t iter,
                                                           int32_t* __restrict__ subs, int32_t* __restrict__ subs_perm,
                                                           int32_t* __restrict__ recheck,
                                                           int* __restrict__ recheck_count) {
    constexpr int KC = DP / 16;
    constexpr int SLOT = mx2_slot_bytes(DP);
    __shared__ __align__(16) unsigned char sp[2 * SLOT];
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end)) return;
    {
        const uint4* src = reinterpret_cast<const uint4*>(SPB + (size_t)(2 * k) * SLOT);
        uint4* dst = reinterpret_cast<uint4*>(sp);
        for (int e = threadIdx.x; e < 2 * SLOT / 16; e += kThreads) dst[e] = src[e];
    }
    __syncthreads();
    (void)tau;
    const int lane = threadIdx.x & 63, h = lane >> 5, pl = lane & 31, wv = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __shared__ int32_t wlist[4][64];
    WaveList rl{wlist[wv], 0};
    for (int p0 = beg + 32 * wv; p0 < end; p0 += 4 * 32) {   // wave-uniform trip count
        const int p = p0 + pl;
        const bool valid = p < end;
        const int idx = valid ? perm[p] : 0;
        float x[KC][8];
        const float* row = X + (int64_t)idx * D;   // idx = 0 for an empty lane
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) load_coords8(row, 16 * kc + 8 * h, D, valid, x[kc]);
        float err;
        const float d = sub_screen<DP>(x, sp, lane, err);
        bool need = false;
        if (valid && h == 0) {
            const float lg = logit_u32<float>(philox4x32_10((uint32_t)idx, 0, iter, TAG_SUB, key0, key1).v[0]);
            if (fabsf(lg - d) > err) {
                const int v = lg < d ? 0 : 1;
                subs[idx] = v;
                subs_perm[p] = v;   // partition order, for the statistics pass (coalesced)
            } else {
                need = true;
            }
        }
        rl.push(need, p, recheck, recheck_count, lane, lt);   // positions: the recheck writes both orders
    }
    rl.flush(recheck, recheck_count, lane);
}

template <int DP>
void launch_sub_gauss_mx(hipStream_t st, int grid, const float* X, int D, int K, int CH, const int32_t* perm,
                         const int32_t* label_start, const int32_t* chunk_start, const unsigned char* SPB, float tau,
                         uint32_t key0, uint32_t key1, uint32_t iter, int32_t* subs, int32_t* subs_perm,
                         int32_t* recheck, int* recheck_count) {
    hipLaunchKernelGGL((k_sub_gauss_mx<DP>), dim3(grid), dim3(kThreads), 0, st, X, D, K, CH, perm, label_start,
                       chunk_start, SPB, tau, key0, key1, iter, subs, subs_perm, recheck, recheck_count);
}

// Exact pruning of the mixed label sweep.  The points are visited in the label partition of the
// previous labels: one chunk holds points of one anchor cluster a, and a wave takes 32 of them.
// After screening the anchor, the wave knows r >= ||x - mu_a|| for its points and T, the smallest
// screened lower bound s_a - e_a of their anchor scores.  For any other cluster k,
//   ||W_k (x - mu_k)|| >= ||W_k (mu_a - mu_k)|| - ||W_k||_2 r   (triangle inequality),
// and with m_ak <= ||W_k (mu_a - mu_k)|| and sig_k >= ||W_k||_2, both reduced / raised by the FP32
// sweep's rounding of W (x - mu) (k_prune_prep), the FP32 sweep's q_k is at least 0.98 (m_ak -
// sig_k r)^2.  A 32-bit uniform's Gumbel noise is below 22.9, so
//   ub_k = c_k - 0.49 max(0, m_ak - sig_k r)^2 + 1e-6 |c_k| + kPruneNoise
// bounds the FP32 sweep's noisy score of cluster k for every point of the wave.  When ub_k < T no
// point can take label k and the cluster is skipped (ub_k joins the runner-up bound U); the other
// clusters are screened as in k_label_gauss_mx, so the decided labels are the FP32 sweep's.

// Screens of the fused sweep for the 32 points of a wave (lane L: point L % 32, coordinates
// 16 kc + 8 (L / 32) .. + 7 in x) against the anchor a (slot sa) and its two sub-clusters (slots ssub),
// from one hi / lo BF16 split of p = fl(x - mu_a):
//   anchor: y = bf16(W_a) p_hi on MFMA, the label screen's y (k_label_gauss_mx rounds p the same way);
//     its bound takes A = F_a ||p||^2 >= sum_j (sum_l |W_jl| |p_l|)^2 (Cauchy-Schwarz per row) in place
//     of the magnitude MFMA;
//   sub-cluster h: W_h (x - mu_h) = W_h p - c_h, c_h = W_h (mu_h - mu_a), as
//     y = -c_h + W_hi p_hi + W_hi p_lo + W_lo p_hi (FP32 accumulation started from -c_h), with
//     a'_j = g_j + sum_l |W_hi,jl| |p_hi,l| (started from g_j).  Against the FP32 sweep's
//     y_j = sum_l W_jl fl(x_l - mu_h,l) the screen is off by at most eps_a a_j + eps_g g'_j, where the
//     split residuals and the FP32 sums of both sides give eps_a and the FP32 roundings of p, of c and
//     of the sweep's residual (|x - mu_h| <= |p| + |mu_h - mu_a|) give eps_g; g_j already carries
//     eps_g / eps_a, so e_j <= eps_a a'_j and the bound on q_h follows as in sub_screen, with its
//     second-order part eps_a^2 sum_j a'_j^2 / 2 taken per point: ||a'|| <= ||g|| + ||W_hi||_F ||p_hi||
//     (triangle inequality, Cauchy-Schwarz per row) and ||p_hi|| <= (1 + 2^-8) ||p||.
// Returns q_a; sets d2 = ||p||^2 (FP32), Aa, and sd = s_0 - s_1 with its bound sderr.
typedef unsigned int mx_u32x4 __attribute__((ext_vector_type(4)));
// The split p = p_hi + p_lo (BF16 pairs packed in 32-bit words) and dd = sum p^2 of this lane's coordinates.
template <int DP>
__device__ __forceinline__ void split_residual(const float (&x)[DP / 16][8], const unsigned char* sa, int lane,
                                               mx_u32x4 (&ph)[DP / 16], mx_u32x4 (&pl)[DP / 16], float& dd) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    constexpr int KC = DP / 16, TB = mx_tiles_bytes(DP);
    const int h = lane >> 5;
    const float4* mu4 = reinterpret_cast<const float4*>(sa + TB);
    dd = 0.f;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
        const float4 ma = mu4[4 * kc + 2 * h], mb = mu4[4 * kc + 2 * h + 1];
        const float mv[8] = {ma.x, ma.y, ma.z, ma.w, mb.x, mb.y, mb.z, mb.w};
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            const float d0 = x[kc][e] - mv[e];
            const float d1 = x[kc][e + 1] - mv[e + 1];
            dd = fmaf(d0, d0, dd);
            dd = fmaf(d1, d1, dd);
            bf16x2 hp;
            hp[0] = (__bf16)d0;
            hp[1] = (__bf16)d1;
            const uint32_t hb = __builtin_bit_cast(uint32_t, hp);
            bf16x2 lp;
            lp[0] = (__bf16)(d0 - __uint_as_float(hb << 16));
            lp[1] = (__bf16)(d1 - __uint_as_float(hb & 0xffff0000u));
            ph[kc][e / 2] = hb;
            pl[kc][e / 2] = __builtin_bit_cast(uint32_t, lp);
        }
    }
}

template <int DP>
__device__ __forceinline__ float fused_screens_split(const mx_u32x4 (&ph)[DP / 16], const mx_u32x4 (&pl)[DP / 16],
                                                     float dd, const unsigned char* sa, const unsigned char* ssub,
                                                     int lane, float& d2, float& Aa, float& sd, float& sderr) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    constexpr int RB = DP / 32, KC = DP / 16, TB = mx_tiles_bytes(DP), SLOT2 = mx2_slot_bytes(DP);
    const int h = lane >> 5;
    // anchor
    const bf16x8* ta = reinterpret_cast<const bf16x8*>(sa);
    float q = 0.f;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        f32x16 acc = {};
#pragma unroll
        for (int kc = 2 * rb; kc < KC; ++kc)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ta[(rb * KC + kc) * 64 + lane],
                                                          __builtin_bit_cast(bf16x8, ph[kc]), acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 16; ++v) q = fmaf(acc[v], acc[v], q);
    }
    d2 = dd + __shfl_xor(dd, 32);
    Aa = reinterpret_cast<const float*>(sa + TB)[DP + 1] * d2 * (1.0f + ((float)DP + 4.0f) * 0x1p-24f);
    // the anchor's screen is finished here, before the sub-clusters' loop (q materialised): sunk past
    // it, its operands would be live through the loop and spill
    asm volatile("" : "+v"(q));
    // the anchor's sub-clusters
    constexpr float eps = sub_rel_eps_a(DP);
    const float heps = 0.5f * eps;
    float s0 = 0.f, s1 = 0.f, dq0 = 0.f, dq1 = 0.f;
#pragma unroll 1
    for (int hh = 0; hh < 2; ++hh) {
        const unsigned char* slot = ssub + hh * SLOT2;
        const bf16x8* thi = reinterpret_cast<const bf16x8*>(slot);
        const bf16x8* tlo = reinterpret_cast<const bf16x8*>(slot + TB);
        const float* rel = reinterpret_cast<const float*>(slot + mx2_rel_offset(DP));   // -c (DP), then g (DP)
        float qq = 0.f, ee = 0.f;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            // accumulators start from -c_j and g_j of this lane's 16 output rows 32 rb + 8 (v / 4) + 4 h + v % 4
            f32x16 y, a;
#pragma unroll
            for (int g4 = 0; g4 < 4; ++g4) {
                const float4 cv = *reinterpret_cast<const float4*>(rel + 32 * rb + 8 * g4 + 4 * h);
                const float4 gv = *reinterpret_cast<const float4*>(rel + DP + 32 * rb + 8 * g4 + 4 * h);
                y[4 * g4] = cv.x; y[4 * g4 + 1] = cv.y; y[4 * g4 + 2] = cv.z; y[4 * g4 + 3] = cv.w;
                a[4 * g4] = gv.x; a[4 * g4 + 1] = gv.y; a[4 * g4 + 2] = gv.z; a[4 * g4 + 3] = gv.w;
            }
#pragma unroll
            for (int kc = 2 * rb; kc < KC; ++kc) {
                const bf16x8 wh = thi[(rb * KC + kc) * 64 + lane];
                const bf16x8 wl = tlo[(rb * KC + kc) * 64 + lane];
                const bf16x8 bh = __builtin_bit_cast(bf16x8, ph[kc]), bl = __builtin_bit_cast(bf16x8, pl[kc]);
                const u32x4 wabs = __builtin_bit_cast(u32x4, wh) & 0x7fff7fffu;
                const u32x4 pabs = ph[kc] & 0x7fff7fffu;
                y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh, bh, y, 0, 0, 0);
                y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh, bl, y, 0, 0, 0);
                y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl, bh, y, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wabs),
                                                            __builtin_bit_cast(bf16x8, pabs), a, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                qq = fmaf(y[v], y[v], qq);
                ee = fmaf(a[v], fabsf(y[v]), ee);
            }
        }
        qq += __shfl_xor(qq, 32);
        // sum_j a'_j^2 <= (||g|| + ||W_hi||_F ||p_hi||)^2, ||p||^2 <= d2 (1 + (DP + 4) 2^-24); 2^-10 covers
        // this line's rounding and the MFMA's of a'
        const float* msub = reinterpret_cast<const float*>(slot + 2 * TB);
        const float an = msub[DP + 1] + msub[DP + 2] * sqrtf(d2 * (1.0f + ((float)DP + 4.0f) * 0x1p-24f)) * (1.0f + 0x1p-7f);
        ee = 2.0f * eps * (ee + __shfl_xor(ee, 32) + heps * an * an * (1.0f + 0x1p-10f));
        const float sv = msub[DP] - 0.5f * qq;
        const float dv = ee + (2.0f * (float)DP + 4.0f) * 0x1p-24f * qq;   // the two sums of squares
        if (hh == 0) { s0 = sv; dq0 = dv; } else { s1 = sv; dq1 = dv; }
    }
    sderr = 0.5f * (dq0 + dq1) * (1.0f + 0x1p-10f) + 0x1p-21f * (fabsf(s0) + fabsf(s1)) + 1e-6f;
    sd = s0 - s1;
    return q + __shfl_xor(q, 32);
}

template <int DP>
__device__ __forceinline__ float fused_screens(const float (&x)[DP / 16][8], const unsigned char* sa,
                                               const unsigned char* ssub, int lane, float& d2, float& Aa,
                                               float& sd, float& sderr) {
    mx_u32x4 ph[DP / 16], pl[DP / 16];
    float dd;
    split_residual<DP>(x, sa, lane, ph, pl, dd);
    return fused_screens_split<DP>(ph, pl, dd, sa, ssub, lane, d2, Aa, sd, sderr);
}

// split_score with each 8-coordinate chunk of the point read where it is used (from cache, row: the
// point's row, ok: a real point) and split on the fly, so that neither the coordinates nor their
// split stay in registers: the same arithmetic in the same order, for the rare refinement path.
template <int DP>
__device__ __forceinline__ float split_score_lean(gfloat* row, int D, bool ok, const bf16x8* thi, const bf16x8* tlo,
                                                  const float* mu, float c, int lane, float& dq) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    constexpr int RB = DP / 32, KC = DP / 16;
    const int h = lane >> 5;
    constexpr float eps = split_screen_eps(DP);
    const float heps = 0.5f * eps;
    float qq = 0.f, ee = 0.f;
#pragma unroll 1
    for (int rb = 0; rb < RB; ++rb) {
        f32x16 y = {}, a = {};
#pragma unroll 1
        for (int kc = 2 * rb; kc < KC; ++kc) {
            float xv[8];
            load_coords8_aligned(row, 16 * kc + 8 * h, D, ok, xv);
            u32x4 dh, dl;
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                const float d0 = xv[e] - mu[16 * kc + 8 * h + e];
                const float d1 = xv[e + 1] - mu[16 * kc + 8 * h + e + 1];
                bf16x2 hp;
                hp[0] = (__bf16)d0;
                hp[1] = (__bf16)d1;
                const uint32_t hb = __builtin_bit_cast(uint32_t, hp);
                bf16x2 lp;
                lp[0] = (__bf16)(d0 - __uint_as_float(hb << 16));
                lp[1] = (__bf16)(d1 - __uint_as_float(hb & 0xffff0000u));
                dh[e / 2] = hb;
                dl[e / 2] = __builtin_bit_cast(uint32_t, lp);
            }
            const bf16x8 wh = thi[(rb * KC + kc) * 64 + lane];
            const bf16x8 wl = tlo[(rb * KC + kc) * 64 + lane];
            const u32x4 wabs = __builtin_bit_cast(u32x4, wh) & 0x7fff7fffu;
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh, __builtin_bit_cast(bf16x8, dh), y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wh, __builtin_bit_cast(bf16x8, dl), y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wl, __builtin_bit_cast(bf16x8, dh), y, 0, 0, 0);
            a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, wabs),
                                                        __builtin_bit_cast(bf16x8, dh & 0x7fff7fffu), a, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            qq = fmaf(y[v], y[v], qq);
            ee = fmaf(a[v], fmaf(heps, a[v], fabsf(y[v])), ee);
        }
    }
    qq += __shfl_xor(qq, 32);
    ee = 2.0f * eps * (ee + __shfl_xor(ee, 32));
    dq = ee + (2.0f * (float)DP + 4.0f) * 0x1p-24f * qq;
    return c - 0.5f * qq;
}

// The sub-label screen fused into the pruned sweep.  A point whose label stays its chunk's cluster
// (the anchor) is screened against the anchor's two sub-clusters in the same pass, with the
// arithmetic and bound of k_sub_gauss_mx; the points whose sub-label the screen leaves undecided, and
// those whose label moved or went to the recheck, are listed (positions in the swept partition) for
// the FP32 draw after the label recheck (k_sub_gauss_seg): they are listed inside the anchor's segment
// of the swept partition, the first kind from its start up (count seg[a]), the second from its end
// down (count seg[K + a]).  skip_same (both forms): every label of a chunk equals
// its anchor before the sweep, so a label that stays is not written again.  (Extra arguments of
// k_label_gauss_mxp.)
struct MxpExtra {
    const unsigned char* SPB;    // 2K sub-cluster screen slots (split-BF16), null: no fused screen
    const int32_t* subs_prev;    // sub-labels before the sweep in the swept partition's order, or null
    int32_t* subs;               // sub-labels by point index, written where they change
    int32_t* subs_old;           // sub-labels in the swept partition's order
    int32_t* sub_list;           // positions left to the FP32 sub-label draw (N entries)
    int* sub_count;              // how many (written by k_sub_gauss_seg)
    int32_t* seg;                // 2K list lengths: per cluster, listed points whose label stayed, then the others
    int skip_same;
    int split;                   // workgroups per partition chunk (the grid is chunks x split)
    const unsigned char* PBlo;   // the clusters' lo tiles (refinement of undecided waves), or null
};

// Workgroups per partition chunk in the pruned sweep (each takes an equal share of the chunk): more
// workgroups shorten the tail, fewer stage the chunk's slots into LDS fewer times.
constexpr int kPruneSplit = 4;

template <int DP, bool FUSE>
__global__ void __launch_bounds__(kThreads, DP <= 64 ? 3 : 2)
k_label_gauss_mxp(const float* __restrict__ X, int D, int K, int CH, const int32_t* __restrict__ perm,
                  const int32_t* __restrict__ label_start, const int32_t* __restrict__ chunk_start,
                  const unsigned char* __restrict__ PB, const float* __restrict__ pm,
                  const float* __restrict__ psc, float tau, uint32_t key0, uint32_t key1, uint32_t iter,
                  int final_iter, int32_t* __restrict__ labels, int32_t* __restrict__ recheck,
                  int* __restrict__ recheck_count, unsigned long long* __restrict__ evals,
                  int* __restrict__ changed, MxpExtra sf) {
    static_assert(!FUSE || DP <= 64, "the fused sub-label screen keeps three slots in LDS");
    constexpr int RB = DP / 32, KC = DP / 16;
    constexpr int SLOT = mx_slot_bytes(DP);
    // one 32-point block per wave: lane L scores point L % 32 on coordinate half L / 32
    const int lane = threadIdx.x & 63, h = lane >> 5, pl = lane & 31, wv = threadIdx.x >> 6;
    int a, cbeg, cend;
    if (!chunk_of(blockIdx.x / sf.split, K, chunk_start, label_start, CH, a, cbeg, cend)) return;
    const int part = blockIdx.x % sf.split, len = cend - cbeg;
    const int beg = cbeg + (int)((int64_t)len * part / sf.split);
    const int end = cbeg + (int)((int64_t)len * (part + 1) / sf.split);
    const float* mrow = pm + (size_t)a * K;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // the anchor's slot, scored by every pass of the workgroup, is staged in LDS once; the other
    // clusters' slots are read from global memory by the few passes that score them
    __shared__ __align__(16) unsigned char sa[SLOT];
    // fused: the anchor's two sub-cluster slots as well (the sub-label screen of the points that stay)
    constexpr int SLOT2 = mx2_slot_bytes(DP);
    __shared__ __align__(16) unsigned char ssub[FUSE ? 2 * SLOT2 : 16];
    {
        const uint4* src = reinterpret_cast<const uint4*>(PB + (size_t)a * SLOT);
        uint4* dst = reinterpret_cast<uint4*>(sa);
        for (int e = threadIdx.x; e < SLOT / 16; e += kThreads) dst[e] = src[e];
        if constexpr (FUSE) {
            const uint4* s2 = reinterpret_cast<const uint4*>(sf.SPB + (size_t)(2 * a) * SLOT2);
            uint4* d2 = reinterpret_cast<uint4*>(ssub);
            for (int e = threadIdx.x; e < 2 * SLOT2 / 16; e += kThreads) d2[e] = s2[e];
        }
    }
    __syncthreads();
    // the anchor's bound inputs of clusters 0..63, one per lane (the same for every pass)
    // (in LDS, not held in registers through the passes)
    __shared__ float bnd0[3][64];
    if (threadIdx.x < 64) {
        const bool in0 = lane < K && lane != a;
        bnd0[0][lane] = in0 ? psc[2 * lane] : 0.f;
        bnd0[1][lane] = in0 ? psc[2 * lane + 1] : 0.f;
        bnd0[2][lane] = in0 ? mrow[lane] : 0.f;
    }
    __syncthreads();
    unsigned long long nev = 0, npass = 0;
    __shared__ int32_t wlist[4][64];
    WaveList rl{wlist[wv], 0};
    __shared__ int32_t wlist2[FUSE ? 4 : 1][64];
    WaveList sl{wlist2[FUSE ? wv : 0], 0};       // fused: undecided sub-labels of points that stay
    __shared__ int32_t wlist3[FUSE ? 4 : 1][64];
    WaveList ml{wlist3[FUSE ? wv : 0], 0};       // fused: points whose label moved or is rechecked

    // the coordinates 16 kc + 8 h .. + 7 of point `pt` (row 0 stands in for an empty lane)
    auto load_row = [&](gfloat* Xs, int pt, bool ok, float (&x)[KC][8]) {
        gfloat* row = Xs + (int64_t)pt * D;
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) load_coords8_aligned(row, 16 * kc + 8 * h, D, ok, x[kc]);   // D % 4 == 0
    };
    int idx_next = (beg + wv * 32 + pl < end) ? perm[beg + wv * 32 + pl] : 0;
    for (int p0 = beg + wv * 32; p0 < end; p0 += 4 * 32) {   // wave-uniform trip count
        const int pos = p0 + pl;
        const bool val = pos < end;
        const int idx = val ? idx_next : 0;
        // the next pass's point, loaded ahead; past the end the load repeats the last entry (a
        // branch-free load, so that waiting for it is left to its first use in the next pass)
        idx_next = perm[pos + 128 < end ? pos + 128 : end - 1];
        const bool own_valid = val && h == 0;
        auto load_x = [&](gfloat* Xs, float (&x)[KC][8]) { load_row(Xs, idx, val, x); };
        // up to 32 dimensions the coordinates stay in registers for the whole pass; wider points are
        // read again by each scoring (from cache after the anchor's), which keeps the registers for
        // the MFMA operands
        constexpr bool kHoldX = DP <= 32;
        float xh[kHoldX ? KC : 1][8];
        if constexpr (kHoldX) load_x((gfloat*)X, xh);
        // screened q of the cluster whose slot is `slot` (the anchor's in LDS, or one in global
        // memory) for this lane's point with coordinates x, sum (x - mu_k)^2 and the bound's sum ee
        auto score_x = [&](const unsigned char* slot, const float (&x)[KC][8], float& d2, float& eo) -> float {
            const bf16x8* tiles = reinterpret_cast<const bf16x8*>(slot);
            const float4* mu4 = reinterpret_cast<const float4*>(slot + mx_tiles_bytes(DP));
            bf16x8 Bv[KC];
            float dd = 0.f;
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                const float4 ma = mu4[4 * kc + 2 * h], mb = mu4[4 * kc + 2 * h + 1];
                const float mv[8] = {ma.x, ma.y, ma.z, ma.w, mb.x, mb.y, mb.z, mb.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float d = x[kc][e] - mv[e];
                    Bv[kc][e] = (__bf16)d;
                    dd = fmaf(d, d, dd);
                }
            }
            float q = 0.f, ee = 0.f;
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                f32x16 acc = {}, aa = {};
#pragma unroll
                for (int kc = 2 * rb; kc < KC; ++kc) {
                    const bf16x8 wt = tiles[(rb * KC + kc) * 64 + lane];
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wt, Bv[kc], acc, 0, 0, 0);
                    bf16x8 bab = bf16x8_abs(Bv[kc]);
                    // at DP = 128 the magnitudes are formed at each use: kept for all row blocks they
                    // would not fit in registers with the operands
                    if constexpr (DP > 96) asm volatile("" : "+v"(bab));
                    aa = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bf16x8_abs(wt), bab, aa, 0, 0, 0);
                }
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    q = fmaf(acc[v], acc[v], q);
                    ee = fmaf(aa[v], aa[v], ee);
                }
            }
            d2 = dd + __shfl_xor(dd, 32);
            eo = ee + __shfl_xor(ee, 32);
            return q + __shfl_xor(q, 32);
        };
        auto score = [&](const unsigned char* slot, float& d2, float& eo) -> float {
            if constexpr (kHoldX) {
                return score_x(slot, xh, d2, eo);
            } else {
                float xr[KC][8];
                gfloat* Xs = (gfloat*)X;
                asm volatile("" : "+s"(Xs));   // a load per scoring, not one hoisted out of the cluster loop
                load_x(Xs, xr);
                return score_x(slot, xr, d2, eo);
            }
        };
        auto noisy = [&](const unsigned char* slot, int k, float q) -> float {
            const float c = reinterpret_cast<const float*>(slot + mx_tiles_bytes(DP))[DP];
            float s = c - 0.5f * q;
            if (!final_iter) {
                const Ctr4 w = philox4x32_10((uint32_t)idx, (uint32_t)(k >> 2), iter, TAG_LABEL, key0, key1);
                s += gumbel_screen(sel4(w, k & 3));
            }
            return s;
        };
        // the anchor, then the wave bounds: r >= ||x - mu_a|| over the points (the FP32 rounding of
        // the distance covered) and T = the smallest screened lower bound s_a - e_a
        float d2 = 0.f, ea = 0.f, qa;
        // fused: the screen of the anchor's sub-clusters for every point, on the coordinates the
        // anchor's scoring read (the few points whose label moves ignore it)
        float sd = 0.f, sderr = 0.f;
        if constexpr (FUSE && !kHoldX) {
            float xa[KC][8];
            load_x((gfloat*)X, xa);
            qa = fused_screens<DP>(xa, sa, ssub, lane, d2, ea, sd, sderr);
        } else if constexpr (FUSE) {
            qa = fused_screens<DP>(xh, sa, ssub, lane, d2, ea, sd, sderr);
        } else {
            qa = score(sa, d2, ea);
        }
        const float* ca = reinterpret_cast<const float*>(sa + mx_tiles_bytes(DP)) + DP;
        float sb = *ca - 0.5f * qa;
        // the anchor's label noise and, fused, the sub-label draw's word come from one Philox
        // evaluation per lane: half 0 draws the former, half 1 (the same points) the latter
        uint32_t sub_word = 0;
        if (FUSE || !final_iter) {
            const bool sub_half = FUSE && h;
            const Ctr4 w = philox4x32_10((uint32_t)idx, sub_half ? 0u : (uint32_t)(a >> 2), iter,
                                         sub_half ? TAG_SUB : TAG_LABEL, key0, key1);
            if constexpr (FUSE) {
                const uint32_t mine = h ? w.v[0] : sel4(w, a & 3);
                const uint32_t other = __shfl_xor(mine, 32);
                if (!final_iter) sb += gumbel_screen(h ? other : mine);
                sub_word = h ? mine : other;
            } else {
                sb += gumbel_screen(sel4(w, a & 3));
            }
        }
        float eb = label_screen_err<DP>(tau, qa, ea, *ca), U = neg_inf<float>();
        float Up = neg_inf<float>();   // the largest bound of the pruned clusters
        int bi = a;
        float rr = own_valid ? sqrtf(d2) : 0.f;
        float tt = own_valid ? sb - eb : __builtin_huge_valf();
        for (int o = 32; o > 0; o >>= 1) {
            rr = fmaxf(rr, __shfl_xor(rr, o));
            tt = fminf(tt, __shfl_xor(tt, o));
        }
        const float rad = rr * (1.0f + (float)(D + 8) * 0x1p-23f);
        const float T = tt;
        unsigned long long ne = 1;
        // 64 clusters per step, one per lane: prune test, then the survivors in index order
        for (int k0 = 0; k0 < K; k0 += 64) {
            const int kk = k0 + lane;
            float ub = neg_inf<float>();
            bool keep = false;
            if (kk < K && kk != a) {
                const float sg = k0 == 0 ? bnd0[0][lane] : psc[2 * kk], ck = k0 == 0 ? bnd0[1][lane] : psc[2 * kk + 1];
                const float lb = fmaxf(0.f, (k0 == 0 ? bnd0[2][lane] : mrow[kk]) - sg * rad);
                ub = ck - 0.49f * lb * lb + 1e-6f * fabsf(ck) + kPruneNoise;
                keep = !(ub < T);
            }
            float upr = keep ? neg_inf<float>() : ub;   // bound of the skipped clusters
            for (int o = 32; o > 0; o >>= 1) upr = fmaxf(upr, __shfl_xor(upr, o));
            U = fmaxf(U, upr);
            Up = fmaxf(Up, upr);
            uint64_t todo = __ballot(keep);
            while (todo) {
                const int k = k0 + __ffsll((unsigned long long)todo) - 1;
                todo &= todo - 1;
                ++ne;
                float unused, ek;
                const unsigned char* slot = PB + (size_t)k * SLOT;
                const float qk = score(slot, unused, ek);
                const float s = noisy(slot, k, qk);
                const float err =
                    label_screen_err<DP>(tau, qk, ek, reinterpret_cast<const float*>(slot + mx_tiles_bytes(DP))[DP]);
                if (s > sb) {
                    U = fmaxf(U, sb + eb);
                    sb = s; eb = err; bi = k;
                } else {
                    U = fmaxf(U, s + err);
                }
            }
        }
        nev += ne;
        ++npass;
        bool sure = sb - eb > U;
        // a wave with undecided points is scored again in split BF16 (split_score_lean, bounds some 2^-7 of the
        // BF16 screen's); a point whose best refined lower bound beats every other refined upper bound
        // and the pruned clusters' bound Up takes that label (the FP32 sweep's: its score is the strict
        // maximum).  Siblings of a split or an over-split start leave most points within the BF16
        // bound of a tie.
        if (sf.PBlo && __ballot(own_valid && !sure)) {
            gfloat* Xs = (gfloat*)X;
            asm volatile("" : "+s"(Xs));   // coordinates read again (from cache), not kept from the screen
            gfloat* row = Xs + (int64_t)idx * D;
            float rs = neg_inf<float>(), re = 0.f, rU = Up;
            int rbi = a;
            auto refine = [&](int k, const unsigned char* slot) {
                const float* mu = reinterpret_cast<const float*>(slot + mx_tiles_bytes(DP));
                float dq;
                float sk = split_score_lean<DP>(row, D, val, reinterpret_cast<const bf16x8*>(slot),
                                                reinterpret_cast<const bf16x8*>(sf.PBlo + (size_t)k * mx_tiles_bytes(DP)),
                                                mu, mu[DP], lane, dq);
                const float e = 0.5f * dq * (1.0f + 0x1p-10f) + 0x1p-21f * (fabsf(sk) + 24.0f) + kNoiseSlack;
                if (!final_iter) {
                    const Ctr4 w = philox4x32_10((uint32_t)idx, (uint32_t)(k >> 2), iter, TAG_LABEL, key0, key1);
                    sk += gumbel_screen(sel4(w, k & 3));
                }
                if (sk > rs) {
                    rU = fmaxf(rU, rs + re);
                    rs = sk; re = e; rbi = k;
                } else {
                    rU = fmaxf(rU, sk + e);
                }
            };
            refine(a, sa);
            for (int k0 = 0; k0 < K; k0 += 64) {   // the kept clusters again (the same test, in index order)
                const int kk = k0 + lane;
                bool keep = false;
                if (kk < K && kk != a) {
                    const float ck = psc[2 * kk + 1];
                    const float lb = fmaxf(0.f, mrow[kk] - psc[2 * kk] * rad);
                    keep = !(ck - 0.49f * lb * lb + 1e-6f * fabsf(ck) + kPruneNoise < T);
                }
                uint64_t todo = __ballot(keep);
                while (todo) {
                    const int k = k0 + __ffsll((unsigned long long)todo) - 1;
                    todo &= todo - 1;
                    refine(k, PB + (size_t)k * SLOT);
                }
            }
            if (!sure && rs - re > rU) {
                sure = true;
                bi = rbi;
            }
        }
        // with skip_same every label in the chunk is the anchor before the sweep: a label that stays is not
        // written again
        if (own_valid && sure && !(sf.skip_same && bi == a)) labels[idx] = bi;
        if constexpr (FUSE) {
            // sub-labels of the points whose label stays the anchor, from its sub-clusters' screen; the others
            // (labels moved or left to the recheck, undecided sub-labels) go to the sub list
            const bool stay = own_valid && sure && bi == a;
            bool undecided = false;
            if (stay) {
                const float lg = logit_u32<float>(sub_word);
                if (fabsf(lg - sd) > sderr) {
                    const int v = lg < sd ? 0 : 1;
                    sf.subs_old[pos] = v;
                    if (!sf.subs_prev || sf.subs_prev[pos] != v) sf.subs[idx] = v;
                } else {
                    undecided = true;
                }
            }
            sl.push(undecided, pos, sf.sub_list + label_start[a], sf.seg + a, lane, lt);
            ml.push(own_valid && !stay, pos, sf.sub_list + label_start[a + 1] - 1, sf.seg + K + a, lane, lt, -1);
        }
        // a label that moved off the chunk's cluster: the partition of the labels before the sweep no longer
        // stands (changed[0]), and the statistics of both clusters are recomputed in full (changed[1 + k])
        if (changed) {
            const bool mv = own_valid && sure && bi != a;
            if (mv) {
                changed[1 + a] = 1;
                changed[1 + bi] = 1;
            }
            if (__ballot(mv) && lane == 0) changed[0] = 1;
        }
        rl.push(own_valid && !sure, (int32_t)idx, recheck, recheck_count, lane, lt);
    }
    rl.flush(recheck, recheck_count, lane);
    if constexpr (FUSE) {
        sl.flush(sf.sub_list + label_start[a], sf.seg + a, lane);
        ml.flush(sf.sub_list + label_start[a + 1] - 1, sf.seg + K + a, lane, -1);
    }
    __shared__ unsigned long long cnt[2][4];   // one pair of counter atomics per workgroup
    if (lane == 0) {
        cnt[0][wv] = nev;
        cnt[1][wv] = npass;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long e = cnt[0][0] + cnt[0][1] + cnt[0][2] + cnt[0][3];
        const unsigned long long pp = cnt[1][0] + cnt[1][1] + cnt[1][2] + cnt[1][3];
        if (pp) {
            atomicAdd(evals, e);
            atomicAdd(evals + 1, pp);
        }
    }
}

// sf.SPB non-null (DP <= 64 only): the sub-label screen fused into the sweep.
template <int DP>
void launch_label_gauss_mxp(hipStream_t st, int grid, const float* X, int D, int K, int CH, const int32_t* perm,
                            const int32_t* label_start, const int32_t* chunk_start, const unsigned char* PB,
                            const float* pm, const float* psc, float tau, uint32_t key0, uint32_t key1, uint32_t iter,
                            int fin, int32_t* labels, int32_t* recheck, int* recheck_count,
                            unsigned long long* evals, int* changed, MxpExtra sf) {
    if constexpr (DP <= 64) {
        if (sf.SPB) {
            hipLaunchKernelGGL((k_label_gauss_mxp<DP, true>), dim3(grid), dim3(kThreads), 0, st, X, D, K, CH, perm,
                               label_start, chunk_start, PB, pm, psc, tau, key0, key1, iter, fin, labels, recheck,
                               recheck_count, evals, changed, sf);
            return;
        }
    }
    hipLaunchKernelGGL((k_label_gauss_mxp<DP, false>), dim3(grid), dim3(kThreads), 0, st, X, D, K, CH, perm,
                       label_start, chunk_start, PB, pm, psc, tau, key0, key1, iter, fin, labels, recheck,
                       recheck_count, evals, changed, sf);
}

#define DPMM_MX_INSTANTIATE(DP)                                                                               \
    template void launch_label_gauss_mx<DP>(hipStream_t, int64_t, int, int, const float*, const unsigned char*, \
                                            float, uint32_t, uint32_t, uint32_t, int, int32_t*, int32_t*, int*); \
    template void launch_sub_gauss_mx<DP>(hipStream_t, int, const float*, int, int, int, const int32_t*,       \
                                          const int32_t*, const int32_t*, const unsigned char*, float, uint32_t, \
                                          uint32_t, uint32_t, int32_t*, int32_t*, int32_t*, int*);             \
    template void launch_label_gauss_mxp<DP>(hipStream_t, int, const float*, int, int, int, const int32_t*,   \
                                             const int32_t*, const int32_t*, const unsigned char*, const float*, \
                                             const float*, float, uint32_t, uint32_t, uint32_t, int, int32_t*,   \
                                             int32_t*, int*, unsigned long long*, int*, MxpExtra);
#define DPMM_MX_EXTERN(DP)                                                                                    \
    extern template void launch_label_gauss_mx<DP>(hipStream_t, int64_t, int, int, const float*,               \
                                                   const unsigned char*, float, uint32_t, uint32_t, uint32_t,  \
                                                   int, int32_t*, int32_t*, int*);                             \
    extern template void launch_sub_gauss_mx<DP>(hipStream_t, int, const float*, int, int, int, const int32_t*, \
                                                 const int32_t*, const int32_t*, const unsigned char*, float,   \
                                                 uint32_t, uint32_t, uint32_t, int32_t*, int32_t*, int32_t*,    \
                                                 int*);                                                         \
    extern template void launch_label_gauss_mxp<DP>(hipStream_t, int, const float*, int, int, int,             \
                                                    const int32_t*, const int32_t*, const int32_t*,            \
                                                    const unsigned char*, const float*, const float*, float,   \
                                                    uint32_t, uint32_t, uint32_t, int, int32_t*, int32_t*, int*, \
                                                    unsigned long long*, int*, MxpExtra);

}  // namespace dpmm
