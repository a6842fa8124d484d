// Templated Gaussian kernels (label sweep, sub-label sweep, FP64 statistics) and their
// launchers; instantiated per padded dimension DP in inst_gauss_*.hip so the heavily unrolled
// bodies compile in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "dpmm_kernels.h"

namespace dpmm {

// Gaussian label sweep of one point per thread (or per lane pair for wide points), the point in
// registers; cluster parameters staged through LDS one cluster at a time.  Every thread of the
// workgroup must call it (it synchronises); threads without a point pass i >= N.
template <typename T, int DP>
__device__ __forceinline__ int label_point(const T* __restrict__ X, int64_t N, int D, int K, const T* __restrict__ P,
                                           int64_t i, int par, T* ps, uint32_t key0, uint32_t key1, uint32_t iter,
                                           int final_iter) {
    constexpr bool kPair = paired_layout<T, DP>();
    constexpr int kCols = kPair ? DP / 2 : DP;
    constexpr int stride = gauss_slot_stride(DP);
    T x[kCols];
#pragma unroll
    for (int c = 0; c < kCols; ++c) {
        const int l = kPair ? pair_col(c, 0) + (par << 2) : c;
        x[c] = (i < N && l < D) ? X[i * D + l] : T(0);
    }
    T best = neg_inf<T>();
    int bi = 0;
    Ctr4 w = {};
    for (int k = 0; k < K; ++k) {
        __syncthreads();
        const T* src = P + (size_t)k * stride;
        for (int e = threadIdx.x; e < stride; e += kThreads) ps[e] = src[e];
        __syncthreads();
        T q;
        if constexpr (kPair) q = whitened_sq_pair<T, DP>(ps, x, par);
        else q = whitened_sq_regs<T, DP>(ps, x);
        T s = ps[gauss_c_offset(DP)] - T(0.5) * q;
        if (!final_iter) {
            if ((k & 3) == 0) w = philox4x32_10((uint32_t)i, (uint32_t)(k >> 2), iter, TAG_LABEL, key0, key1);
            s += gumbel<T>(w.v[0]);
            rotate_words(w);
        }
        if (s > best || k == 0) { best = s; bi = k; }
    }
    return bi;
}

template <typename T, int DP>
__global__ void __launch_bounds__(kThreads) k_label_gauss(const T* __restrict__ X, int64_t N, int D, int K,
                                                          const T* __restrict__ P, uint32_t key0, uint32_t key1,
                                                          uint32_t iter, int final_iter, int32_t* __restrict__ labels) {
    constexpr bool kPair = paired_layout<T, DP>();
    constexpr int kPts = kPair ? kThreads / 2 : kThreads;
    __shared__ __align__(16) T ps[gauss_slot_stride(DP)];
    const int par = kPair ? (threadIdx.x & 1) : 0;
    const int64_t i = (int64_t)blockIdx.x * kPts + (kPair ? (threadIdx.x >> 1) : threadIdx.x);
    const int bi = label_point<T, DP>(X, N, D, K, P, i, par, ps, key0, key1, iter, final_iter);
    if (i < N && par == 0) labels[i] = bi;
}

// The same FP32/FP64 sweep over a list of point indices (the points the BF16 screen could not
// decide); workgroups stride over the list, whose length is read from device memory.
template <typename T, int DP>
__global__ void __launch_bounds__(kThreads) k_label_gauss_list(const T* __restrict__ X, int64_t N, int D, int K,
                                                               const T* __restrict__ P, uint32_t key0, uint32_t key1,
                                                               uint32_t iter, int final_iter,
                                                               const int32_t* __restrict__ list,
                                                               const int* __restrict__ list_count,
                                                               int32_t* __restrict__ labels) {
    constexpr bool kPair = paired_layout<T, DP>();
    constexpr int kPts = kPair ? kThreads / 2 : kThreads;
    __shared__ __align__(16) T ps[gauss_slot_stride(DP)];
    const int par = kPair ? (threadIdx.x & 1) : 0;
    const int local = kPair ? (threadIdx.x >> 1) : threadIdx.x;
    const int n = *list_count;
    for (int b0 = blockIdx.x * kPts; b0 < n; b0 += gridDim.x * kPts) {
        const int pos = b0 + local;
        const int64_t i = pos < n ? (int64_t)list[pos] : N;
        const int bi = label_point<T, DP>(X, N, D, K, P, i, par, ps, key0, key1, iter, final_iter);
        if (i < N && par == 0) labels[i] = bi;
    }
}

// The FP32 re-scoring of listed points with the bounds of the pruned sweep (pm, psc: whitened mean
// distances and spectral norms from k_prune_prep).  Each point's anchor a (its label before the sweep,
// 0 when that lies outside [0, K)) is scored first, with r >= ||x - mu_a||; cluster k is then scored
// only if some point of the workgroup cannot rule it out, i.e. c_k - 0.49 max(0, m_ak - sig_k r)^2
// + 1e-6 |c_k| + kPruneNoise is not below that point's noisy anchor score.  A skipped cluster's FP32
// score lies below the anchor's for every point of the workgroup, so the first maximum in index order
// over the scored clusters is the FP32 sweep's label (k_label_gauss).  Noise words are taken
// directly (word k % 4 of block k / 4), the values k_label_gauss consumes in turn.
template <typename T, int DP>
__global__ void __launch_bounds__(kThreads) k_label_gauss_listp(const T* __restrict__ X, int64_t N, int D, int K,
                                                                const T* __restrict__ P, const float* __restrict__ pm,
                                                                const float* __restrict__ psc, uint32_t key0,
                                                                uint32_t key1, uint32_t iter, int final_iter,
                                                                const int32_t* __restrict__ list,
                                                                const int* __restrict__ list_count,
                                                                int32_t* __restrict__ labels,
                                                                int* __restrict__ changed) {
    constexpr bool kPair = paired_layout<T, DP>();
    constexpr int kPts = kPair ? kThreads / 2 : kThreads;
    constexpr int kCols = kPair ? DP / 2 : DP;
    constexpr int stride = gauss_slot_stride(DP);
    __shared__ __align__(16) T ps[stride];
    const int par = kPair ? (threadIdx.x & 1) : 0;
    const int local = kPair ? (threadIdx.x >> 1) : threadIdx.x;
    const int n = *list_count;
    for (int b0 = blockIdx.x * kPts; b0 < n; b0 += gridDim.x * kPts) {   // uniform trip count
        const int pos = b0 + local;
        const int64_t i = pos < n ? (int64_t)list[pos] : N;
        const bool valid = i < N;
        T x[kCols];
#pragma unroll
        for (int c = 0; c < kCols; ++c) {
            const int l = kPair ? pair_col(c, 0) + (par << 2) : c;
            x[c] = (valid && l < D) ? X[i * D + l] : T(0);
        }
        const int a0 = valid ? labels[i] : 0;   // the label before the sweep
        int a = a0;
        if ((unsigned)a >= (unsigned)K) a = 0;
        const T* pa = P + (size_t)a * stride;
        T qa;
        if constexpr (kPair) qa = whitened_sq_pair<T, DP>(pa, x, par);
        else qa = whitened_sq_regs<T, DP>(pa, x);
        T sa = pa[gauss_c_offset(DP)] - T(0.5) * qa;
        if (!final_iter)
            sa += gumbel<T>(sel4(philox4x32_10((uint32_t)i, (uint32_t)(a >> 2), iter, TAG_LABEL, key0, key1), a & 3));
        float d2 = 0.f;
#pragma unroll
        for (int c = 0; c < kCols; ++c) {
            const int l = kPair ? pair_col(c, 0) + (par << 2) : c;
            const float d = l < D ? (float)(x[c] - pa[gauss_mu_offset(DP) + l]) : 0.f;
            d2 = fmaf(d, d, d2);
        }
        if constexpr (kPair) d2 += __shfl_xor(d2, 1);
        const float r = sqrtf(d2) * (1.0f + (float)(D + 8) * 0x1p-23f);
        const float* mrow = pm + (size_t)a * K;
        T best = neg_inf<T>();
        int bi = 0;
        bool first = true;
        for (int k = 0; k < K; ++k) {
            bool need = false;
            if (valid) {
                if (k == a) {
                    need = true;
                } else {
                    const float ck = psc[2 * k + 1];
                    const float lb = fmaxf(0.f, mrow[k] - psc[2 * k] * r);
                    need = !(ck - 0.49f * lb * lb + 1e-6f * fabsf(ck) + kPruneNoise < (float)sa);
                }
            }
            if (!__syncthreads_or(need)) continue;   // also the barrier after the last cluster's readers
            const T* src = P + (size_t)k * stride;
            for (int e = threadIdx.x; e < stride; e += kThreads) ps[e] = src[e];
            __syncthreads();
            T q;
            if constexpr (kPair) q = whitened_sq_pair<T, DP>(ps, x, par);
            else q = whitened_sq_regs<T, DP>(ps, x);
            T s = ps[gauss_c_offset(DP)] - T(0.5) * q;
            if (!final_iter)
                s += gumbel<T>(sel4(philox4x32_10((uint32_t)i, (uint32_t)(k >> 2), iter, TAG_LABEL, key0, key1), k & 3));
            if (s > best || first) { best = s; bi = k; first = false; }
        }
        if (valid && par == 0) labels[i] = bi;
        // a label that moved: the partition of the labels before the sweep no longer stands (changed[0]),
        // and the statistics of both clusters are recomputed in full (changed[1 + k]); the cluster it
        // left is the one the swept partition filed it under (a, 0 for a label outside [0, K))
        if (changed) {
            const bool mv = valid && par == 0 && bi != a0;
            if (mv) {
                changed[1 + bi] = 1;
                changed[1 + a] = 1;
            }
            if (__ballot(mv) && (threadIdx.x & 63) == 0) changed[0] = 1;
        }
    }
}

template <typename T, int DP>
__global__ void __launch_bounds__(kThreads) k_sub_gauss(const T* __restrict__ X, int D, int K, int CH,
                                                        const int32_t* __restrict__ perm,
                                                        const int32_t* __restrict__ label_start,
                                                        const int32_t* __restrict__ chunk_start,
                                                        const T* __restrict__ SP, uint32_t key0, uint32_t key1,
                                                        uint32_t iter, int32_t* __restrict__ subs) {
    constexpr bool kPair = paired_layout<T, DP>();
    constexpr int kPts = kPair ? kThreads / 2 : kThreads;
    constexpr int kCols = kPair ? DP / 2 : DP;
    constexpr int stride = gauss_slot_stride(DP);
    __shared__ __align__(16) T ps[2 * stride];
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end)) return;
    const T* src = SP + (size_t)(2 * k) * stride;
    for (int e = threadIdx.x; e < 2 * stride; e += kThreads) ps[e] = src[e];
    __syncthreads();
    const int par = kPair ? (threadIdx.x & 1) : 0;
    const int lane_pt = kPair ? (threadIdx.x >> 1) : threadIdx.x;
    // every lane of a pair runs the loop the same number of times (the DPP swap needs both)
    for (int p0 = beg; p0 < end; p0 += kPts) {
        const int p = p0 + lane_pt;
        const bool valid = p < end;
        const int idx = valid ? perm[p] : 0;
        T x[kCols];
#pragma unroll
        for (int c = 0; c < kCols; ++c) {
            const int l = kPair ? pair_col(c, 0) + (par << 2) : c;
            x[c] = (valid && l < D) ? X[(int64_t)idx * D + l] : T(0);
        }
        T sc[2];
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {   // one sub-cluster at a time keeps a single residual live
            const T* pp = ps + h * stride;
            T q;
            if constexpr (kPair) q = whitened_sq_pair<T, DP>(pp, x, par);
            else q = whitened_sq_regs<T, DP>(pp, x);
            sc[h] = pp[gauss_c_offset(DP)] - T(0.5) * q;
        }
        if (valid && par == 0) {
            const uint32_t w = philox4x32_10((uint32_t)idx, 0, iter, TAG_SUB, key0, key1).v[0];
            subs[idx] = (logit_u32<T>(w) < sc[0] - sc[1]) ? 0 : 1;
        }
    }
}

// Statistics of both sub-clusters of one chunk on FP64 MFMA (v_mfma_f64_16x16x4f64).  Batches of
// B points are staged in LDS as FP64; each wave owns (sub-cluster, 16 x 16 tile) pairs of the upper
// block triangle of the scatter and accumulates 4 points per MFMA: lane L supplies coordinate L % 16
// of point L / 16 of the sub-cluster's compacted list as both operands, and output element v of
// lane L is row 4 v + L / 16, column L % 16 of the tile.  The next batch's point indices, sub-labels
// and coordinates are loaded into registers while the current batch is multiplied, so the gather's
// latency overlaps the MFMA work.  Writes the partials of (chunk, 0) and (chunk, 1).
template <typename TX, int DP>
__global__ void __launch_bounds__(kThreads) k_stats_gauss_mfma(const TX* __restrict__ X, int D, int K, int CH,
                                                               const int32_t* __restrict__ perm,
                                                               const int32_t* __restrict__ label_start,
                                                               const int32_t* __restrict__ chunk_start,
                                                               const int32_t* __restrict__ subs,
                                                               const int32_t* __restrict__ subs_perm,
                                                               double* __restrict__ partials,
                                                               const int32_t* __restrict__ dirty, StatsGate sg) {
    typedef double f64x4 __attribute__((ext_vector_type(4)));
    constexpr int T = DP / 16;
    constexpr int NTILE = T * (T + 1) / 2;
    constexpr int TPW = (2 * NTILE + 3) / 4;          // (sub-cluster, tile) pairs per wave
    constexpr int B = DP > 64 ? 32 : 64;               // points per batch
    constexpr int PER = B * DP / kThreads;             // staged coordinates per thread
    constexpr size_t E = 1 + DP + (size_t)DP * DP;
    static_assert(B * DP % (4 * kThreads) == 0, "batch must tile the workgroup in groups of 4");
    constexpr int XS = DP + 1;                         // padded row: the 4 points of an MFMA step
    __shared__ __align__(16) double xs[(B + 1) * XS];  // start in different LDS banks; row B is zero
    __shared__ int idn[B];          // point indices of the batch being loaded
    __shared__ int sbb[B];          // sub-labels of the staged batch
    __shared__ int lst[2][B + 4];
    __shared__ int nls[2];
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end)) return;
    if (dirty && !dirty[k]) return;   // statistics of this cluster are kept from the last pass
    const int shift = stats_delta_shift(sg, k, label_start);   // kFullPass: every point
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int kp = lane >> 4, r = lane & 15;
    int us[TPW], ca[TPW], cb[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        int t = wv + 4 * u;
        us[u] = -1; ca[u] = 0; cb[u] = 0;
        if (t < 2 * NTILE) {
            us[u] = t / NTILE;
            t -= us[u] * NTILE;
            int row = 0;
            while (t >= T - row) { t -= T - row; ++row; }
            ca[u] = 16 * row + r; cb[u] = 16 * (row + t) + r;
        }
    }
    f64x4 acc[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) acc[u] = f64x4{0.0, 0.0, 0.0, 0.0};
    // first moments: wave w, lane (kp, r) sums coordinate 16 ti + r (ti = w, w + 4, ...) of its points
    constexpr int TS = (T + 3) / 4;
    double sums[2][TS];
#pragma unroll
    for (int sub = 0; sub < 2; ++sub)
#pragma unroll
        for (int u = 0; u < TS; ++u) sums[sub][u] = 0.0;
    int64_t cnt0 = 0, cnt1 = 0;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    for (int e = tid; e < XS; e += kThreads) xs[B * XS + e] = 0.0;   // the row that pads a group of 4
    // prefetch pipeline: indices two batches ahead, sub-labels and coordinates one batch ahead
    int id1 = (tid < B && beg + tid < end) ? perm[beg + tid] : -1;
    int id2 = (tid < B && beg + B + tid < end) ? perm[beg + B + tid] : -1;
    // sub-labels in partition order when the sub-label sweep left them (coalesced), else gathered
    auto sub_at = [&](int pos, int id) {
        if (id < 0) return -1;
        const int sv = subs_perm ? subs_perm[pos] : subs[id];
        return shift != kFullPass && sg.subs_prev[pos + shift] == sv ? -1 : sv;   // moved points only
    };
    int sb1 = sub_at(beg + tid, id1);
    if (tid < B) idn[tid] = id1;
    __syncthreads();
    // thread tid stages coordinates 4 e4 .. 4 e4 + 3 of the batch (e4 = tid + kThreads j): point
    // 4 e4 / DP, coordinates (4 e4) % DP onwards, loaded 4 at a time
    TX xv[PER / 4][4];
    auto load_batch = [&]() {
#pragma unroll
        for (int j = 0; j < PER / 4; ++j) {
            const int e = 4 * (tid + kThreads * j), q = e / DP, l = e - q * DP;
            const int id = idn[q];
            load_coords4(X + (int64_t)(id >= 0 ? id : 0) * D, l, D, id >= 0, xv[j]);
        }
    };
    load_batch();
    for (int b0 = beg; b0 < end; b0 += B) {
        __syncthreads();   // the previous batch's readers of xs / idn are done
#pragma unroll
        for (int j = 0; j < PER / 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int f = 4 * (tid + kThreads * j), q = f / DP, l = f - q * DP;
                xs[q * XS + l + e] = (double)xv[j][e];
            }
        if (tid < B) {
            sbb[tid] = sb1;
            idn[tid] = id2;   // indices of the next batch
        }
        __syncthreads();
        if (wv == 0) {   // compact by sub-label, index order kept
            const int sb = lane < B ? sbb[lane] : -1;
            const uint64_t m0 = __ballot(sb == 0), m1 = __ballot(sb == 1);
            if (sb == 0) lst[0][__popcll(m0 & lt)] = lane;
            if (sb == 1) lst[1][__popcll(m1 & lt)] = lane;
            const int c0 = __popcll(m0), c1 = __popcll(m1);
            if (lane < 4) { lst[0][c0 + lane] = 0; lst[1][c1 + lane] = 0; }   // in-range pads
            if (lane == 0) { nls[0] = c0; nls[1] = c1; }
        }
        // next batch: coordinates and sub-labels from the indices now in idn, indices after that
        load_batch();
        sb1 = sub_at(b0 + B + tid, id2);
        id2 = (tid < B && b0 + 2 * B + tid < end) ? perm[b0 + 2 * B + tid] : -1;
        __syncthreads();
        const int n0 = nls[0], n1 = nls[1];
        cnt0 += n0;
        cnt1 += n1;
        // lane (kp, r) holds coordinates r, 16 + r, ... of point kp of each group of 4: operand of
        // every tile; (sub-cluster, tile) pair p belongs to wave p % 4, accumulator p / 4
#pragma unroll
        for (int sub = 0; sub < 2; ++sub) {
            const int ns = sub ? n1 : n0;
            const int* L = lst[sub];
            for (int g = 0; g < ns; g += 4) {
                const int q = g + kp;
                const double* row = xs + (q < ns ? L[q] : B) * XS + r;
                double v[T];
#pragma unroll
                for (int ti = 0; ti < T; ++ti) v[ti] = row[16 * ti];
#pragma unroll
                for (int ti = 0; ti < T; ++ti)
                    if ((ti & 3) == wv) sums[sub][ti >> 2] += v[ti];
#pragma unroll
                for (int ti = 0; ti < T; ++ti)
#pragma unroll
                    for (int tj = ti; tj < T; ++tj) {
                        const int pr = sub * NTILE + ti * T - ti * (ti - 1) / 2 + (tj - ti);
                        if ((pr & 3) == wv)
                            acc[pr >> 2] = __builtin_amdgcn_mfma_f64_16x16x4f64(v[ti], v[tj], acc[pr >> 2], 0, 0, 0);
                    }
            }
        }
    }
    double* out0 = partials + (size_t)(2 * blockIdx.x) * E;
    if (tid == 0) { out0[0] = (double)cnt0; out0[E] = (double)cnt1; }
    // sums over the 4 points of each group (lanes r, r + 16, r + 32, r + 48), fixed order
#pragma unroll
    for (int sub = 0; sub < 2; ++sub)
#pragma unroll
        for (int u = 0; u < TS; ++u) {
            const int ti = 4 * u + wv;
            double v = sums[sub][u];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (kp == 0 && ti < T) out0[sub * E + 1 + 16 * ti + r] = v;
        }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        if (us[u] < 0) continue;
        double* S = out0 + us[u] * E + 1 + DP;
        const int ti = (ca[u] - r) / 16, tj = (cb[u] - r) / 16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int ra = 16 * ti + 4 * v + kp, cbb = 16 * tj + r;
            S[(size_t)ra * DP + cbb] = acc[u][v];
            S[(size_t)cbb * DP + ra] = acc[u][v];
        }
    }
}

// Statistics of both sub-clusters of one chunk on FP64 MFMA without LDS staging (DP <= 64).
// Coordinates are taken in a permuted order: with T = DP / 16, lane (kp, r) reads the T consecutive
// coordinates T r .. T r + T - 1 of point kp of a group of 4 in one vector load, and its e-th value
// is element r of "tile" e, the 16 coordinates T r' + e.  Tile pair (e1 <= e2) accumulates
// sum x[T a + e1] x[T b + e2] in one v_mfma_f64_16x16x4f64 per 4 points; output element v of lane
// (kp, r) is a = 4 v + kp, b = r, written back to scatter entries (T a + e1, T b + e2) and their
// mirror.  The sub-clusters are done one after the other (one set of accumulators): each wave
// owns 64-point batches b = wave, wave + 4, ... of the chunk, moves the batch's points of the
// sub-cluster to the front with one lane permute (ds_permute of a stable partition), then gathers
// 4 groups of 4 rows at a time straight from global memory into MFMA operands.  The four waves'
// accumulators are summed through LDS in fixed order (waves 3, 2, 1, then 0).  Each batch's indices
// and sub-labels are loaded one batch ahead.  In update mode (StatsGate) a chunk of a cluster that kept
// its points takes one pass over the points whose sub-label moved, the first operand signed (+ for a
// point now in sub-cluster 0, - in 1): slot 0 then holds the change of (k, 0), which (k, 1) loses.
template <typename TX, int DP>
__global__ void __launch_bounds__(kThreads, sizeof(TX) == 4 ? 3 : 2) k_stats_gauss_wave(const TX* __restrict__ X, int D, int K, int CH,
                                                               const int32_t* __restrict__ perm,
                                                               const int32_t* __restrict__ label_start,
                                                               const int32_t* __restrict__ chunk_start,
                                                               const int32_t* __restrict__ subs,
                                                               const int32_t* __restrict__ subs_perm,
                                                               double* __restrict__ partials,
                                                               const int32_t* __restrict__ dirty, StatsGate sg) {
    typedef double f64x4 __attribute__((ext_vector_type(4)));
    typedef TX txv __attribute__((ext_vector_type(DP / 16 == 4 ? 4 : 2)));
    constexpr int T = DP / 16;
    constexpr int NT = T * (T + 1) / 2;
    constexpr size_t E = 1 + DP + (size_t)DP * DP;
    constexpr int NV = 4 * NT + T;   // doubles per lane summed across the waves
    static_assert(DP % 16 == 0 && DP <= 64, "wave statistics kernel covers DP = 16 .. 64");
    __shared__ double red[NV * 64];
    __shared__ int64_t cnt_w[4];
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end)) return;
    if (dirty && !dirty[k]) return;   // statistics of this cluster are kept from the last pass
    const int shift = stats_delta_shift(sg, k, label_start);   // kFullPass: every point
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kp = lane >> 4, r = lane & 15;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const bool vec = (T == 2 || T == 4) && D % T == 0;   // rows aligned to T elements
    const int c0 = T * r;
    double* out = partials + (size_t)(2 * blockIdx.x) * E;

    // update mode (shift != kFullPass): one pass over the points whose sub-label moved, each signed (+ into
    // sub-cluster 0, - into 1), giving the change of (k, 0) in slot 0; full mode: one pass per sub-cluster
    const bool upd = shift != kFullPass && subs_perm;   // update mode reads the sub-labels in partition order
    for (int s = 0; s < (upd ? 1 : 2); ++s, out += E) {
        f64x4 acc[NT];
        double sm[T];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int e = 0; e < T; ++e) sm[e] = 0.0;
        int64_t cnt = 0;
        // a point's index, its sub-label and (update mode) its previous sub-label, loaded one batch ahead;
        // sub-labels in partition order when the sub-label sweep left them (a coalesced read that does not
        // wait for perm), else gathered by index
        auto fetch = [&](int p, int& idx, int& sn, int& sp) {
            const bool in = p < end;
            idx = in ? perm[p] : 0;
            sn = in ? (subs_perm ? subs_perm[p] : subs[idx]) : -1;
            sp = (in && upd) ? sg.subs_prev[p + shift] : -1;
        };
        int idx_n, sn_n, sp_n;
        fetch(beg + 64 * wv + lane, idx_n, sn_n, sp_n);
        for (int b0 = beg + 64 * wv; b0 < end; b0 += 4 * 64) {
            const int p = b0 + lane;
            const int idx = idx_n, sn = sn_n, sp = sp_n;
            fetch(p + 4 * 64, idx_n, sn_n, sp_n);
            const bool mine = p < end && (upd ? sn != sp : sn == s);
            const uint64_t m = __ballot(mine);
            const int ns = __popcll(m);
            if (upd) {   // signed count: into 0 minus into 1
                const int n1 = __popcll(__ballot(mine && sn == 1));
                cnt += ns - 2 * n1;
            } else {
                cnt += ns;
            }
            // stable partition: the selected points first (the rest are never used); bit 31 of the value
            // marks a point that moved into sub-cluster 1
            const int dst = mine ? __popcll(m & lt) : ns + __popcll(~m & lt);
            const int cval = __builtin_amdgcn_ds_permute(dst << 2, idx | ((upd && sn == 1) ? (int)0x80000000u : 0));
            for (int g = 0; g < ns; g += 16) {
                TX xv[4][T];
                double sgn[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {   // coordinates c0 .. c0 + T - 1 of 4 rows (zero past D)
                    const int q = g + 4 * u + kp;
                    const int val = __shfl(cval, q < ns ? q : (g + 4 * u < ns ? ns - 1 : 0));
                    sgn[u] = val < 0 ? -1.0 : 1.0;
                    const TX* row = X + (int64_t)(val & 0x7fffffff) * D + c0;
                    if (g + 4 * u >= ns) {   // a group past the selected points: not loaded, not used
#pragma unroll
                        for (int e = 0; e < T; ++e) xv[u][e] = TX(0);
                    } else if (vec && c0 + T <= D) {
                        if constexpr (T == 2 || T == 4) {
                            const txv w = *reinterpret_cast<const txv*>(row);
#pragma unroll
                            for (int e = 0; e < T; ++e) xv[u][e] = w[e];
                        }
                    } else {
#pragma unroll
                        for (int e = 0; e < T; ++e) xv[u][e] = c0 + e < D ? row[e] : TX(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (g + 4 * u >= ns) break;
                    const bool ok = g + 4 * u + kp < ns;
                    double v[T];
#pragma unroll
                    for (int e = 0; e < T; ++e) {
                        v[e] = ok ? (double)xv[u][e] : 0.0;
                        sm[e] += sgn[u] * v[e];
                    }
                    int t = 0;
#pragma unroll
                    for (int e1 = 0; e1 < T; ++e1)
#pragma unroll
                        for (int e2 = e1; e2 < T; ++e2, ++t)
                            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sgn[u] * v[e1], v[e2], acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();   // the previous sub-cluster's readers of red are done
        if (lane == 0) cnt_w[wv] = cnt;
        for (int w = 3; w >= 1; --w) {
            if (wv == w) {
                int i = 0;
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int c = 0; c < 4; ++c, ++i)
                        red[i * 64 + lane] = (w == 3 ? 0.0 : red[i * 64 + lane]) + acc[t][c];
#pragma unroll
                for (int e = 0; e < T; ++e, ++i) red[i * 64 + lane] = (w == 3 ? 0.0 : red[i * 64 + lane]) + sm[e];
            }
            __syncthreads();
        }
        if (wv == 0) {
            if (lane == 0) out[0] = (double)(cnt_w[3] + cnt_w[2] + cnt_w[1] + cnt_w[0]);
            double* S = out + 1 + DP;
            int i = 0, t = 0;
#pragma unroll
            for (int e1 = 0; e1 < T; ++e1)
#pragma unroll
                for (int e2 = e1; e2 < T; ++e2, ++t) {
#pragma unroll
                    for (int c = 0; c < 4; ++c, ++i) {
                        const double val = red[i * 64 + lane] + acc[t][c];
                        const int ra = T * (4 * c + kp) + e1, cb = T * r + e2;
                        S[(size_t)ra * DP + cb] = val;
                        S[(size_t)cb * DP + ra] = val;
                    }
                }
            // first moments: the 4 points of each group (lanes r, r + 16, r + 32, r + 48), fixed order
#pragma unroll
            for (int e = 0; e < T; ++e, ++i) {
                double v = red[i * 64 + lane] + sm[e];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                if (kp == 0) out[1 + c0 + e] = v;
            }
        }
    }
}

// Exact sub-label draw of listed points (the undecided points of the mixed sub-label screen), with
// the arithmetic of k_sub_gauss: each point's two sub-cluster slots (of its label) are read from
// global memory; workgroups stride over the list, whose length is read from device memory.
template <typename T, int DP>
__global__ void __launch_bounds__(kThreads) k_sub_gauss_list(const T* __restrict__ X, int D, int K,
                                                             const T* __restrict__ SP,
                                                             const int32_t* __restrict__ labels,
                                                             const int32_t* __restrict__ list,
                                                             const int* __restrict__ list_count, uint32_t key0,
                                                             uint32_t key1, uint32_t iter,
                                                             int32_t* __restrict__ subs,
                                                             const int32_t* __restrict__ perm,
                                                             int32_t* __restrict__ subs_perm) {
    constexpr bool kPair = paired_layout<T, DP>();
    constexpr int kPts = kPair ? kThreads / 2 : kThreads;
    constexpr int kCols = kPair ? DP / 2 : DP;
    constexpr int stride = gauss_slot_stride(DP);
    const int par = kPair ? (threadIdx.x & 1) : 0;
    const int local = kPair ? (threadIdx.x >> 1) : threadIdx.x;
    const int n = *list_count;
    for (int b0 = blockIdx.x * kPts; b0 < n; b0 += gridDim.x * kPts) {   // uniform trip count per pair
        const int pos = b0 + local;
        const bool valid = pos < n;
        // with perm, list entries are positions in the partition (the sub-label is also written there)
        const int ent = valid ? list[pos] : 0;
        const int idx = perm ? perm[ent] : ent;
        const int k0 = valid ? labels[idx] : 0;
        const bool ok = valid && (unsigned)k0 < (unsigned)K;   // labels outside [0, K) are left alone
        const int k = ok ? k0 : 0;
        T x[kCols];
#pragma unroll
        for (int c = 0; c < kCols; ++c) {
            const int l = kPair ? pair_col(c, 0) + (par << 2) : c;
            x[c] = (ok && l < D) ? X[(int64_t)idx * D + l] : T(0);
        }
        T sc[2];
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            const T* pp = SP + (size_t)(2 * k + h) * stride;
            T q;
            if constexpr (kPair) q = whitened_sq_pair<T, DP>(pp, x, par);
            else q = whitened_sq_regs<T, DP>(pp, x);
            sc[h] = pp[gauss_c_offset(DP)] - T(0.5) * q;
        }
        if (ok && par == 0) {
            const uint32_t w = philox4x32_10((uint32_t)idx, 0, iter, TAG_SUB, key0, key1).v[0];
            const int v = (logit_u32<T>(w) < sc[0] - sc[1]) ? 0 : 1;
            subs[idx] = v;
            if (subs_perm) subs_perm[ent] = v;
        }
    }
}

// FP32 sub-label draw of the points the fused label + sub-label sweep (k_label_gauss_mxp) left, listed
// per cluster inside the cluster's segment [label_start[a], label_start[a + 1]) of the swept partition:
// seg[a] points whose label stayed a from the segment's start up, seg[K + a] others (label moved, or
// decided by the FP32 label recheck) from its end down.  The 2K lists are cut into pieces of kThreads
// points (one per thread) and each workgroup takes a contiguous range of the pieces, so a piece of a
// first list shares a's two sub-cluster slots, staged in LDS once per run of pieces of one cluster;
// a lane draws its point with the arithmetic of k_sub_gauss (whitened_sq_lane), for the second lists
// from its own label's slots in global memory.  Sub-labels go to subs (by index) and subs_perm (by
// position); labels outside [0, K) are left alone.  *total: how many points were listed.
constexpr int kSegMaxLists = 4096;   // 2 K
template <typename T, int DP>
__global__ void __launch_bounds__(kThreads) k_sub_gauss_seg(const T* __restrict__ X, int D, int K,
                                                            const int32_t* __restrict__ perm,
                                                            const int32_t* __restrict__ label_start,
                                                            const T* __restrict__ SP,
                                                            const int32_t* __restrict__ labels,
                                                            const int32_t* __restrict__ list,
                                                            const int32_t* __restrict__ seg, uint32_t key0,
                                                            uint32_t key1, uint32_t iter, int32_t* __restrict__ subs,
                                                            int32_t* __restrict__ subs_perm, int* __restrict__ total) {
    constexpr int stride = gauss_slot_stride(DP);
    constexpr bool kLds = sizeof(T) == 4 && DP <= 64;   // the fused sweep's widths
    __shared__ __align__(16) T ws[kLds ? 2 * stride : 4];
    __shared__ int P[kSegMaxLists + 1];   // exclusive prefix of the lists' piece counts
    __shared__ int wsum[2][kThreads / 64];
    const int nl = 2 * K, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // pieces per list, scanned over the workgroup: thread t owns lists [t m, t m + m)
    const int m = (nl + kThreads - 1) / kThreads;
    int pcs = 0, ents = 0;
    for (int u = 0; u < m; ++u) {
        const int t = tid * m + u;
        const int c = t < nl ? seg[t] : 0;
        pcs += (c + kThreads - 1) / kThreads;
        ents += c;
    }
    int inc = pcs, tot = ents;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o);
        if (lane >= o) inc += v;
        tot += __shfl_xor(tot, o);
    }
    if (lane == 63) wsum[0][wv] = inc;
    if (lane == 0) wsum[1][wv] = tot;
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wv; ++w) off += wsum[0][w];
    int run = off + inc - pcs;
    for (int u = 0; u < m; ++u) {
        const int t = tid * m + u;
        if (t < nl) {
            P[t] = run;
            const int c = seg[t];
            run += (c + kThreads - 1) / kThreads;
        }
    }
    if (tid == kThreads - 1) P[nl] = run;
    __syncthreads();
    if (blockIdx.x == 0 && tid == 0) {
        int e = 0;
        for (int w = 0; w < kThreads / 64; ++w) e += wsum[1][w];
        *total = e;
    }
    const int npieces = P[nl];
    const int q0 = (int)((int64_t)npieces * blockIdx.x / gridDim.x);
    const int q1 = (int)((int64_t)npieces * (blockIdx.x + 1) / gridDim.x);
    int staged = -1;
    for (int q = q0; q < q1; ++q) {   // uniform over the workgroup
        int lo = 0, hi = nl;          // the list holding piece q: the last t with P[t] <= q
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (P[mid] <= q) lo = mid;
            else hi = mid;
        }
        const bool stays = lo < K;
        const int a = stays ? lo : lo - K;
        const int i = (q - P[lo]) * kThreads + tid;
        const bool in = i < seg[lo];
        if constexpr (kLds) {
            if (stays && staged != a) {
                __syncthreads();   // the last piece's readers are done
                const float4* src = reinterpret_cast<const float4*>(SP + (size_t)(2 * a) * stride);
                float4* dst = reinterpret_cast<float4*>(ws);
                static_assert(stride % 4 == 0, "16-byte slot rows");
                for (int e = tid; e < 2 * stride / 4; e += kThreads) dst[e] = src[e];
                __syncthreads();
                staged = a;
            }
        }
        const int ent = in ? (stays ? list[label_start[a] + i] : list[label_start[a + 1] - 1 - i]) : 0;
        const int idx = in ? perm[ent] : 0;
        int k = a;
        bool ok = in;
        if (!stays) {
            const int k0 = in ? labels[idx] : 0;
            ok = in && (unsigned)k0 < (unsigned)K;   // labels outside [0, K) are left alone
            k = ok ? k0 : 0;
        }
        T sc[2];
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            // the coordinates are read for each sub-cluster (the second time from cache): kept across
            // both, they would hold 64 more registers
            T x[DP];
            if constexpr (sizeof(T) == 4) {
#pragma unroll
                for (int l = 0; l < DP; l += 4) {
                    const bool c = ok && l < D;   // D % 4 == 0 (the fused sweep's condition): 16-byte rows
                    const float4 v = *reinterpret_cast<const float4*>(X + (int64_t)(c ? idx : 0) * D + (c ? l : 0));
                    x[l] = c ? v.x : T(0);
                    x[l + 1] = c ? v.y : T(0);
                    x[l + 2] = c ? v.z : T(0);
                    x[l + 3] = c ? v.w : T(0);
                }
            } else {
#pragma unroll
                for (int l = 0; l < DP; ++l) x[l] = (ok && l < D) ? X[(int64_t)idx * D + l] : T(0);
            }
            if (kLds && stays) {
                const T* pp = ws + h * stride;
                sc[h] = pp[gauss_c_offset(DP)] - T(0.5) * whitened_sq_lane<T, DP>(pp, x);
            } else {
                const T* pp = SP + (size_t)(2 * k + h) * stride;
                sc[h] = pp[gauss_c_offset(DP)] - T(0.5) * whitened_sq_lane<T, DP>(pp, x);
            }
        }
        if (ok) {
            const uint32_t w = philox4x32_10((uint32_t)idx, 0, iter, TAG_SUB, key0, key1).v[0];
            const int v = (logit_u32<T>(w) < sc[0] - sc[1]) ? 0 : 1;
            subs[idx] = v;
            subs_perm[ent] = v;
        }
    }
}

// Noise-free scores of points [first, first + n) with the sweeps' arithmetic (the FP32 / FP64
// expression c - q / 2 of k_label_gauss and k_sub_gauss): which = 0 writes the n x K matrix of
// cluster scores (log weight + log N(x | mu_k, Sigma_k)), which = 1 the n x 2 scores of the two
// sub-clusters of each point's label (log sub-weight + log-likelihood).  Slots are read from global
// memory; workgroups stride over the points with a uniform trip count per lane pair.
template <typename T, int DP>
__global__ void __launch_bounds__(kThreads) k_score_gauss(const T* __restrict__ X, int D, int K,
                                                          const T* __restrict__ P, const T* __restrict__ SP,
                                                          const int32_t* __restrict__ labels, int64_t first,
                                                          int64_t n, int which, T* __restrict__ out) {
    constexpr bool kPair = paired_layout<T, DP>();
    constexpr int kPts = kPair ? kThreads / 2 : kThreads;
    constexpr int kCols = kPair ? DP / 2 : DP;
    constexpr int stride = gauss_slot_stride(DP);
    const int par = kPair ? (threadIdx.x & 1) : 0;
    const int local = kPair ? (threadIdx.x >> 1) : threadIdx.x;
    for (int64_t b0 = (int64_t)blockIdx.x * kPts; b0 < n; b0 += (int64_t)gridDim.x * kPts) {
        const int64_t pos = b0 + local;
        const bool valid = pos < n;
        const int64_t idx = valid ? first + pos : 0;
        T x[kCols];
#pragma unroll
        for (int c = 0; c < kCols; ++c) {
            const int l = kPair ? pair_col(c, 0) + (par << 2) : c;
            x[c] = (valid && l < D) ? X[idx * D + l] : T(0);
        }
        const int kl = (which == 1 && valid) ? labels[idx] : 0;
        const int cnt = which == 1 ? 2 : K;
#pragma unroll 1
        for (int k = 0; k < cnt; ++k) {
            const T* pp = which == 1 ? SP + (size_t)(2 * kl + k) * stride : P + (size_t)k * stride;
            T q;
            if constexpr (kPair) q = whitened_sq_pair<T, DP>(pp, x, par);
            else q = whitened_sq_regs<T, DP>(pp, x);
            const T s = pp[gauss_c_offset(DP)] - T(0.5) * q;
            if (valid && par == 0) out[pos * cnt + k] = s;
        }
    }
}

// Partial (count, sum, scatter) of the points of sub-cluster s in one chunk, FP64.  Threads own
// 4x4 blocks of the upper block triangle; points staged in LDS in batches, compacted by ballot.
template <typename TX, int DP>
__global__ void __launch_bounds__(kThreads) k_stats_gauss(const TX* __restrict__ X, int D, int K, int CH,
                                                          const int32_t* __restrict__ perm,
                                                          const int32_t* __restrict__ label_start,
                                                          const int32_t* __restrict__ chunk_start,
                                                          const int32_t* __restrict__ subs,
                                                          const int32_t* __restrict__ subs_perm,
                                                          double* __restrict__ partials,
                                                          const int32_t* __restrict__ dirty, StatsGate sg) {
    constexpr int NB = DP / 4;
    constexpr int NTRI = NB * (NB + 1) / 2;
    constexpr int R = (NTRI + kThreads - 1) / kThreads;
    constexpr int BATCH = DP > 64 ? 32 : 64;
    constexpr size_t E = 1 + DP + (size_t)DP * DP;
    __shared__ __align__(16) double xs[BATCH * DP];
    __shared__ int idl[64];
    __shared__ int nsel;
    const int c = blockIdx.x >> 1, s = blockIdx.x & 1;
    int k, beg, end;
    if (!chunk_of(c, K, chunk_start, label_start, CH, k, beg, end)) return;
    if (dirty && !dirty[k]) return;   // statistics of this cluster are kept from the last pass
    const int shift = stats_delta_shift(sg, k, label_start);   // kFullPass: every point
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int bi[R], bj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int b = threadIdx.x + r * kThreads;
        bi[r] = -1; bj[r] = -1;
        if (b < NTRI) {   // row-major upper block triangle
            int row = 0;
            while (b >= NB - row) { b -= NB - row; ++row; }
            bi[r] = row; bj[r] = row + b;
        }
    }
    double acc[R][16];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[r][e] = 0.0;
    double sx = 0.0;
    int64_t cnt = 0;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int b0 = beg; b0 < end; b0 += BATCH) {
        if (wv == 0) {
            const int p = b0 + lane;
            int idx = -1;
            bool f = false;
            if (lane < BATCH && p < end) {
                idx = perm[p];
                f = (subs_perm ? subs_perm[p] : subs[idx]) == s && !(shift != kFullPass && sg.subs_prev[p + shift] == s);
            }
            const uint64_t m = __ballot(f);
            if (f) idl[__popcll(m & lt)] = idx;
            if (lane == 0) nsel = __popcll(m);
        }
        __syncthreads();
        const int n = nsel;
        for (int e = threadIdx.x; e < n * DP; e += kThreads) {
            const int pp = e / DP, l = e - pp * DP;
            xs[pp * DP + l] = l < D ? (double)X[(int64_t)idl[pp] * D + l] : 0.0;
        }
        __syncthreads();
        for (int pp = 0; pp < n; ++pp) {
            const double* xp = xs + pp * DP;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (bi[r] < 0) continue;
                double a[4], b[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) { a[u] = xp[4 * bi[r] + u]; b[u] = xp[4 * bj[r] + u]; }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[r][u * 4 + v] = fma(a[u], b[v], acc[r][u * 4 + v]);
            }
            if (threadIdx.x < DP) sx += xp[threadIdx.x];
        }
        cnt += n;
        __syncthreads();
    }
    double* out = partials + (size_t)blockIdx.x * E;
    if (threadIdx.x == 0) out[0] = (double)cnt;
    if (threadIdx.x < DP) out[1 + threadIdx.x] = sx;
    double* S = out + 1 + DP;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (bi[r] < 0) continue;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int a = 4 * bi[r] + u, b = 4 * bj[r] + v;
                S[(size_t)a * DP + b] = acc[r][u * 4 + v];
                S[(size_t)b * DP + a] = acc[r][u * 4 + v];
            }
    }
}


template <typename T, int DP>
constexpr int label_points_per_block() { return paired_layout<T, DP>() ? kThreads / 2 : kThreads; }

template <typename T, int DP>
void launch_label_gauss(hipStream_t st, int grid, const T* X, int64_t N, int D, int K, const T* P, uint32_t key0,
                        uint32_t key1, uint32_t iter, int fin, int32_t* labels) {
    hipLaunchKernelGGL((k_label_gauss<T, DP>), dim3(grid), dim3(kThreads), 0, st, X, N, D, K, P, key0, key1, iter,
                       fin, labels);
}

template <typename T, int DP>
void launch_label_gauss_list(hipStream_t st, int grid, const T* X, int64_t N, int D, int K, const T* P, uint32_t key0,
                             uint32_t key1, uint32_t iter, int fin, const int32_t* list, const int* count,
                             int32_t* labels) {
    hipLaunchKernelGGL((k_label_gauss_list<T, DP>), dim3(grid), dim3(kThreads), 0, st, X, N, D, K, P, key0, key1,
                       iter, fin, list, count, labels);
}

template <typename T, int DP>
void launch_label_gauss_listp(hipStream_t st, int grid, const T* X, int64_t N, int D, int K, const T* P,
                              const float* pm, const float* psc, uint32_t key0, uint32_t key1, uint32_t iter, int fin,
                              const int32_t* list, const int* count, int32_t* labels, int* changed) {
    hipLaunchKernelGGL((k_label_gauss_listp<T, DP>), dim3(grid), dim3(kThreads), 0, st, X, N, D, K, P, pm, psc, key0,
                       key1, iter, fin, list, count, labels, changed);
}

template <typename T, int DP>
void launch_sub_gauss_list(hipStream_t st, int grid, const T* X, int D, int K, const T* SP, const int32_t* labels,
                           const int32_t* list, const int* count, uint32_t key0, uint32_t key1, uint32_t iter,
                           int32_t* subs, const int32_t* perm, int32_t* subs_perm) {
    hipLaunchKernelGGL((k_sub_gauss_list<T, DP>), dim3(grid), dim3(kThreads), 0, st, X, D, K, SP, labels, list,
                       count, key0, key1, iter, subs, perm, subs_perm);
}

template <typename T, int DP>
void launch_sub_gauss_seg(hipStream_t st, int grid, const T* X, int D, int K, const int32_t* perm,
                          const int32_t* label_start, const T* SP, const int32_t* labels, const int32_t* list,
                          const int32_t* seg, uint32_t key0, uint32_t key1, uint32_t iter, int32_t* subs,
                          int32_t* subs_perm, int* total) {
    hipLaunchKernelGGL((k_sub_gauss_seg<T, DP>), dim3(grid), dim3(kThreads), 0, st, X, D, K, perm, label_start, SP,
                       labels, list, seg, key0, key1, iter, subs, subs_perm, total);
}

template <typename T, int DP>
void launch_sub_gauss(hipStream_t st, int grid, const T* X, int D, int K, int CH, const int32_t* perm,
                      const int32_t* label_start, const int32_t* chunk_start, const T* SP, uint32_t key0,
                      uint32_t key1, uint32_t iter, int32_t* subs) {
    hipLaunchKernelGGL((k_sub_gauss<T, DP>), dim3(grid), dim3(kThreads), 0, st, X, D, K, CH, perm, label_start,
                       chunk_start, SP, key0, key1, iter, subs);
}

template <typename T, int DP>
void launch_score_gauss(hipStream_t st, int grid, const T* X, int D, int K, const T* P, const T* SP,
                        const int32_t* labels, int64_t first, int64_t n, int which, T* out) {
    hipLaunchKernelGGL((k_score_gauss<T, DP>), dim3(grid), dim3(kThreads), 0, st, X, D, K, P, SP, labels, first, n,
                       which, out);
}

// nchunks: upper bound of the chunk count (one MFMA block per chunk, one vector block per
// (chunk, sub-cluster)).
template <typename TX, int DP>
void launch_stats_gauss(hipStream_t st, int nchunks, const TX* X, int D, int K, int CH, const int32_t* perm,
                        const int32_t* label_start, const int32_t* chunk_start, const int32_t* subs,
                        const int32_t* subs_perm, double* partials, const int32_t* dirty, StatsGate sg) {
    if constexpr (DP % 16 == 0 && DP <= 64)
        hipLaunchKernelGGL((k_stats_gauss_wave<TX, DP>), dim3(nchunks), dim3(kThreads), 0, st, X, D, K, CH, perm,
                           label_start, chunk_start, subs, subs_perm, partials, dirty, sg);
    else if constexpr (DP % 16 == 0)
        hipLaunchKernelGGL((k_stats_gauss_mfma<TX, DP>), dim3(nchunks), dim3(kThreads), 0, st, X, D, K, CH, perm,
                           label_start, chunk_start, subs, subs_perm, partials, dirty, sg);
    else
        hipLaunchKernelGGL((k_stats_gauss<TX, DP>), dim3(2 * nchunks), dim3(kThreads), 0, st, X, D, K, CH, perm,
                           label_start, chunk_start, subs, subs_perm, partials, dirty, sg);
}

#define DPMM_GAUSS_INSTANTIATE(T, DP)                                                                         \
    template void launch_label_gauss<T, DP>(hipStream_t, int, const T*, int64_t, int, int, const T*, uint32_t, \
                                            uint32_t, uint32_t, int, int32_t*);                                \
    template void launch_label_gauss_list<T, DP>(hipStream_t, int, const T*, int64_t, int, int, const T*,      \
                                                 uint32_t, uint32_t, uint32_t, int, const int32_t*, const int*, \
                                                 int32_t*);                                                    \
    template void launch_sub_gauss<T, DP>(hipStream_t, int, const T*, int, int, int, const int32_t*,           \
                                          const int32_t*, const int32_t*, const T*, uint32_t, uint32_t,        \
                                          uint32_t, int32_t*);                                                 \
    template void launch_stats_gauss<T, DP>(hipStream_t, int, const T*, int, int, int, const int32_t*,         \
                                            const int32_t*, const int32_t*, const int32_t*, const int32_t*,    \
                                            double*, const int32_t*, StatsGate);                               \
    template void launch_sub_gauss_list<T, DP>(hipStream_t, int, const T*, int, int, const T*, const int32_t*, \
                                               const int32_t*, const int*, uint32_t, uint32_t, uint32_t,       \
                                               int32_t*, const int32_t*, int32_t*);                            \
    template void launch_score_gauss<T, DP>(hipStream_t, int, const T*, int, int, const T*, const T*,          \
                                            const int32_t*, int64_t, int64_t, int, T*);                        \
    template void launch_sub_gauss_seg<T, DP>(hipStream_t, int, const T*, int, int, const int32_t*,            \
                                              const int32_t*, const T*, const int32_t*, const int32_t*,       \
                                              const int32_t*, uint32_t, uint32_t, uint32_t, int32_t*,         \
                                              int32_t*, int*);                                                 \
    template void launch_label_gauss_listp<T, DP>(hipStream_t, int, const T*, int64_t, int, int, const T*,     \
                                                  const float*, const float*, uint32_t, uint32_t, uint32_t, int, \
                                                  const int32_t*, const int*, int32_t*, int*);

#define DPMM_GAUSS_EXTERN(T, DP)                                                                              \
    extern template void launch_label_gauss<T, DP>(hipStream_t, int, const T*, int64_t, int, int, const T*,    \
                                                   uint32_t, uint32_t, uint32_t, int, int32_t*);               \
    extern template void launch_label_gauss_list<T, DP>(hipStream_t, int, const T*, int64_t, int, int,         \
                                                        const T*, uint32_t, uint32_t, uint32_t, int,           \
                                                        const int32_t*, const int*, int32_t*);                 \
    extern template void launch_sub_gauss<T, DP>(hipStream_t, int, const T*, int, int, int, const int32_t*,    \
                                                 const int32_t*, const int32_t*, const T*, uint32_t, uint32_t, \
                                                 uint32_t, int32_t*);                                          \
    extern template void launch_stats_gauss<T, DP>(hipStream_t, int, const T*, int, int, int, const int32_t*,  \
                                                   const int32_t*, const int32_t*, const int32_t*,             \
                                                   const int32_t*, double*, const int32_t*, StatsGate);        \
    extern template void launch_sub_gauss_list<T, DP>(hipStream_t, int, const T*, int, int, const T*,          \
                                                      const int32_t*, const int32_t*, const int*, uint32_t,    \
                                                      uint32_t, uint32_t, int32_t*, const int32_t*, int32_t*); \
    extern template void launch_score_gauss<T, DP>(hipStream_t, int, const T*, int, int, const T*, const T*,   \
                                                   const int32_t*, int64_t, int64_t, int, T*);                 \
    extern template void launch_sub_gauss_seg<T, DP>(hipStream_t, int, const T*, int, int, const int32_t*,    \
                                                     const int32_t*, const T*, const int32_t*, const int32_t*, \
                                                     const int32_t*, uint32_t, uint32_t, uint32_t, int32_t*,   \
                                                     int32_t*, int*);                                          \
    extern template void launch_label_gauss_listp<T, DP>(hipStream_t, int, const T*, int64_t, int, int,        \
                                                         const T*, const float*, const float*, uint32_t,       \
                                                         uint32_t, uint32_t, int, const int32_t*, const int*,  \
                                                         int32_t*, int*);

}  // namespace dpmm
