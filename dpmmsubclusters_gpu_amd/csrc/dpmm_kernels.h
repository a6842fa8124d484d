// Device building blocks shared by the sweep, partition, statistics and relabel kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpmm_rng.h"

namespace dpmm {

constexpr int kThreads = 256;

// Packed upper-triangular whitening factor W = R^T (R the lower Cholesky factor of the
// precision matrix), row j stored from column (j & ~3) to DP-1 so every row starts on a
// 16-byte boundary: entries below the diagonal inside that first group are zero.
// Offset of row j: sum over rows r < j of DP - (r & ~3); the 4-row groups g < j / 4 drop 16 g
// columns in all and the j % 4 rows of the last group 4 (j / 4) each.
__host__ __device__ constexpr int tri_row_offset(int DP, int j) {
    const int q = j >> 2, rem = j & 3;
    return j * DP - 8 * q * (q - 1) - 4 * rem * q;
}
__host__ __device__ constexpr int tri_size(int DP) { return tri_row_offset(DP, DP); }
// Gaussian parameter slot: [W packed | mu (DP) | c | pad] in elements of the compute type.
__host__ __device__ constexpr int gauss_slot_stride(int DP) { return ((tri_size(DP) + DP + 1) + 3) & ~3; }
__host__ __device__ constexpr int gauss_mu_offset(int DP) { return tri_size(DP); }
__host__ __device__ constexpr int gauss_c_offset(int DP) { return tri_size(DP) + DP; }

// -log(-log u), u = (w + 0.5) 2^-32, with -log u taken through log1p on the upper half so
// the exponential variate keeps full relative precision near u = 1.
__device__ __forceinline__ float gumbel_f32(uint32_t w) {
    float e;
    if (w < 0x80000000u) e = -logf(((float)w + 0.5f) * 0x1p-32f);
    else e = -log1pf(-(((float)(0xFFFFFFFFu - w)) + 0.5f) * 0x1p-32f);
    return -logf(e);
}
__device__ __forceinline__ double gumbel_f64(uint32_t w) {
    double e;
    if (w < 0x80000000u) e = -log(((double)w + 0.5) * 0x1p-32);
    else e = -log1p(-(((double)(0xFFFFFFFFu - w)) + 0.5) * 0x1p-32);
    return -log(e);
}
template <typename T> __device__ __forceinline__ T gumbel(uint32_t w);
template <> __device__ __forceinline__ float gumbel<float>(uint32_t w) { return gumbel_f32(w); }
template <> __device__ __forceinline__ double gumbel<double>(uint32_t w) { return gumbel_f64(w); }

// logit(u) = log(u) - log(1 - u) for u = (w + 0.5) 2^-32.
template <typename T> __device__ __forceinline__ T logit_u32(uint32_t w);
template <> __device__ __forceinline__ float logit_u32<float>(uint32_t w) {
    return logf((float)w + 0.5f) - logf((float)(0xFFFFFFFFu - w) + 0.5f);
}
template <> __device__ __forceinline__ double logit_u32<double>(uint32_t w) {
    return log((double)w + 0.5) - log((double)(0xFFFFFFFFu - w) + 0.5);
}

// Shift the next word of a Philox block into v[0] (static indices keep the block in registers).
__device__ __forceinline__ void rotate_words(Ctr4& c) {
    c.v[0] = c.v[1];
    c.v[1] = c.v[2];
    c.v[2] = c.v[3];
}

__device__ __forceinline__ uint32_t sel4(const Ctr4& c, int u) {
    const uint32_t a = (u & 1) ? c.v[1] : c.v[0];
    const uint32_t b = (u & 1) ? c.v[3] : c.v[2];
    return (u & 2) ? b : a;
}

template <typename T> __device__ __forceinline__ T neg_inf();
template <> __device__ __forceinline__ float neg_inf<float>() { return -__builtin_huge_valf(); }
template <> __device__ __forceinline__ double neg_inf<double>() { return -__builtin_huge_val(); }

// Mahalanobis form q = || W (x - mu) ||^2 with W packed upper triangular in LDS.
// Small points keep the centred residual d = x - mu in registers next to x.
template <typename T, int DP>
__device__ __forceinline__ T whitened_sq_regs(const T* __restrict__ ps, const T (&x)[DP]) {
    T d[DP];
#pragma unroll
    for (int l = 0; l < DP; ++l) d[l] = x[l] - ps[gauss_mu_offset(DP) + l];
    T q = 0;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
        const int l0 = j & ~3;
        const T* row = ps + tri_row_offset(DP, j) - l0;
        T y = 0;
#pragma unroll
        for (int l = l0; l < DP; ++l) y = fma(row[l], d[l], y);
        q = fma(y, y, q);
    }
    return q;
}

// Wide points (x and the residual would not both fit in registers) are split across a lane
// pair: lane parity `par` holds the coordinates l = 8 b + 4 par + u (u < 4) of every 8-block
// (local index c = 4 b + u), forms its half of each row sum, and the pair adds the halves with a
// DPP swap, so both lanes end with the same q.
__host__ __device__ constexpr int pair_col(int c, int par) { return ((c >> 2) << 3) + (par << 2) + (c & 3); }
// Points whose x (DP elements of T) exceeds 256 bytes take the lane-pair layout.
template <typename T, int DP> __host__ __device__ constexpr bool paired_layout() { return DP * sizeof(T) > 256; }

__device__ __forceinline__ float pair_swap(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));   // quad_perm 1,0,3,2
}
__device__ __forceinline__ double pair_swap(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0xB1, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <typename T, int DP>
__device__ __forceinline__ T whitened_sq_pair(const T* __restrict__ ps, const T (&xh)[DP / 2], int par) {
    constexpr int HC = DP / 2;
    const T* mu = ps + gauss_mu_offset(DP);
    T d[HC];
#pragma unroll
    for (int c = 0; c < HC; ++c) d[c] = xh[c] - mu[pair_col(c, 0) + (par << 2)];
    const T keep = par ? T(1) : T(0);   // the 4-block just below an odd row block belongs to odd lanes
    T q = 0;
#pragma unroll
    for (int j = 0; j < DP; ++j) {
        const int jb = j >> 2;
        const T* row = ps + tri_row_offset(DP, j) - (j & ~3) + (par << 2);
        T y = 0;
#pragma unroll
        for (int b = 0; b < HC / 4; ++b) {
            if (2 * b + 1 < jb) continue;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                T dv = d[4 * b + u];
                if (2 * b + 1 == jb) dv *= keep;
                y = fma(row[(b << 3) + u], dv, y);
            }
        }
        y += pair_swap(y);
        q = fma(y, y, q);
    }
    return q;
}

// whitened_sq_regs's arithmetic for a slot `ps` that is the same across the wave (its rows become scalar
// loads, SGPR operands of the FMAs): one row per trip of a loop that is not unrolled, so that only that
// row's loads are in flight; row j's terms run over columns (j & ~3) .. DP - 1 in order.  The rows of a
// 4-block share their first column, which keeps the column loop's bounds constant.
template <typename T, int DP>
__device__ __forceinline__ T whitened_sq_uniform(const T* __restrict__ ps, const T (&x)[DP]) {
    T d[DP];
#pragma unroll
    for (int l = 0; l < DP; ++l) d[l] = x[l] - ps[gauss_mu_offset(DP) + l];
    T q = 0;
#pragma unroll
    for (int l0 = 0; l0 < DP; l0 += 4) {   // the four rows j = l0 .. l0 + 3 share their first column
#pragma unroll 1
        for (int j = l0; j < l0 + 4; ++j) {
            const T* row = ps + tri_row_offset(DP, j) - l0;
            T y = 0;
#pragma unroll
            for (int l = l0; l < DP; ++l) y = fma(row[l], d[l], y);
            q = fma(y, y, q);
        }
    }
    return q;
}

// One lane's q with the arithmetic of the FP32 sweep's layout for DP: whitened_sq_regs's, or for a paired
// layout both halves of whitened_sq_pair formed in this lane, each in the order its lane of the pair
// uses, and added as the pair adds them (the same value in either lane).  x: the point's DP
// coordinates, zero past D.
template <typename T, int DP>
__device__ __forceinline__ T whitened_sq_lane(const T* __restrict__ ps, const T (&x)[DP]) {
    if constexpr (!paired_layout<T, DP>()) {
        // whitened_sq_regs's arithmetic, one row's loads at a time (hoisted over the whole triangle
        // they would take every register)
        T d[DP];
#pragma unroll
        for (int l = 0; l < DP; ++l) d[l] = x[l] - ps[gauss_mu_offset(DP) + l];
        T q = 0;
#pragma unroll
        for (int j = 0; j < DP; ++j) {
            const int l0 = j & ~3;
            const T* row = ps + tri_row_offset(DP, j) - l0;
            T y = 0;
#pragma unroll
            for (int l = l0; l < DP; ++l) y = fma(row[l], d[l], y);
            q = fma(y, y, q);
            asm volatile("" ::: "memory");
        }
        return q;
    } else {
        constexpr int HC = DP / 2;
        const T* mu = ps + gauss_mu_offset(DP);
        T d[2][HC];
#pragma unroll
        for (int par = 0; par < 2; ++par)
#pragma unroll
            for (int c = 0; c < HC; ++c) d[par][c] = x[pair_col(c, 0) + (par << 2)] - mu[pair_col(c, 0) + (par << 2)];
        T q = 0;
#pragma unroll
        for (int j = 0; j < DP; ++j) {
            const int jb = j >> 2;
            T yh[2];
#pragma unroll
            for (int par = 0; par < 2; ++par) {
                const T* row = ps + tri_row_offset(DP, j) - (j & ~3) + (par << 2);
                const T keep = par ? T(1) : T(0);
                T y = 0;
#pragma unroll
                for (int b = 0; b < HC / 4; ++b) {
                    if (2 * b + 1 < jb) continue;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        T dv = d[par][4 * b + u];
                        if (2 * b + 1 == jb) dv *= keep;
                        y = fma(row[(b << 3) + u], dv, y);
                    }
                }
                yh[par] = y;
            }
            const T y = yh[0] + yh[1];
            q = fma(y, y, q);
            asm volatile("" ::: "memory");   // one row's loads at a time, as above
        }
        return q;
    }
}

template <typename T, int DP>
__device__ __forceinline__ T whitened_sq(const T* __restrict__ ps, const T (&x)[DP]) {
    return whitened_sq_regs<T, DP>(ps, x);
}

// Coordinates l .. l + 7 of a point row of D floats, zero past D or when !valid; `row` must exist
// either way (callers pass row 0 for an empty lane).  With D a multiple of 4 the row and every
// 4-aligned l are 16-byte aligned: two unconditional 16-byte loads (offset 0 stands in past D);
// otherwise eight unconditional 4-byte loads from clamped offsets.  No branches around the loads.
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const float gfloat;        // a float in global memory
typedef __attribute__((address_space(1))) const f32x4v gf32x4v;
template <typename F, typename V>   // F: float or gfloat, V: the matching 16-byte vector type
__device__ __forceinline__ void load_coords8_t(F* __restrict__ row, int l, int D, bool valid, float (&o)[8]) {
    if ((D & 3) == 0) {
        const bool a = valid && l + 4 <= D, b = valid && l + 8 <= D;
        const f32x4v lo = *(V*)(row + (a ? l : 0));
        const f32x4v hi = *(V*)(row + (b ? l + 4 : 0));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            o[e] = a ? lo[e] : 0.f;
            o[e + 4] = b ? hi[e] : 0.f;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = l + e;
            const float v = row[c < D ? c : 0];
            o[e] = (valid && c < D) ? v : 0.f;
        }
    }
}
__device__ __forceinline__ void load_coords8(const float* __restrict__ row, int l, int D, bool valid, float (&o)[8]) {
    load_coords8_t<const float, const f32x4v>(row, l, D, valid, o);
}
// Only for D a multiple of 4: the two 16-byte loads without the other path.
__device__ __forceinline__ void load_coords8_aligned(gfloat* __restrict__ row, int l, int D, bool valid,
                                                     float (&o)[8]) {
    const bool a = valid && l + 4 <= D, b = valid && l + 8 <= D;
    const f32x4v lo = *(gf32x4v*)(row + (a ? l : 0));
    const f32x4v hi = *(gf32x4v*)(row + (b ? l + 4 : 0));
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        o[e] = a ? lo[e] : 0.f;
        o[e + 4] = b ? hi[e] : 0.f;
    }
}
// The same for 4 coordinates l .. l + 3.
__device__ __forceinline__ void load_coords4(const float* __restrict__ row, int l, int D, bool valid, float (&o)[4]) {
    if ((D & 3) == 0) {
        const bool a = valid && l + 4 <= D;
        const f32x4v v = *reinterpret_cast<const f32x4v*>(row + (a ? l : 0));
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = a ? v[e] : 0.f;
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int c = l + e;
            const float v = row[c < D ? c : 0];
            o[e] = (valid && c < D) ? v : 0.f;
        }
    }
}
__device__ __forceinline__ void load_coords4(const double* __restrict__ row, int l, int D, bool valid, double (&o)[4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int c = l + e;
        const double v = row[c < D ? c : 0];
        o[e] = (valid && c < D) ? v : 0.0;
    }
}

// Upper bound of the Gumbel noise of a 32-bit uniform (22.87) plus slack, used by the exact
// pruning bounds of the mixed label sweep.
constexpr float kPruneNoise = 24.0f;

// Update mode of a statistics pass (all null: every chunk summed in full).  With subs_prev, the
// sub-labels of the previous pass in the order of its partition (segment of cluster k starting at
// prev_start[k]), a chunk of a cluster whose points all kept their label (moved[k] == 0) counts a point
// for sub-cluster s only when it moved into s: its segment holds the same points in the same order, at
// offset prev_start[k] - label_start[k] in the previous partition, and the partials are the moved-in
// sums from which the statistics are updated.  Chunks of clusters with moved[k] != 0 are summed in full.
struct StatsGate {
    const int32_t* subs_prev;
    const int32_t* moved;
    const int32_t* prev_start;
};
// The offset of cluster k's points in the previous partition (negative when its segment moved up), or
// kFullPass when its chunks are summed in full.
constexpr int kFullPass = INT32_MIN;
__device__ __forceinline__ int stats_delta_shift(const StatsGate& g, int k, const int32_t* __restrict__ label_start) {
    if (!g.subs_prev || g.moved[k]) return kFullPass;
    return g.prev_start[k] - label_start[k];
}

// Chunk c of the partition: the cluster k whose segment holds it and its point range [beg, end)
// in perm.  False past the last chunk.
__device__ __forceinline__ bool chunk_of(int c, int K, const int32_t* __restrict__ chunk_start,
                                         const int32_t* __restrict__ label_start, int CH, int& k, int& beg, int& end) {
    if (c >= chunk_start[K]) return false;
    int lo = 0, hi = K - 1;   // last k with chunk_start[k] <= c
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (chunk_start[mid] <= c) lo = mid; else hi = mid - 1;
    }
    k = lo;
    beg = label_start[k] + (c - chunk_start[k]) * CH;
    end = min(beg + CH, label_start[k + 1]);
    return true;
}

}  // namespace dpmm
