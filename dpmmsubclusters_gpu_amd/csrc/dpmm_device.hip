// MI355X (gfx950) device layer of the DPMM sub-cluster sweep and its C ABI (include/dpmm_hip.h).
//
// One iteration's device work, all on one stream:
//   label sweep     X read once; every point scored against all K clusters (whitened
//                   residual || W_k (x - mu_k) ||^2 or x . log theta_k), Gumbel-max draw
//   partition       block histograms of labels -> scans -> stable scatter of point indices
//                   (ballot ranks, no global atomics): segments of points per cluster
//   sub-label sweep per segment chunk: the cluster's two sub-cluster parameters in LDS,
//                   logistic draw of each point's sub-label
//   statistics      per (chunk, sub-cluster): count, FP64 sum and FP64 scatter matrix,
//                   then a fixed-order reduction over the chunks of each cluster
//   relabel         one table-driven pass for split / merge / reset / remove
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dpmm_hip.h"
#include "dpmm_kernels.h"
#include "dpmm_device.h"
#include "dpmm_gauss.h"
#include "dpmm_mixed.h"
#include "dpmm_niw.h"
#include "dpmm_mult.h"
#include "dpmm_step.h"

using namespace dpmm;

namespace dpmm {
DPMM_GAUSS_EXTERN(float, 4)
DPMM_GAUSS_EXTERN(float, 8)
DPMM_GAUSS_EXTERN(float, 16)
DPMM_GAUSS_EXTERN(float, 32)
DPMM_GAUSS_EXTERN(float, 48)
DPMM_GAUSS_EXTERN(float, 64)
DPMM_GAUSS_EXTERN(float, 96)
DPMM_GAUSS_EXTERN(float, 128)
DPMM_GAUSS_EXTERN(double, 4)
DPMM_GAUSS_EXTERN(double, 8)
DPMM_GAUSS_EXTERN(double, 16)
DPMM_GAUSS_EXTERN(double, 32)
DPMM_GAUSS_EXTERN(double, 48)
DPMM_GAUSS_EXTERN(double, 64)
DPMM_MX_EXTERN(32)
DPMM_MX_EXTERN(64)
DPMM_MX_EXTERN(96)
DPMM_MX_EXTERN(128)
}  // namespace dpmm

namespace {
// Allocation fault injection (dpmm_debug_fail_allocations): the number of device allocations that still
// succeed before every further one fails with hipErrorOutOfMemory; -1: off.
std::atomic<int64_t> g_alloc_budget{-1};
template <typename T>
hipError_t dev_malloc(T** p, size_t bytes) {
    int64_t b = g_alloc_budget.load();
    while (b >= 0) {
        if (b == 0) return hipErrorOutOfMemory;
        if (g_alloc_budget.compare_exchange_weak(b, b - 1)) break;
    }
    return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
thread_local std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(DPMM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));            \
    } while (0)

constexpr int kBlockPts = 4096;    // points per partition block (256 threads x 16)
constexpr int kMaxK = 2048;        // clusters per context
// clusters of a multinomial context: the FP32 label sweep keeps 32 documents' counts (65 floats
// apart) and K x 33 floats of cluster terms in 64 KiB of LDS
constexpr int kMaxKMult = (16384 - 32 * 65) / 33;
}  // namespace

// ============================================================================ kernels ======

__global__ void __launch_bounds__(kThreads) k_init_labels(int32_t* labels, int32_t* subs, int64_t N, int K0,
                                                          uint32_t key0, uint32_t key1) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const Ctr4 w = philox4x32_10((uint32_t)i, 0, 0, TAG_INIT, key0, key1);
    int l = (int)(u01_f64(w.v[0]) * (double)K0);
    labels[i] = l < K0 - 1 ? l : K0 - 1;
    subs[i] = (int32_t)(w.v[1] & 1u);
}

// Multinomial label sweep: 64 points x K clusters per workgroup, x . log theta accumulated over
// 32-dimension chunks staged in LDS.  Thread (p, g) owns clusters g, g + 4, ...
template <int KPT>
__global__ void __launch_bounds__(kThreads) k_label_mult(const int32_t* __restrict__ X, int64_t N, int D, int K,
                                                         const float* __restrict__ P, int stride, uint32_t key0,
                                                         uint32_t key1, uint32_t iter, int final_iter,
                                                         int32_t* __restrict__ labels) {
    extern __shared__ __align__(16) unsigned char smem[];
    float* xs = reinterpret_cast<float*>(smem);        // [32][65]
    float* th = xs + 32 * 65;                           // [K][33]
    __shared__ float bs[4][64];
    __shared__ int bk[4][64];
    const int p = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * 64;
    float acc[KPT];
#pragma unroll
    for (int j = 0; j < KPT; ++j) acc[j] = 0.f;
    for (int d0 = 0; d0 < D; d0 += 32) {
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 32; e += kThreads) {
            const int pp = e >> 5, dd = e & 31;
            const int64_t i = base + pp;
            xs[dd * 65 + pp] = (i < N && d0 + dd < D) ? (float)X[i * D + d0 + dd] : 0.f;
        }
        for (int e = threadIdx.x; e < K * 32; e += kThreads) {
            const int kk = e >> 5, dd = e & 31;
            th[kk * 33 + dd] = (d0 + dd < D) ? P[(size_t)kk * stride + d0 + dd] : 0.f;
        }
        __syncthreads();
        for (int dd = 0; dd < 32; ++dd) {
            const float xv = xs[dd * 65 + p];
#pragma unroll
            for (int j = 0; j < KPT; ++j) {
                const int k = g + 4 * j;
                if (k < K) acc[j] = fmaf(xv, th[k * 33 + dd], acc[j]);
            }
        }
    }
    const int64_t i = base + p;
    float best = neg_inf<float>();
    int bi = 0x7fffffff;
#pragma unroll
    for (int j = 0; j < KPT; ++j) {
        const int k = g + 4 * j;
        if (k < K) {
            float s = acc[j] + P[(size_t)k * stride + D];
            if (!final_iter) s += gumbel_f32(philox4x32_10((uint32_t)i, (uint32_t)j, iter, TAG_LABEL, key0, key1).v[g]);
            if (s > best || bi == 0x7fffffff) { best = s; bi = k; }
        }
    }
    bs[g][p] = best;
    bk[g][p] = bi;
    __syncthreads();
    if (g == 0 && i < N) {
        float b = bs[0][p];
        int kb = bk[0][p];
        for (int gg = 1; gg < 4; ++gg) {
            const int k2 = bk[gg][p];
            if (k2 == 0x7fffffff) continue;
            const float s2 = bs[gg][p];
            if (kb == 0x7fffffff || s2 > b || (s2 == b && k2 < kb)) { b = s2; kb = k2; }
        }
        labels[i] = kb;
    }
}

// ---- deterministic partition of points by label -----------------------------------------------
// clamp != 0 files labels outside [0, K) under label 0 instead of leaving them out.
// cond (null: always) on the partition kernels: they run only when *cond != 0, a device flag that the
// pruned label sweep sets when a label moved; otherwise the partition of the labels before the sweep
// stands (a stable sort of unchanged labels gives it again).
__global__ void __launch_bounds__(kThreads) k_block_hist(const int32_t* __restrict__ labels, int64_t N, int K,
                                                         int nblk, int clamp, int32_t* __restrict__ hist,
                                                         const int* __restrict__ cond) {
    if (cond && !*cond) return;
    __shared__ int cnt[kMaxK];
    for (int k = threadIdx.x; k < K; k += kThreads) cnt[k] = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * kBlockPts;
    for (int r = 0; r < kBlockPts / kThreads; ++r) {
        const int64_t i = b0 + r * kThreads + threadIdx.x;
        if (i < N) {
            int L = labels[i];
            if (clamp && (unsigned)L >= (unsigned)K) L = 0;
            if ((unsigned)L < (unsigned)K) atomicAdd(&cnt[L], 1);   // out-of-range labels are left out
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += kThreads) hist[(size_t)k * nblk + blockIdx.x] = cnt[k];
}

// Exclusive scan of one label's row of block counts (one workgroup per label).
__global__ void __launch_bounds__(kThreads) k_scan_rows(int32_t* __restrict__ hist, int nblk,
                                                        int32_t* __restrict__ rowtot, const int* __restrict__ cond) {
    if (cond && !*cond) return;
    __shared__ int buf[kThreads];
    __shared__ int carry;
    int32_t* row = hist + (size_t)blockIdx.x * nblk;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblk; b0 += kThreads) {
        const int b = b0 + threadIdx.x;
        const int v = b < nblk ? row[b] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < kThreads; off <<= 1) {
            const int t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        const int incl = buf[threadIdx.x];
        if (b < nblk) row[b] = carry + incl - v;
        __syncthreads();
        if (threadIdx.x == kThreads - 1) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) rowtot[blockIdx.x] = carry;
}

// Segment starts per label and chunk starts (ceil(n_k / CH) chunks per label).
__global__ void k_scan_labels(const int32_t* __restrict__ rowtot, int K, int CH, int32_t* __restrict__ label_start,
                              int32_t* __restrict__ chunk_start, const int* __restrict__ cond) {
    if (threadIdx.x != 0 || (cond && !*cond)) return;
    int s = 0, c = 0;
    for (int k = 0; k < K; ++k) {
        label_start[k] = s;
        chunk_start[k] = c;
        const int n = rowtot[k];
        s += n;
        c += (n + CH - 1) / CH;
    }
    label_start[K] = s;
    chunk_start[K] = c;
}

// Stable scatter: within a block, points keep index order per label (ranks within a wave, waves in
// order, rounds in order).  A lane's rank among the wave's lanes of its label comes from a 64-bit
// lane mask per (wave, label) in LDS that every lane ORs its bit into: rank = popcount of the lower
// lanes' bits, count = popcount of all, whatever the number of distinct labels in the wave.
// Dynamic LDS: base[K] ints, wc[4][K] ints, then (8-byte aligned) masks[4][K].
__global__ void __launch_bounds__(kThreads) k_scatter(const int32_t* __restrict__ labels, int64_t N, int K, int nblk,
                                                      int clamp, const int32_t* __restrict__ hist,
                                                      const int32_t* __restrict__ label_start,
                                                      int32_t* __restrict__ perm, const int* __restrict__ cond) {
    if (cond && !*cond) return;
    extern __shared__ __align__(16) unsigned char smem[];
    int* base = reinterpret_cast<int*>(smem);
    int* wc = base + K;   // [4][K]
    unsigned long long* masks = reinterpret_cast<unsigned long long*>(smem + (((size_t)5 * K * 4 + 15) & ~(size_t)15));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long* wmask = masks + (size_t)wv * K;
    for (int k = threadIdx.x; k < K; k += kThreads) base[k] = label_start[k] + hist[(size_t)k * nblk + blockIdx.x];
    for (int e = threadIdx.x; e < 4 * K; e += kThreads) masks[e] = 0ull;
    const int64_t b0 = (int64_t)blockIdx.x * kBlockPts;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kBlockPts / kThreads; ++r) {
        for (int e = threadIdx.x; e < 4 * K; e += kThreads) wc[e] = 0;
        __syncthreads();
        const int64_t i = b0 + r * kThreads + threadIdx.x;
        int Li = i < N ? labels[i] : -1;
        if (clamp && i < N && (unsigned)Li >= (unsigned)K) Li = 0;
        const bool valid = (unsigned)Li < (unsigned)K;
        int rank = 0, cnt = 0;
        if (valid) atomicOr(&wmask[Li], 1ull << lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (valid) {
            const unsigned long long m = wmask[Li];
            rank = __popcll(m & lt);
            cnt = __popcll(m);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (valid && rank == 0) {
            wc[wv * K + Li] = cnt;
            wmask[Li] = 0ull;   // every lane of the label has read it
        }
        __syncthreads();
        if (valid) {
            int pos = base[Li] + rank;
            for (int w2 = 0; w2 < wv; ++w2) pos += wc[w2 * K + Li];
            perm[pos] = (int32_t)i;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < K; k += kThreads) base[k] += wc[k] + wc[K + k] + wc[2 * K + k] + wc[3 * K + k];
        __syncthreads();
    }
}

// ---- multinomial rows --------------------------------------------------------------------------
// A document's counts are read V words at a time: V = 4 (one 16-byte load per lane) when the
// vocabulary size is a multiple of 4, so every row starts 16-byte aligned, else V = 1.
typedef int32_t i32x4v __attribute__((ext_vector_type(4)));
template <int V> struct IRow;
template <> struct IRow<1> {
    typedef int32_t T;
    static __device__ __forceinline__ int32_t at(T x, int) { return x; }
};
template <> struct IRow<4> {
    typedef i32x4v T;
    static __device__ __forceinline__ int32_t at(T x, int v) { return x[v]; }
};

// Lane `lane`'s share of the two sub-cluster dot products of one document: words
// lane*V + 64*V*step + v in increasing (step, v) order, zero counts skipped (so a -inf log theta of
// an absent word adds nothing).  The sub-label sweep and the score entry point both sum this way.
template <int V>
__device__ __forceinline__ void mult_sub_partial(typename IRow<V>::T x, int col, const float* __restrict__ th,
                                                 int stride, float& a0, float& a1) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const float xv = (float)IRow<V>::at(x, v);
        if (xv != 0.f) { a0 = fmaf(xv, th[col + v], a0); a1 = fmaf(xv, th[stride + col + v], a1); }
    }
}

// ---- sub-label sweeps over segments ----------------------------------------------------------
// Multinomial: one wavefront per document, lanes split the vocabulary V words at a time; each wave
// takes two documents at once and issues all of a 256 V-word stretch of both rows before using
// them, so eight row loads per wave are in flight.
template <int V>
__global__ void __launch_bounds__(kThreads) k_sub_mult(const int32_t* __restrict__ X, int D, int K, int CH,
                                                       const int32_t* __restrict__ perm,
                                                       const int32_t* __restrict__ label_start,
                                                       const int32_t* __restrict__ chunk_start,
                                                       const float* __restrict__ SP, int stride, uint32_t key0,
                                                       uint32_t key1, uint32_t iter, int32_t* __restrict__ subs) {
    typedef typename IRow<V>::T VT;
    extern __shared__ __align__(16) unsigned char smem[];
    float* th = reinterpret_cast<float*>(smem);   // [2][stride]
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end)) return;
    const float* src = SP + (size_t)(2 * k) * stride;
    for (int e = threadIdx.x; e < 2 * stride; e += kThreads) th[e] = src[e];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int p = beg + wv; p < end; p += 8) {
        const bool two = p + 4 < end;
        const int i0 = perm[p];
        const int i1 = two ? perm[p + 4] : i0;
        const VT* r0 = reinterpret_cast<const VT*>(X + (int64_t)i0 * D);
        const VT* r1 = reinterpret_cast<const VT*>(X + (int64_t)i1 * D);
        float a00 = 0.f, a01 = 0.f, a10 = 0.f, a11 = 0.f;
        for (int c0 = 0; c0 < D; c0 += 4 * 64 * V) {
            VT x0[4], x1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int col = c0 + (u * 64 + lane) * V;
                const int g = col < D ? col / V : 0;
                x0[u] = r0[g];
                x1[u] = r1[g];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int col = c0 + (u * 64 + lane) * V;
                if (col < D) {
                    mult_sub_partial<V>(x0[u], col, th, stride, a00, a01);
                    mult_sub_partial<V>(x1[u], col, th, stride, a10, a11);
                }
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            a00 += __shfl_xor(a00, off); a01 += __shfl_xor(a01, off);
            a10 += __shfl_xor(a10, off); a11 += __shfl_xor(a11, off);
        }
        if (lane == 0) {
            uint32_t w = philox4x32_10((uint32_t)i0, 0, iter, TAG_SUB, key0, key1).v[0];
            subs[i0] = (logit_u32<float>(w) < (a00 + th[D]) - (a01 + th[stride + D])) ? 0 : 1;
            if (two) {
                w = philox4x32_10((uint32_t)i1, 0, iter, TAG_SUB, key0, key1).v[0];
                subs[i1] = (logit_u32<float>(w) < (a10 + th[D]) - (a11 + th[stride + D])) ? 0 : 1;
            }
        }
    }
}

// Noise-free multinomial scores of documents [first, first + n), one wave per document, with the
// sweeps' FP32 arithmetic: which = 0 the n x K cluster scores of k_label_mult (per cluster the
// fmaf chain over the vocabulary in order, then the slot's constant), which = 1 the n x 2 scores of
// the two sub-clusters of the document's label as k_sub_mult<V> forms them (mult_sub_partial's
// lane partial sums, butterfly reduction, constant).
template <int V>
__global__ void __launch_bounds__(kThreads) k_score_mult(const int32_t* __restrict__ X, int D, int K,
                                                         const float* __restrict__ P, const float* __restrict__ SP,
                                                         int stride, const int32_t* __restrict__ labels,
                                                         int64_t first, int64_t n, int which,
                                                         float* __restrict__ out) {
    typedef typename IRow<V>::T VT;
    const int lane = threadIdx.x & 63;
    for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < n; w += (int64_t)gridDim.x * 4) {
        const int64_t i = first + w;
        const int32_t* x = X + i * (int64_t)D;
        if (which == 0) {
            for (int k = lane; k < K; k += 64) {
                const float* th = P + (size_t)k * stride;
                float a = 0.f;
                for (int d = 0; d < D; ++d) a = fmaf((float)x[d], th[d], a);
                out[w * K + k] = a + th[D];
            }
        } else {
            const float* th = SP + (size_t)(2 * labels[i]) * stride;
            float a0 = 0.f, a1 = 0.f;
            for (int col = lane * V; col < D; col += 64 * V)
                mult_sub_partial<V>(reinterpret_cast<const VT*>(x)[col / V], col, th, stride, a0, a1);
            for (int off = 32; off > 0; off >>= 1) { a0 += __shfl_xor(a0, off); a1 += __shfl_xor(a1, off); }
            if (lane == 0) {
                out[w * 2] = a0 + th[D];
                out[w * 2 + 1] = a1 + th[stride + D];
            }
        }
    }
}

// ---- sufficient statistics --------------------------------------------------------------------
// Multinomial: exact integer sums per vocabulary entry, both sub-clusters of a chunk in one block.
// Thread t owns words [t V, t V + V) of each 256 V-word stretch and every wave walks all of the
// chunk's documents: a wave reads 64 (index, sub-label) pairs with one load each, then gathers the
// rows eight at a time with the document index broadcast from a lane (readlane, so the row address
// is scalar) and adds each row into the sub-cluster's int64 accumulators.  Integer sums are exact,
// so the order does not matter; partials are written as (2 chunk + sub-cluster) rows.
template <int V>
__global__ void __launch_bounds__(kThreads) k_stats_mult(const int32_t* __restrict__ X, int D, int K, int CH,
                                                         const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ label_start,
                                                         const int32_t* __restrict__ chunk_start,
                                                         const int32_t* __restrict__ subs, int E,
                                                         double* __restrict__ partials,
                                                         const int32_t* __restrict__ dirty) {
    typedef typename IRow<V>::T VT;
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end)) return;
    if (dirty && !dirty[k]) return;   // statistics of this cluster are kept from the last pass
    double* out0 = partials + (size_t)(2 * blockIdx.x) * E;
    double* out1 = out0 + E;
    const int lane = threadIdx.x & 63;
    for (int c0 = 0; c0 < D; c0 += kThreads * V) {
        const int col = c0 + threadIdx.x * V;
        const int g = col < D ? col / V : 0;
        int64_t a0[V], a1[V];
#pragma unroll
        for (int v = 0; v < V; ++v) { a0[v] = 0; a1[v] = 0; }
        int64_t n0 = 0, n1 = 0;
        for (int p0 = beg; p0 < end; p0 += 64) {
            const int p = p0 + lane;
            const int idx = p < end ? perm[p] : 0;
            const int s = p < end ? subs[idx] : -1;
            n0 += __popcll(__ballot(s == 0));
            n1 += __popcll(__ballot(s == 1));
            const int n = min(64, end - p0);
            for (int j = 0; j < n; j += 8) {
                VT xv[8];
                int sv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int jj = min(j + u, n - 1);
                    const int ij = __builtin_amdgcn_readlane(idx, jj);
                    const int sj = __builtin_amdgcn_readlane(s, jj);
                    sv[u] = j + u < n ? sj : -1;
                    xv[u] = reinterpret_cast<const VT*>(X + (int64_t)ij * D)[g];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const int64_t x = IRow<V>::at(xv[u], v);
                        a0[v] += sv[u] == 0 ? x : 0;
                        a1[v] += sv[u] == 1 ? x : 0;
                    }
                }
            }
        }
        if (col < D) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                out0[1 + col + v] = (double)a0[v];
                out1[1 + col + v] = (double)a1[v];
            }
        }
        if (c0 == 0 && threadIdx.x == 0) { out0[0] = (double)n0; out1[0] = (double)n1; }
    }
}

// Fixed-order reduction of the chunk partials of each (cluster, sub-cluster): one thread per
// statistic entry, summing that entry over the cluster's chunks in chunk order.
__global__ void __launch_bounds__(kThreads) k_stats_reduce(const double* __restrict__ partials,
                                                           const int32_t* __restrict__ chunk_start, int K, int E,
                                                           double* __restrict__ stats,
                                                           const int32_t* __restrict__ dirty) {
    const int q = blockIdx.x, k = q >> 1, s = q & 1;
    const int e = blockIdx.y * kThreads + threadIdx.x;
    if (e >= E || (dirty && !dirty[k])) return;
    const int c0 = chunk_start[k], c1 = chunk_start[k + 1];
    double v = 0.0;
    const double* src = partials + (size_t)s * E + e;
    int c = c0;
    for (; c + 8 <= c1; c += 8) {   // eight loads in flight, added in chunk order
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = src[(size_t)(2 * (c + u)) * E];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += t[u];
    }
    for (; c < c1; ++c) v += src[(size_t)(2 * c) * E];
    stats[(size_t)q * E + e] = v;
}

// The reduction of a pass in update mode (StatsGate): a cluster whose points kept their labels gets
// d, the change of entry (k, 0), summed over its chunks in chunk order: entry (k, 0) gains d and (k, 1)
// loses it.  With signed_slot the chunk partials of slot 0 hold d's terms (points moved into 0 counted
// +, into 1 -); else slot s holds the sums of the points that moved into s and d = slot 0 - slot 1.  A
// cluster with moved[k] != 0 gets the full sums of both slots as k_stats_reduce forms them.
__global__ void __launch_bounds__(kThreads) k_stats_update(const double* __restrict__ partials,
                                                           const int32_t* __restrict__ chunk_start, int K, int E,
                                                           double* __restrict__ stats,
                                                           const int32_t* __restrict__ moved, int signed_slot,
                                                           const int32_t* __restrict__ dirty) {
    const int k = blockIdx.x;
    const int e = blockIdx.y * kThreads + threadIdx.x;
    if (e >= E || (dirty && !dirty[k])) return;
    const int c0 = chunk_start[k], c1 = chunk_start[k + 1];
    const bool full = moved[k] != 0, two = full || !signed_slot;
    double v0 = 0.0, v1 = 0.0;
    const double* src = partials + e;
    int c = c0;
    for (; c + 8 <= c1; c += 8) {   // eight chunks' loads in flight, added in chunk order
        double t0[8], t1[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            t0[u] = src[(size_t)(2 * (c + u)) * E];
            t1[u] = two ? src[(size_t)(2 * (c + u) + 1) * E] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            v0 += t0[u];
            v1 += t1[u];
        }
    }
    for (; c < c1; ++c) {
        v0 += src[(size_t)(2 * c) * E];
        if (two) v1 += src[(size_t)(2 * c + 1) * E];
    }
    if (full) {
        stats[(size_t)(2 * k) * E + e] = v0;
        stats[(size_t)(2 * k + 1) * E + e] = v1;
    } else {
        const double d = signed_slot ? v0 : v0 - v1;
        stats[(size_t)(2 * k) * E + e] += d;
        stats[(size_t)(2 * k + 1) * E + e] -= d;
    }
}

// After a relabel that redrew the sub-labels of the clusters flagged in `mask` (resets), the sub-labels
// in partition order of those clusters' chunks: the old ones into prev, the new ones (by index) into cur.
__global__ void __launch_bounds__(kThreads) k_refresh_subs_perm(int K, int CH, const int32_t* __restrict__ perm,
                                                                const int32_t* __restrict__ label_start,
                                                                const int32_t* __restrict__ chunk_start,
                                                                const int32_t* __restrict__ subs,
                                                                const int32_t* __restrict__ mask,
                                                                int32_t* __restrict__ cur, int32_t* __restrict__ prev) {
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end) || !mask[k]) return;
    for (int p = beg + threadIdx.x; p < end; p += kThreads) {
        prev[p] = cur[p];
        cur[p] = subs[perm[p]];
    }
}

// Sub-labels in the order of the current partition from those the fused label + sub-label sweep left
// in the order of the partition it swept (segment starts old_start): a cluster no label moved into or
// out of (moved[k] == 0) holds the same points in the same order there, at offset old_start[k] -
// label_start[k]; the points of the other clusters are gathered by index.
__global__ void __launch_bounds__(kThreads) k_subs_to_partition(int K, int CH, const int32_t* __restrict__ perm,
                                                                const int32_t* __restrict__ label_start,
                                                                const int32_t* __restrict__ chunk_start,
                                                                const int32_t* __restrict__ old_start,
                                                                const int32_t* __restrict__ moved,
                                                                const int32_t* __restrict__ subs_old,
                                                                const int32_t* __restrict__ subs,
                                                                int32_t* __restrict__ cur) {
    int k, beg, end;
    if (!chunk_of(blockIdx.x, K, chunk_start, label_start, CH, k, beg, end)) return;
    if (moved[k]) {
        for (int p = beg + threadIdx.x; p < end; p += kThreads) cur[p] = subs[perm[p]];
    } else {
        const int shift = old_start[k] - label_start[k];
        for (int p = beg + threadIdx.x; p < end; p += kThreads) cur[p] = subs_old[p + shift];
    }
}

__global__ void k_copy_int(const int32_t* __restrict__ src, int32_t* __restrict__ dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_zero_int(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Largest value of an int32 array (one atomic per workgroup).
__global__ void __launch_bounds__(kThreads) k_max_i32(const int32_t* __restrict__ x, int64_t n, int* __restrict__ out) {
    int m = 0;
    for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n; e += (int64_t)gridDim.x * kThreads)
        m = max(m, x[e]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    __shared__ int wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, max(max(wm[0], wm[1]), max(wm[2], wm[3])));
}

// ---- relabel -----------------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads) k_relabel(int32_t* __restrict__ labels, int32_t* __restrict__ subs,
                                                      int64_t N, const int32_t* __restrict__ table, int K,
                                                      uint32_t key0, uint32_t key1, uint32_t iter) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= N) return;
    const int k = labels[i], s = subs[i];
    if ((unsigned)k >= (unsigned)K || (unsigned)s > 1u) return;
    const int act = table[2 * K + k];
    labels[i] = table[2 * k + s];
    int ns = s;
    if (act == 1 || act == 2) {
        ns = (int)(philox4x32_10((uint32_t)i, 0, iter, act == 1 ? TAG_RESET : TAG_SPLITSUB, key0, key1).v[0] & 1u);
    } else if (act == 3) {
        ns = 0;
    } else if (act == 4) {
        ns = 1;
    }
    subs[i] = ns;
}

// Bounds for the pruned mixed label sweep (k_label_gauss_mxp), from the FP32 slots of the K clusters
// (their values are exact in FP64): psc[2k] = sig_k >= ||W_k||_2 + g ||W_k||_F, psc[2k + 1] = c_k and
// pm[a K + k] <= ||W_k (mu_a - mu_k)|| - g ||W_k||_F ||mu_a - mu_k||, with g = (D + 2) u / (1 - (D + 2) u),
// u = 2^-24, bounding the FP32 sweep's rounding of W (x - mu) by g ||W||_F ||x - mu||.
// ||W||_2^2 = lambda_max(W^T W): a power iteration gives an estimate s2 from below, and t = f s2
// (f = 1.05, 1.2, 1.5, 3 in turn) is certified by an FP64 Cholesky of t I - W^T W (its success bounds
// lambda_max by t plus the factorisation's backward error, added); without a certificate the bound
// ||W||_1 ||W||_inf is kept.  One workgroup per cluster k; LDS: W (D x D floats), then four FP64
// difference vectors of max(D, 64) entries per wave, two power-iteration vectors and a packed D x D
// matrix.
__global__ void __launch_bounds__(kThreads) k_prune_prep(const float* __restrict__ P, int K, int D, int DP,
                                                         float* __restrict__ pm, float* __restrict__ psc,
                                                         double* __restrict__ pv, int kwarm) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int DS = D + 1;   // padded row stride: rows of W start in different LDS banks
    float* W = reinterpret_cast<float*>(smem);
    double* diff = reinterpret_cast<double*>(smem + (((size_t)D * DS * 4 + 15) & ~(size_t)15));
    double* v = diff + 16 * (D > 64 ? D : 64);
    double* uu = v + D;
    double* M = uu + D;
    __shared__ double red[4], red_c[4], red_r[4];
    __shared__ double sh[2];
    const int k = blockIdx.x;
    const int stride = gauss_slot_stride(DP);
    const float* slot = P + (size_t)k * stride;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // row j of W from the slot (its columns j & ~3 onwards are stored): thread t takes rows
    // t / 64, t / 64 + 4, ..., lane l its columns l, l + 64; the loads of 8 rows are issued before
    // their stores
    constexpr int kRowBatch = 8;
    for (int j0 = wv; j0 < D; j0 += 4 * kRowBatch) {
        float w[kRowBatch][2];
#pragma unroll
        for (int i = 0; i < kRowBatch; ++i) {
            const int j = j0 + 4 * i;
            const float* src = slot + tri_row_offset(DP, j < D ? j : 0) - ((j < D ? j : 0) & ~3);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int l = lane + 64 * u;
                w[i][u] = (j < D && l < D && l >= j) ? src[l] : 0.f;
            }
        }
#pragma unroll
        for (int i = 0; i < kRowBatch; ++i) {
            const int j = j0 + 4 * i;
            if (j >= D) break;
#pragma unroll
            for (int u = 0; u < 2; ++u)
                if (lane + 64 * u < D) W[j * DS + lane + 64 * u] = w[i][u];
        }
    }
    __syncthreads();
    // fallback bound ||W||_2^2 <= ||W||_1 ||W||_inf, and ||W||_F^2
    double f2 = 0.0, c1 = 0.0, r1 = 0.0;
    for (int i = threadIdx.x; i < D; i += kThreads) {
        double cs = 0.0, rw = 0.0;
        for (int j = 0; j < D; ++j) {
            cs += fabs((double)W[j * DS + i]);
            rw += fabs((double)W[i * DS + j]);
            f2 = fma((double)W[i * DS + j], (double)W[i * DS + j], f2);
        }
        c1 = fmax(c1, cs);
        r1 = fmax(r1, rw);
    }
    for (int o = 32; o > 0; o >>= 1) {
        c1 = fmax(c1, __shfl_xor(c1, o));
        r1 = fmax(r1, __shfl_xor(r1, o));
    }
    if (lane == 0) {
        red_c[wv] = c1;
        red_r[wv] = r1;
    }
    f2 = block_sum(f2, red);   // synchronises, so the maxima above are visible too
    c1 = fmax(fmax(red_c[0], red_c[1]), fmax(red_c[2], red_c[3]));
    r1 = fmax(fmax(red_r[0], red_r[1]), fmax(red_r[2], red_r[3]));
    double sig2 = c1 * r1 * (1.0 + 1e-12);
    // power iteration on W^T W (u = W v, v <- W^T u / ||W^T u||, s2 = ||W v||^2 for unit v): TPR
    // threads per row, each summing every TPR-th term, combined by butterfly shuffles
    const int TPR = D > 64 ? 2 : (D > 32 ? 4 : 8);
    const int row = threadIdx.x / TPR, part = threadIdx.x % TPR;
    // start: this slot's vector of the last call (the parameters of a cluster change little between
    // iterations, so its dominant direction does too) with fewer steps, else a fixed vector; the
    // estimate only chooses the certificate's t, so a poor start costs time, not correctness
    bool warm = k < kwarm;
    {
        const double w0 = warm && row < D && part == 0 ? pv[(size_t)k * DP + row] : 0.0;
        const double nw = block_sum(w0 * w0, red);
        warm = warm && nw > 0.0 && nw < 1e300;   // uniform: every thread holds the same sum
        const double v0 = row < D && part == 0 ? (warm ? w0 : 1.0 + 0.125 * (row % 7)) : 0.0;
        const double nv = warm ? nw : block_sum(v0 * v0, red);
        if (row < D && part == 0) v[row] = v0 / sqrt(nv);
    }
    __syncthreads();
    double s2 = 0.0;
    const int nit = warm ? 6 : 14;
    for (int itp = 0; itp < nit; ++itp) {
        double uj = 0.0;
        if (row < D)
#pragma unroll 8
            for (int l = row + part; l < D; l += TPR) uj = fma((double)W[row * DS + l], v[l], uj);
        for (int o = 1; o < TPR; o <<= 1) uj += __shfl_xor(uj, o);
        if (row < D && part == 0) uu[row] = uj;
        if (itp == nit - 1) s2 = block_sum(row < D && part == 0 ? uj * uj : 0.0, red);   // also publishes uu
        else __syncthreads();                                                      // publishes uu
        double zi = 0.0;
        if (row < D)
#pragma unroll 8
            for (int j = part; j <= row; j += TPR) zi = fma((double)W[j * DS + row], uu[j], zi);
        for (int o = 1; o < TPR; o <<= 1) zi += __shfl_xor(zi, o);
        const double nz = block_sum(row < D && part == 0 ? zi * zi : 0.0, red);
        if (row < D && part == 0 && nz > 0.0) v[row] = zi / sqrt(nz);
        __syncthreads();
    }
    if (row < D && part == 0) pv[(size_t)k * DP + row] = v[row];
    for (int tr = 0; tr < 4; ++tr) {   // uniform: s2 and sig2 are the same in every thread
        const double tt = s2 * (tr == 0 ? 1.05 : tr == 1 ? 1.2 : tr == 2 ? 1.5 : 3.0);
        if (!(tt > 0.0) || tt >= sig2) break;
        const int T = D * (D + 1) / 2;
        for (int e = threadIdx.x; e < T; e += kThreads) {
            int i, j;
            tri_ij(e, i, j);   // j <= i
            double lam = 0.0;
#pragma unroll 8
            for (int l = 0; l <= j; ++l) lam = fma((double)W[l * DS + i], (double)W[l * DS + j], lam);
            M[e] = (i == j ? tt : 0.0) - lam;
        }
        __syncthreads();
        if (chol_packed(M, D, sh)) {
            sig2 = tt + 1e-12 * (f2 + (double)D * tt);
            break;
        }
    }
    const double u = 0x1p-24;
    const double gam = (D + 2) * u / (1.0 - (D + 2) * u);
    const double F = sqrt(f2) * (1.0 + 1e-12);
    if (threadIdx.x == 0) {
        psc[2 * k] = __double2float_ru(sqrt(sig2) * (1.0 + 1e-12) + gam * F);
        psc[2 * k + 1] = slot[gauss_c_offset(DP)];
    }
    const float* muk = slot + gauss_mu_offset(DP);
    auto sync_wave = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto put_bound = [&](int a, double sy, double se) {
        const double m = sqrt(sy) * (1.0 - 1e-12) - gam * F * sqrt(se) * (1.0 + 1e-12);
        pm[(size_t)a * K + k] = a == k ? 0.f : __double2float_rd(fmax(m, 0.0));
    };
    if (D <= 64) {
        // lane j keeps row j of W (zero left of the diagonal and past column D, so the full sum over
        // 64 columns is the triangular one) in registers; each wave scores 4 anchors at a time (a,
        // a + 4, a + 8, a + 12) from 4 difference vectors of 64 entries (zero past D) in its own LDS,
        // the loads of their means issued together, and reads them 8 columns at a time
        float wr[64];
#pragma unroll
        for (int l = 0; l < 64; ++l) wr[l] = (lane < D && l < D) ? W[lane * DS + l] : 0.f;
        const float mkl = lane < D ? muk[lane] : 0.f;
        const int mu_off = gauss_mu_offset(DP);
        double* dv = diff + (size_t)wv * 4 * 64;
        for (int a0 = wv; a0 < K; a0 += 16) {
            float ma[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int a = a0 + 4 * g;
                ma[g] = (a < K && lane < D) ? P[(size_t)a * stride + mu_off + lane] : 0.f;
            }
            double se[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const double d = lane < D ? (double)ma[g] - (double)mkl : 0.0;
                dv[g * 64 + lane] = d;
                se[g] = d * d;
            }
            sync_wave();
            double y[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int l0 = 0; l0 < 64; l0 += 8) {
                double dd[4][8];
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int i = 0; i < 8; ++i) dd[g][i] = dv[g * 64 + l0 + i];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const double wl = (double)wr[l0 + i];
#pragma unroll
                    for (int g = 0; g < 4; ++g) y[g] = fma(wl, dd[g][i], y[g]);
                }
            }
            double sy[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) sy[g] = y[g] * y[g];
            for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    sy[g] += __shfl_xor(sy[g], o);
                    se[g] += __shfl_xor(se[g], o);
                }
            }
            if (lane == 0) {
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    if (a0 + 4 * g < K) put_bound(a0 + 4 * g, sy[g], se[g]);
            }
            sync_wave();   // dv is read before the next anchors' differences overwrite it
        }
        return;
    }
    double* dv = diff + (size_t)wv * D;
    // anchors a = wave, wave + 4, ...: each wave works on its own difference vector with wave-level
    // ordering only, and loads the next anchor's mean while it finishes the current one
    float mk[2], ma[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int l = lane + 64 * u;
        mk[u] = l < D ? muk[l] : 0.f;
        ma[u] = (wv < K && l < D) ? P[(size_t)wv * stride + gauss_mu_offset(DP) + l] : 0.f;
    }
    for (int a = wv; a < K; a += 4) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
            if (lane + 64 * u < D) dv[lane + 64 * u] = (double)ma[u] - (double)mk[u];
        if (a + 4 < K) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int l = lane + 64 * u;
                ma[u] = l < D ? P[(size_t)(a + 4) * stride + gauss_mu_offset(DP) + l] : 0.f;
            }
        }
        sync_wave();
        double sy = 0.0, se = 0.0;
        for (int j = lane; j < D; j += 64) {
            double y = 0.0;
#pragma unroll 8
            for (int l = j; l < D; ++l) y = fma((double)W[j * DS + l], dv[l], y);
            sy = fma(y, y, sy);
            se = fma(dv[j], dv[j], se);
        }
        for (int o = 32; o > 0; o >>= 1) {
            sy += __shfl_xor(sy, o);
            se += __shfl_xor(se, o);
        }
        if (lane == 0) put_bound(a, sy, se);
        sync_wave();   // dv is read before the next anchor's difference overwrites it
    }
}

// ---- verbose diagnostics: label x class contingency table ------------------------------------
// Counts of (label, ground-truth class) pairs for the NMI that dp_parallel reports per iteration in
// verbose mode (dp_parallel_sampling.cpp:164-180).  Each workgroup histograms a grid-strided share of
// the points into an LDS table (LDS atomics only) and writes it as one partial row; k_table_reduce
// sums the rows in workgroup order, so the table is exact and deterministic.
__global__ void __launch_bounds__(kThreads) k_contingency(const int32_t* __restrict__ labels,
                                                          const int32_t* __restrict__ classes, int64_t N, int K,
                                                          int G, int32_t* __restrict__ partials) {
    extern __shared__ int tab[];   // K x G
    const int KG = K * G;
    for (int e = threadIdx.x; e < KG; e += kThreads) tab[e] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < N; i += (int64_t)gridDim.x * kThreads) {
        const int L = labels[i], g = classes[i];
        if ((unsigned)L < (unsigned)K && (unsigned)g < (unsigned)G) atomicAdd(&tab[L * G + g], 1);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < KG; e += kThreads) partials[(size_t)blockIdx.x * KG + e] = tab[e];
}

__global__ void __launch_bounds__(kThreads) k_table_reduce(const int32_t* __restrict__ partials, int rows, int KG,
                                                           int64_t* __restrict__ table) {
    const int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= KG) return;
    int64_t s = 0;
    for (int b = 0; b < rows; ++b) s += partials[(size_t)b * KG + e];
    table[e] = s;
}

// ============================================================================ context ======

struct dpmm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t N = 0;
    int D = 0, DP = 0, dtype = 0, prior = 0, precision = DPMM_PREC_FP32;
    uint64_t seed = 0;
    void* dX = nullptr;
    bool owns_X = false;
    int32_t* dLabels = nullptr;
    int32_t* dSubs = nullptr;
    // sub-labels in the order of the current partition (dPerm), left by the mixed sub-label sweep for the
    // statistics pass; valid until the labels, the sub-labels or the partition change
    int32_t* dSubsPerm = nullptr;
    size_t subsPermCap = 0;
    bool subs_perm_valid = false;
    // parameters (compute type: double for FP64 precision, float otherwise)
    int K = 0, Kcap = 0, stride = 0;
    size_t capBytes = 0;   // bytes of dP (dSP holds twice as many)
    int kcap_limit = 0;   // dpmm_limit_clusters (0: the kernels' own cap)
    bool params_f64 = false;
    void* dP = nullptr;
    void* dSP = nullptr;
    std::vector<unsigned char> hP, hSP;
    // partition
    int nblk = 0, CH = 0, Kpart = 0, Kpartition = 0;
    bool partition_valid = false;
    int32_t* dHist = nullptr;
    size_t histCap = 0;
    int32_t* dRowtot = nullptr;
    int32_t* dLabelStart = nullptr;
    int32_t* dChunkStart = nullptr;
    int32_t* dPerm = nullptr;
    int32_t* dTable = nullptr;
    size_t tableCap = 0;
    // incremental statistics (behind the pruned mixed label sweep): that sweep sets dFlags[0] when a label
    // moved and dFlags[1 + k] for both clusters of every move; dSubsPermPrev keeps the previous pass's
    // sub-labels in the order of its partition, whose segment starts dPrevStart keeps.  inc_base: dStats
    // holds the statistics of the current partition with the sub-labels in dSubsPerm.
    int* dFlags = nullptr;     // then, from dFlags + 1 + K, the fused sweep's 2K sub-label list lengths
    size_t flagsCap = 0;
    int32_t* dSubsPermPrev = nullptr;
    size_t subsPermPrevCap = 0;
    int32_t* dPrevStart = nullptr;
    size_t prevStartCap = 0;
    bool label_flag = false;   // the last label sweep set dFlags[0] and left the partition of the labels before it
    bool inc_base = false;
    bool inc_ready = false;    // this iteration's sub-label sweep moved the baseline sub-labels to dSubsPermPrev
    int stats_refresh = 8;     // statistics passes per full pass (<= 1: every pass full)
    int inc_passes = 0;        // passes since the last full one
    StatsGate stats_gate{};    // of the statistics kernels being enqueued
    bool last_inc = false;     // the last statistics pass enqueued the update from the moved points
    bool timed_inc = false;    // so did the last pass the statistics phase events time (not a resets-only pass)
    // a relabel that redrew only the sub-labels of some clusters (resets) came after a pass that left
    // inc_base: the next pass updates those clusters from their moved sub-labels
    bool reset_update_ok = false;
    // statistics
    double* dPartials = nullptr;
    size_t partialsCap = 0;
    double* dStats = nullptr;
    size_t statsCap = 0;
    double* hStats = nullptr;
    size_t hStatsCap = 0;
    // mixed precision: BF16 screen slots (padded dimension DPM, a multiple of 32) + recheck list
    int DPM = 0;
    void* dPB = nullptr;
    size_t pbCap = 0;
    std::vector<unsigned char> hPB;
    void* dSPB = nullptr;   // the 2K sub-clusters' BF16 slots
    // lo tiles of the K clusters' W (W - bf16(W) in BF16, the slot's tile order): the pruned sweep's
    // split-precision refinement of the points its BF16 screen leaves undecided
    void* dPBlo = nullptr;
    size_t pbloCap = 0;
    std::vector<unsigned char> hPBlo;
    size_t spbCap = 0;
    std::vector<unsigned char> hSPB;
    bool spb_valid = false;
    int32_t* dRecheck = nullptr;
    int* dRecheckCount = nullptr;
    // sub-label screen fused into the pruned label sweep (DPM <= 64): the sub-labels it leaves in the
    // order of the partition it swept, and its list of positions left to the FP32 draw
    int fuse_mode = 1;            // 0: separate sub-label sweep
    int prune_split_fused = 2;    // workgroups per partition chunk of the fused sweep
    int prune_split = 0;          // dpmm_set_sweep_workgroups (0: the defaults)
    int refine_mode = 1;          // split-BF16 refinement of undecided waves in the pruned sweep
    bool fused_subs = false;      // the last label sweep drew the sub-labels as well, for iteration fused_iter
    uint32_t fused_iter = 0;
    int32_t* dSubsOld = nullptr;
    size_t subsOldCap = 0;
    int32_t* dSubList = nullptr;
    size_t subListCap = 0;
    float tau = 1.0f;   // scale of the mixed label screens' error bound (1: the derived bound)
    // exact pruning of the mixed Gaussian label sweep: bounds of the current parameters
    int prune_mode = 1;           // 0 off, 1 adaptive, 2 always
    bool prune_valid = false;
    int prune_skip = 0;           // adaptive: unpruned sweeps left before pruning is tried again
    unsigned long long* hEvals = nullptr;   // pinned copy of dEvals, read at the next sweep
    // pinned staging of the per-iteration transfers of the log-ML batch and the parameter draws
    // (members, log weights, results, draw status): a copy from pageable memory blocks the host
    // and goes through a driver bounce buffer.  Each call that uses them ends synchronised.
    void* pinMembers = nullptr;
    size_t pinMembersCap = 0;
    void* pinLogW = nullptr;
    size_t pinLogWCap = 0;
    void* pinOut = nullptr;
    size_t pinOutCap = 0;
    hipEvent_t evEvals = nullptr;
    bool evals_pending = false;
    int evals_K = 0;
    bool evals_copy_due = false;   // a captured pruned sweep's counters still to be copied back
    // graph mode: the sweep segment recorded into a hipGraph each iteration, replayed through one
    // executable graph that is updated in place while the recorded topology stays the same
    bool graph_mode = false;
    int32_t* dClasses = nullptr;   // ground-truth classes of the points (verbose NMI)
    int nClasses = 0;
    hipGraphExec_t gexec = nullptr;
    int64_t graph_instantiations = 0, graph_updates = 0;
    float* dPruneM = nullptr;     // K x K: row a, column k = lower bound of ||W_k (mu_a - mu_k)||
    size_t pruneMCap = 0;
    float* dPruneSC = nullptr;    // 2K: sigma_k, c_k
    size_t pruneSCCap = 0;
    double* dPruneV = nullptr;    // kMaxK x DP: each slot's last power-iteration vector (warm start)
    size_t pruneVCap = 0;
    int prune_v_k = 0;            // slots holding a vector
    unsigned long long* dEvals = nullptr;   // cluster evaluations, wave passes of the last pruned sweep
    size_t evalsCap = 0;
    int last_pruned_K = 0;        // K of the last pruned sweep, 0 if the last sweep was not pruned
    // dStats holds the statistics of the current labels and sub-labels, for Kpartition clusters
    bool stats_cached = false;
    int Kstats = 0;   // clusters whose statistics dStats holds (entries 0 .. 2 Kstats - 1)
    // After a relabel that only redraws the sub-labels of some clusters (a reset), the chunk
    // partials of every other cluster are still those of the current labels: the next statistics
    // pass recomputes the flagged clusters only.
    bool partials_valid = false;
    int partialsK = 0;
    bool dirty_pending = false;
    std::vector<int32_t> dirty_host;
    int32_t* dDirty = nullptr;
    size_t dirtyCap = 0;
    const int32_t* stats_mask = nullptr;   // dDirty during a partial pass, else null
    // prior and buffers of the batched conjugate work
    bool has_prior = false;
    PriorArgs prior_args{};
    double* dPrior = nullptr;   // m0 (D) | psi0 (D x D) or alpha0 (D)
    // the outlier cluster's prior (dpmm_set_outlier_prior), selected per job / part
    bool has_prior1 = false;
    PriorArgs prior_args1{};
    double* dPrior1 = nullptr;
    int32_t* dPsel = nullptr;     // per-job prior selectors of a log-ML batch
    size_t pselCap = 0;
    int32_t* dCtl = nullptr;      // per-part draw control of a parameter draw
    size_t ctlCap = 0;
    double* dDrawsPrev = nullptr; // the previous draw call's rows, for parts that keep their parameters
    size_t drawsPrevCap = 0;
    int32_t* dMembers = nullptr;
    size_t membersCap = 0;
    double* dLM = nullptr;
    size_t lmCap = 0;
    double* dDraws = nullptr;
    size_t drawsCap = 0;
    int32_t* dDrawStatus = nullptr;
    size_t drawStatusCap = 0;
    double* dLogW = nullptr;
    size_t logwCap = 0;
    float* dFeat = nullptr;
    size_t featCap = 0;
    // mixed multinomial sweep: split-BF16 tile block for clusters in NCB blocks of 32
    unsigned char* dMB = nullptr;
    size_t mbCap = 0;
    int mm_ncb = 0;        // 0: no tile block for the current parameters
    int max_count = -1;    // largest count in X (multinomial), -1 until measured
    int Kdraw = 0;
    // device-resident sampler iteration (dpmm_step.h): the cluster table, its history rings (step_B
    // doubles per cluster) and their compaction copies, sized for kMaxK clusters; the log-ML job list and
    // the accepted pairs grow with K
    StepCluster* dStepCl = nullptr;
    StepCluster* dStepClTmp = nullptr;
    double* dStepHist = nullptr;
    double* dStepHistTmp = nullptr;
    int step_B = 0;
    double* dStepW = nullptr;         // mixture weights of the last draw
    int32_t* dStepMem = nullptr;      // statistics entries of the draw's 3K parts
    int32_t* dStepCand = nullptr;     // merge candidates, then the touched clusters
    int32_t* dStepCounts = nullptr;   // jobs, candidates, pairs the merge bound kept, pairs it rejected
    int32_t* dStepMap = nullptr;      // log-ML row of each pair job the merge bound kept
    size_t stepMapCap = 0;
    double* dStepBnd = nullptr;       // the merge bound's per-cluster means (K x D) and scalars (K x 6)
    size_t stepBndCap = 0;
    int32_t* dStepTab = nullptr;      // relabel tables: resets + splits (3 kMaxK), merges + removals (3 kMaxK), aux (2 kMaxK)
    int32_t* dStepJobs = nullptr;
    size_t stepJobsCap = 0;
    int32_t* dStepAcc = nullptr;
    size_t stepAccCap = 0;
    StepSummary* dStepSum = nullptr;
    StepSummary* hStepSum = nullptr;  // pinned
    int drawStride() const {
        return prior == DPMM_PRIOR_GAUSSIAN ? ((D + D * (D + 1) / 2 + 1) + 1) & ~1 : (D + 1) & ~1;
    }
    hipEvent_t ev[8] = {};   // phases: label sweep 0-1, sub-labels 2-3, statistics 4-5; the sweep's main kernel 6-7
    float ms_label = 0, ms_sub = 0, ms_stats = 0;
    bool ev_label = false, ev_sub = false, ev_stats = false, ev_main = false;
    uint32_t key0() const { return (uint32_t)seed; }
    uint32_t key1() const { return (uint32_t)(seed >> 32); }
    size_t statEntries() const { return prior == DPMM_PRIOR_GAUSSIAN ? 1 + DP + (size_t)DP * DP : 1 + (size_t)D; }
};

namespace {

int pick_dp(int D) {
    static const int dps[] = {4, 8, 16, 32, 48, 64, 96, 128};
    for (int dp : dps)
        if (D <= dp) return dp;
    return -1;
}

// Set while dpmm_sweep_iteration records a sweep into a hipGraph: nothing may allocate, synchronise
// or record timing events meanwhile.
static thread_local bool t_capturing = false;

// Grows a device buffer to n elements.  The new buffer is allocated before the old one is released:
// when the allocation fails, the old buffer and its capacity stay, so no state of the context ever
// refers to a released or missing buffer.
template <typename T>
int grow(T*& p, size_t& cap, size_t n) {
    if (n <= cap && p) return DPMM_OK;
    if (t_capturing) return fail(DPMM_ESTATE, "buffer growth while a sweep is being captured");
    T* q = nullptr;
    if (dev_malloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return fail(DPMM_ENOMEM, "hipMalloc failed");
    if (p) (void)hipFree(p);
    p = q;
    cap = n;
    return DPMM_OK;
}

int ensure_partition(dpmm_ctx* c, int K) {
    c->nblk = (int)((c->N + kBlockPts - 1) / kBlockPts);
    int rc = grow(c->dHist, c->histCap, (size_t)K * c->nblk);
    if (rc) return rc;
    if (c->Kpart < K || !c->dLabelStart) {
        if (t_capturing) return fail(DPMM_ESTATE, "buffer growth while a sweep is being captured");
        // all three new buffers first; the old ones stay when any allocation fails
        const int cap = std::max(K, 64) + 1;
        int32_t* nb[3] = {nullptr, nullptr, nullptr};
        for (auto& q : nb)
            if (dev_malloc(&q, cap * sizeof(int32_t)) != hipSuccess) {
                for (auto& r : nb)
                    if (r) (void)hipFree(r);
                return fail(DPMM_ENOMEM, "hipMalloc failed");
            }
        for (int32_t* old : {c->dRowtot, c->dLabelStart, c->dChunkStart})
            if (old) (void)hipFree(old);
        c->dRowtot = nb[0];
        c->dLabelStart = nb[1];
        c->dChunkStart = nb[2];
        c->Kpart = cap - 1;
    }
    return DPMM_OK;
}

// Points per partition chunk, aiming at about 4096 chunks from 2^22 points on (2048 below, and for
// DP > 64): more, smaller workgroups shorten the tail of the per-chunk kernels, fewer keep the chunk
// reduction small.  Measured per sampler iteration: N = 1e7, D = 64: 4.03-4.09 ms with 4096 chunks,
// 4.13 with 2048, 4.33 with 8192; N = 1e6, D = 32: 0.576 ms with 2048, 0.597 with 4096; N = 5e6,
// D = 128: 39.0 ms with 2048, 39.6 with 1024.
int chunk_size(const dpmm_ctx* c) {
    const int64_t target = c->DP > 64 || c->N < (int64_t(1) << 22) ? 2048 : 4096;
    int64_t ch = (c->N + target - 1) / target;
    ch = std::max<int64_t>(ch, 256);
    ch = (ch + 63) / 64 * 64;
    return (int)ch;
}

int chunks_upper(const dpmm_ctx* c, int K) { return (int)((c->N + c->CH - 1) / c->CH) + K; }

int allow_lds(const void* kernel, size_t bytes);

// cond (null: always): the partition kernels run only when *cond != 0 on the device (see k_block_hist).
int build_partition(dpmm_ctx* c, int K, int clamp, const int* cond = nullptr) {
    int rc = ensure_partition(c, K);
    if (rc) return rc;
    c->partials_valid = false;   // the chunk layout may change
    c->CH = chunk_size(c);
    hipLaunchKernelGGL(k_block_hist, dim3(c->nblk), dim3(kThreads), 0, c->stream, c->dLabels, c->N, K, c->nblk, clamp,
                       c->dHist, cond);
    hipLaunchKernelGGL(k_scan_rows, dim3(K), dim3(kThreads), 0, c->stream, c->dHist, c->nblk, c->dRowtot, cond);
    hipLaunchKernelGGL(k_scan_labels, dim3(1), dim3(64), 0, c->stream, c->dRowtot, K, c->CH, c->dLabelStart,
                       c->dChunkStart, cond);
    const size_t scatter_lds = (((size_t)5 * K * 4 + 15) & ~(size_t)15) + (size_t)4 * K * 8;
    int lrc = allow_lds(reinterpret_cast<const void*>(k_scatter), scatter_lds);
    if (lrc) return lrc;
    hipLaunchKernelGGL(k_scatter, dim3(c->nblk), dim3(kThreads), scatter_lds, c->stream, c->dLabels, c->N, K, c->nblk,
                       clamp, c->dHist, c->dLabelStart, c->dPerm, cond);
    HIPCHK(hipGetLastError());
    c->partition_valid = true;
    c->Kpartition = K;
    c->stats_cached = false;
    c->partials_valid = false;
    c->subs_perm_valid = false;   // positions changed
    return DPMM_OK;
}

int create_common(const void* X, int64_t N, int D, int dtype, int prior, uint64_t seed, bool device_ptr,
                  dpmm_ctx** out) {
    if (!out) return fail(DPMM_EINVAL, "out is null");
    *out = nullptr;
    if (!X || N <= 0 || N >= (int64_t)INT32_MAX || D <= 0) return fail(DPMM_EINVAL, "bad X / N / D");
    if (prior != DPMM_PRIOR_GAUSSIAN && prior != DPMM_PRIOR_MULTINOMIAL) return fail(DPMM_EINVAL, "bad prior");
    if (prior == DPMM_PRIOR_GAUSSIAN && dtype != DPMM_DTYPE_F32 && dtype != DPMM_DTYPE_F64)
        return fail(DPMM_EINVAL, "Gaussian data must be f32 or f64");
    if (prior == DPMM_PRIOR_MULTINOMIAL && dtype != DPMM_DTYPE_I32)
        return fail(DPMM_EINVAL, "multinomial data must be int32 counts");
    if (prior == DPMM_PRIOR_GAUSSIAN && pick_dp(D) < 0) return fail(DPMM_EINVAL, "Gaussian D > 128 unsupported");
    if (prior == DPMM_PRIOR_GAUSSIAN && dtype == DPMM_DTYPE_F64 && D > 64)
        return fail(DPMM_EINVAL, "f64 Gaussian data supports D <= 64");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return fail(DPMM_EHIP, "no HIP device");
    dpmm_ctx* c = new dpmm_ctx();
    c->N = N; c->D = D; c->dtype = dtype; c->prior = prior; c->seed = seed;
    c->DP = prior == DPMM_PRIOR_GAUSSIAN ? pick_dp(D) : D;
    c->DPM = (D + 31) / 32 * 32;
    c->precision = dtype == DPMM_DTYPE_F64 ? DPMM_PREC_FP64 : DPMM_PREC_FP32;
    (void)hipGetDevice(&c->device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(DPMM_EHIP, "stream create failed");
    }
    const size_t esz = dtype == DPMM_DTYPE_F64 ? 8 : 4;
    const size_t bytes = (size_t)N * D * esz;
    if (device_ptr) {
        c->dX = const_cast<void*>(X);
    } else {
        if (dev_malloc(&c->dX, bytes) != hipSuccess) { dpmm_destroy(c); return fail(DPMM_ENOMEM, "X alloc"); }
        c->owns_X = true;
        if (hipMemcpy(c->dX, X, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            dpmm_destroy(c);
            return fail(DPMM_EHIP, "X copy");
        }
    }
    if (dev_malloc(&c->dLabels, N * sizeof(int32_t)) != hipSuccess ||
        dev_malloc(&c->dSubs, N * sizeof(int32_t)) != hipSuccess ||
        dev_malloc(&c->dPerm, N * sizeof(int32_t)) != hipSuccess) {
        dpmm_destroy(c);
        return fail(DPMM_ENOMEM, "label alloc");
    }
    (void)hipMemsetAsync(c->dLabels, 0, N * sizeof(int32_t), c->stream);
    (void)hipMemsetAsync(c->dSubs, 0, N * sizeof(int32_t), c->stream);
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    *out = c;
    return DPMM_OK;
}

// ---- parameter packing -----------------------------------------------------------------------
bool chol_lower(int D, const double* A, std::vector<double>& L) {
    L.assign(A, A + (size_t)D * D);
    for (int j = 0; j < D; ++j) {
        double s = L[(size_t)j * D + j];
        for (int k = 0; k < j; ++k) s -= L[(size_t)j * D + k] * L[(size_t)j * D + k];
        if (!(s > 0.0) || !std::isfinite(s)) return false;
        const double d = std::sqrt(s);
        L[(size_t)j * D + j] = d;
        for (int i = j + 1; i < D; ++i) {
            double t = L[(size_t)i * D + j];
            for (int k = 0; k < j; ++k) t -= L[(size_t)i * D + k] * L[(size_t)j * D + k];
            L[(size_t)i * D + j] = t / d;
        }
    }
    return true;
}

template <typename T>
void pack_gauss_slot(int D, int DP, const double* mu, const double* R, double c, T* dst) {
    const int stride = gauss_slot_stride(DP);
    std::fill(dst, dst + stride, T(0));
    for (int j = 0; j < D; ++j) {
        const int l0 = j & ~3;
        T* row = dst + tri_row_offset(DP, j) - l0;
        for (int l = j; l < D; ++l) row[l] = (T)R[(size_t)l * D + j];   // W = R^T
    }
    for (int l = 0; l < D; ++l) dst[gauss_mu_offset(DP) + l] = (T)mu[l];
    dst[gauss_c_offset(DP)] = (T)c;
}

int ensure_param_buffers(dpmm_ctx* c, int K, size_t slot_bytes) {
    if (K > kMaxK) return fail(DPMM_EINVAL, "K exceeds 2048");
    const size_t need = (size_t)K * slot_bytes;
    if (c->capBytes < need || !c->dP) {
        // both new buffers first; the old ones (and the parameters in them) stay when either fails
        const int cap = std::max(K, 64);
        void* np_ = nullptr;
        void* nsp = nullptr;
        if (dev_malloc(&np_, (size_t)cap * slot_bytes) != hipSuccess ||
            dev_malloc(&nsp, (size_t)2 * cap * slot_bytes) != hipSuccess) {
            if (np_) (void)hipFree(np_);
            return fail(DPMM_ENOMEM, "param alloc");
        }
        if (c->dP) (void)hipFree(c->dP);
        if (c->dSP) (void)hipFree(c->dSP);
        c->dP = np_;
        c->dSP = nsp;
        c->Kcap = cap;
        c->capBytes = (size_t)cap * slot_bytes;
    }
    return DPMM_OK;
}

uint16_t f32_to_bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// BF16 screen slot of one cluster: W = R^T tiles in MFMA lane order, mu, score constant.
void pack_mx_slot(int D, int DPM, const double* mu, const double* R, double cst, unsigned char* dst,
                  unsigned char* lo_dst) {
    std::memset(dst, 0, mx_slot_bytes(DPM));
    std::memset(lo_dst, 0, mx_tiles_bytes(DPM));
    uint16_t* tiles = reinterpret_cast<uint16_t*>(dst);
    uint16_t* lo = reinterpret_cast<uint16_t*>(lo_dst);
    for (int j = 0; j < D; ++j)
        for (int l = j; l < D; ++l) {
            const float w = (float)R[(size_t)l * D + j];
            const uint16_t hb = f32_to_bf16_rne(w);
            uint32_t hu = (uint32_t)hb << 16;
            float hf;
            std::memcpy(&hf, &hu, 4);
            tiles[mx_tile_index(DPM, j, l)] = hb;
            lo[mx_tile_index(DPM, j, l)] = f32_to_bf16_rne(w - hf);
        }
    float* m = reinterpret_cast<float*>(dst + mx_tiles_bytes(DPM));
    for (int l = 0; l < D; ++l) m[l] = (float)mu[l];
    m[DPM] = (float)cst;
    // F >= ||W||_F^2 of the FP32 W (rounded up)
    double F = 0.0;
    for (int j = 0; j < D; ++j)
        for (int l = j; l < D; ++l) {
            const double w = (double)(float)R[(size_t)l * D + j];
            F += w * w;
        }
    m[DPM + 1] = (float)(F * (1.0 + 0x1p-20));
}

// Recheck list of the mixed sweeps.
int ensure_recheck(dpmm_ctx* c) {
    if (!c->dRecheck || !c->dRecheckCount) {
        if (dev_malloc(&c->dRecheck, c->N * sizeof(int32_t)) != hipSuccess ||
            dev_malloc(&c->dRecheckCount, 2 * sizeof(int)) != hipSuccess) {
            if (c->dRecheck) (void)hipFree(c->dRecheck);
            if (c->dRecheckCount) (void)hipFree(c->dRecheckCount);
            c->dRecheck = nullptr;
            c->dRecheckCount = nullptr;
            return fail(DPMM_ENOMEM, "recheck alloc");
        }
    }
    return DPMM_OK;
}

// BF16 screen slots for K clusters and 2K sub-clusters, and the recheck list.
int ensure_mixed_buffers(dpmm_ctx* c, int K) {
    unsigned char* pb = reinterpret_cast<unsigned char*>(c->dPB);
    int rc = grow(pb, c->pbCap, (size_t)K * mx_slot_bytes(c->DPM));
    c->dPB = pb;
    if (rc) return rc;
    unsigned char* spb = reinterpret_cast<unsigned char*>(c->dSPB);
    rc = grow(spb, c->spbCap, (size_t)2 * K * mx2_slot_bytes(c->DPM));
    c->dSPB = spb;
    if (rc) return rc;
    unsigned char* plo = reinterpret_cast<unsigned char*>(c->dPBlo);
    rc = grow(plo, c->pbloCap, (size_t)K * mx_tiles_bytes(c->DPM));
    c->dPBlo = plo;
    if (rc) return rc;
    return ensure_recheck(c);
}

// Split-BF16 tiles of the K clusters' log theta (cluster k's D values at src + k * row_stride).
int pack_mult_mixed(dpmm_ctx* c, int K, const double* src, size_t row_stride, const double* logw, int use_weight) {
    c->mm_ncb = 0;
    if (K > 256) return DPMM_OK;   // the FP32 sweep serves larger K
    const int ncb = (K + 31) / 32;
    int rc = grow(c->dMB, c->mbCap, mm_block_bytes(c->D, ncb));
    if (rc) return rc;
    rc = ensure_recheck(c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mult_mx_clear, dim3(256), dim3(kThreads), 0, c->stream, c->dMB, c->D, ncb);
    hipLaunchKernelGGL(k_pack_mult_mx, dim3(K), dim3(kThreads), 0, c->stream, src, row_stride, c->D, ncb, logw,
                       use_weight, c->dMB);
    HIPCHK(hipGetLastError());
    c->mm_ncb = ncb;
    return DPMM_OK;
}

// Split-BF16 slot of one sub-cluster for the sub-label screen: hi and lo tiles of W = R^T (of the
// FP32 value the FP32 sweep uses), mu, score constant.
void pack_mx2_slot(int D, int DPM, const double* mu, const double* R, double cst, const double* parent_mu,
                   unsigned char* dst) {
    std::memset(dst, 0, mx2_slot_bytes(DPM));
    uint16_t* hi = reinterpret_cast<uint16_t*>(dst);
    uint16_t* lo = reinterpret_cast<uint16_t*>(dst + mx_tiles_bytes(DPM));
    for (int j = 0; j < D; ++j)
        for (int l = j; l < D; ++l) {
            const float w = (float)R[(size_t)l * D + j];
            const uint16_t h = f32_to_bf16_rne(w);
            uint32_t hb = (uint32_t)h << 16;
            float hf;
            std::memcpy(&hf, &hb, 4);
            hi[mx_tile_index(DPM, j, l)] = h;
            lo[mx_tile_index(DPM, j, l)] = f32_to_bf16_rne(w - hf);
        }
    float* m = reinterpret_cast<float*>(dst + 2 * mx_tiles_bytes(DPM));
    for (int l = 0; l < D; ++l) m[l] = (float)mu[l];
    m[DPM] = (float)cst;
    // the screen relative to the parent cluster's mean: -c and g (see mx2_slot_bytes)
    float* rel = reinterpret_cast<float*>(dst + mx2_rel_offset(DPM));
    const double kappa = (double)sub_rel_eps_g(DPM) / (double)sub_rel_eps_a(DPM);
    double G = 0.0, H = 0.0;
    for (int j = 0; j < D; ++j) {
        double c = 0.0, b = 0.0;
        for (int l = j; l < D; ++l) {
            const double w = (double)(float)R[(size_t)l * D + j];
            const double dl = (double)(float)mu[l] - (double)(float)parent_mu[l];
            c += w * dl;
            b += std::fabs(w) * std::fabs(dl);
            const uint32_t hb = (uint32_t)f32_to_bf16_rne((float)w) << 16;
            float hf;
            std::memcpy(&hf, &hb, 4);
            H += (double)hf * (double)hf;
        }
        rel[j] = (float)(-c);
        const float g = (float)(kappa * (b + std::fabs(c)) * (1.0 + 0x1p-20) + 0x1p-120);
        rel[DPM + j] = g;
        G += (double)g * (double)g;
    }
    // ||g|| and ||W_hi||_F, rounded up
    m[DPM + 1] = (float)(std::sqrt(G) * (1.0 + 0x1p-20));
    m[DPM + 2] = (float)(std::sqrt(H) * (1.0 + 0x1p-20));
}

int set_gauss(dpmm_ctx* c, int K, const double* mu, const double* mat, const double* logdet, const double* weights,
              const double* smu, const double* smat, const double* slogdet, const double* lrw, bool is_factor) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    if (c->prior != DPMM_PRIOR_GAUSSIAN) return fail(DPMM_EINVAL, "context is not Gaussian");
    if (K <= 0 || !mu || !mat || !logdet || !weights || !smu || !smat || !slogdet || !lrw)
        return fail(DPMM_EINVAL, "bad Gaussian params");
    const int D = c->D, DP = c->DP;
    for (size_t e = 0; e < (size_t)K * D; ++e)
        if (!std::isfinite(mu[e]) || !std::isfinite(smu[e]) || !std::isfinite(smu[(size_t)K * D + e]))
            return fail(DPMM_EINVAL, "a mean is NaN or infinite");
    for (int k = 0; k < K; ++k)
        if (!std::isfinite(logdet[k]) || !std::isfinite(slogdet[2 * k]) || !std::isfinite(slogdet[2 * k + 1]) ||
            !(weights[k] >= 0.0) || !(lrw[2 * k] >= 0.0) || !(lrw[2 * k + 1] >= 0.0) || std::isinf(weights[k]))
            return fail(DPMM_EINVAL, "a log determinant or weight is NaN, infinite or negative");
    const bool f64 = c->precision == DPMM_PREC_FP64;
    const int stride = gauss_slot_stride(DP);
    const size_t esz = f64 ? 8 : 4;
    c->params_f64 = f64;
    c->stride = stride;
    int rc = ensure_param_buffers(c, K, (size_t)stride * esz);
    if (rc) return rc;
    c->hP.resize((size_t)K * stride * esz);
    c->hSP.resize((size_t)2 * K * stride * esz);
    std::vector<double> R;
    auto one = [&](const double* m, const double* A, double ld, double logw, unsigned char* dst) -> bool {
        if (is_factor) R.assign(A, A + (size_t)D * D);
        else if (!chol_lower(D, A, R)) return false;
        // Score constant without the common -D^2 log(2 pi) / 2 of cudaKernel_gaussian.cu:75,108:
        // label and sub-label draws only compare scores of the same point.
        const double cst = -0.5 * ld + logw;
        if (f64) pack_gauss_slot<double>(D, DP, m, R.data(), cst, reinterpret_cast<double*>(dst));
        else pack_gauss_slot<float>(D, DP, m, R.data(), cst, reinterpret_cast<float*>(dst));
        return true;
    };
    const bool mixed = c->precision == DPMM_PREC_MIXED;
    const int mslot = mx_slot_bytes(c->DPM);
    c->spb_valid = false;
    c->fused_subs = false;
    if (mixed) {
        c->hPB.assign((size_t)K * mslot, 0);
        c->hPBlo.assign((size_t)K * mx_tiles_bytes(c->DPM), 0);
        c->hSPB.assign((size_t)2 * K * mx2_slot_bytes(c->DPM), 0);
    }
    for (int k = 0; k < K; ++k) {
        if (!std::isfinite(logdet[k])) return fail(DPMM_EINVAL, "non-finite logdet");
        if (!one(mu + (size_t)k * D, mat + (size_t)k * D * D, logdet[k], std::log(weights[k]),
                 c->hP.data() + (size_t)k * stride * esz))
            return fail(DPMM_ENOTSPD, "cluster precision matrix not SPD");
        if (mixed)
            pack_mx_slot(D, c->DPM, mu + (size_t)k * D, R.data(), -0.5 * logdet[k] + std::log(weights[k]),
                         c->hPB.data() + (size_t)k * mslot, c->hPBlo.data() + (size_t)k * mx_tiles_bytes(c->DPM));
    }
    if (mixed) {
        int rc2 = ensure_mixed_buffers(c, K);
        if (rc2) return rc2;
        HIPCHK(hipMemcpyAsync(c->dPB, c->hPB.data(), c->hPB.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->dPBlo, c->hPBlo.data(), c->hPBlo.size(), hipMemcpyHostToDevice, c->stream));
    }
    for (int q = 0; q < 2 * K; ++q) {
        if (!one(smu + (size_t)q * D, smat + (size_t)q * D * D, slogdet[q], std::log(lrw[q]),
                 c->hSP.data() + (size_t)q * stride * esz))
            return fail(DPMM_ENOTSPD, "sub-cluster precision matrix not SPD");
        if (mixed)
            pack_mx2_slot(D, c->DPM, smu + (size_t)q * D, R.data(), -0.5 * slogdet[q] + std::log(lrw[q]),
                          mu + (size_t)(q / 2) * D, c->hSPB.data() + (size_t)q * mx2_slot_bytes(c->DPM));
    }
    if (mixed) {
        HIPCHK(hipMemcpyAsync(c->dSPB, c->hSPB.data(), c->hSPB.size(), hipMemcpyHostToDevice, c->stream));
        c->spb_valid = true;
    }
    HIPCHK(hipMemcpyAsync(c->dP, c->hP.data(), c->hP.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dSP, c->hSP.data(), c->hSP.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->K = K;
    c->prune_valid = false;
    return DPMM_OK;
}

template <typename T, int DP>
void run_label_gauss(dpmm_ctx* c, uint32_t iter, int fin) {
    constexpr int ppb = label_points_per_block<T, DP>();
    const int grid = (int)((c->N + ppb - 1) / ppb);
    launch_label_gauss<T, DP>(c->stream, grid, reinterpret_cast<const T*>(c->dX), c->N, c->D, c->K,
                              reinterpret_cast<const T*>(c->dP), c->key0(), c->key1(), iter, fin, c->dLabels);
}

template <typename T, int DP>
void run_sub_gauss(dpmm_ctx* c, uint32_t iter) {
    launch_sub_gauss<T, DP>(c->stream, chunks_upper(c, c->K), reinterpret_cast<const T*>(c->dX), c->D, c->K, c->CH,
                            c->dPerm, c->dLabelStart, c->dChunkStart, reinterpret_cast<const T*>(c->dSP), c->key0(),
                            c->key1(), iter, c->dSubs);
}

template <typename TX, int DP>
void run_stats_gauss(dpmm_ctx* c) {
    const int K = c->Kpartition;
    launch_stats_gauss<TX, DP>(c->stream, chunks_upper(c, K), reinterpret_cast<const TX*>(c->dX), c->D, K, c->CH,
                               c->dPerm, c->dLabelStart, c->dChunkStart, c->dSubs,
                               c->subs_perm_valid ? c->dSubsPerm : nullptr, c->dPartials, c->stats_mask, c->stats_gate);
}

#define DPMM_DISPATCH_DP(DPV, FN, ...)            \
    switch (DPV) {                                \
        case 4: FN<4>(__VA_ARGS__); break;        \
        case 8: FN<8>(__VA_ARGS__); break;        \
        case 16: FN<16>(__VA_ARGS__); break;      \
        case 32: FN<32>(__VA_ARGS__); break;      \
        case 48: FN<48>(__VA_ARGS__); break;      \
        case 64: FN<64>(__VA_ARGS__); break;      \
        case 96: FN<96>(__VA_ARGS__); break;      \
        default: FN<128>(__VA_ARGS__); break;     \
    }

#define DPMM_DISPATCH_DP64(DPV, FN, ...)          \
    switch (DPV) {                                \
        case 4: FN<4>(__VA_ARGS__); break;        \
        case 8: FN<8>(__VA_ARGS__); break;        \
        case 16: FN<16>(__VA_ARGS__); break;      \
        case 32: FN<32>(__VA_ARGS__); break;      \
        case 48: FN<48>(__VA_ARGS__); break;      \
        default: FN<64>(__VA_ARGS__); break;      \
    }

template <int DP> void lg_f32(dpmm_ctx* c, uint32_t it, int f) { run_label_gauss<float, DP>(c, it, f); }
template <int DP> void lg_f64(dpmm_ctx* c, uint32_t it, int f) { run_label_gauss<double, DP>(c, it, f); }
template <int DP> void sg_f32(dpmm_ctx* c, uint32_t it) { run_sub_gauss<float, DP>(c, it); }
template <int DP> void sg_f64(dpmm_ctx* c, uint32_t it) { run_sub_gauss<double, DP>(c, it); }
template <int DP> void st_f32(dpmm_ctx* c) { run_stats_gauss<float, DP>(c); }
template <int DP> void st_f64(dpmm_ctx* c) { run_stats_gauss<double, DP>(c); }

template <int DP> void lgl_f32(dpmm_ctx* c, uint32_t it, int f, int grid) {
    launch_label_gauss_list<float, DP>(c->stream, grid, reinterpret_cast<const float*>(c->dX), c->N, c->D, c->K,
                                       reinterpret_cast<const float*>(c->dP), c->key0(), c->key1(), it, f,
                                       c->dRecheck, c->dRecheckCount, c->dLabels);
}

template <int DP> void lglp_f32(dpmm_ctx* c, uint32_t it, int f, int grid) {
    launch_label_gauss_listp<float, DP>(c->stream, grid, reinterpret_cast<const float*>(c->dX), c->N, c->D, c->K,
                                        reinterpret_cast<const float*>(c->dP), c->dPruneM, c->dPruneSC, c->key0(),
                                        c->key1(), it, f, c->dRecheck, c->dRecheckCount, c->dLabels, c->dFlags);
}

template <int DP> void sgl_f32(dpmm_ctx* c, uint32_t it, int grid) {
    launch_sub_gauss_list<float, DP>(c->stream, grid, reinterpret_cast<const float*>(c->dX), c->D, c->K,
                                     reinterpret_cast<const float*>(c->dSP), c->dLabels, c->dRecheck,
                                     c->dRecheckCount + 1, c->key0(), c->key1(), it, c->dSubs, c->dPerm,
                                     c->dSubsPerm);
}
// the lists of the fused sweep (per workgroup of its grid): positions in the partition it swept,
// sub-labels into dSubsOld
template <int DP> void sgsf_f32(dpmm_ctx* c, uint32_t it) {
    static_assert(2 * kMaxK <= kSegMaxLists, "list counts of the fused sweep");
    launch_sub_gauss_seg<float, DP>(c->stream, 1024, reinterpret_cast<const float*>(c->dX), c->D, c->K, c->dPerm,
                                    c->dLabelStart, reinterpret_cast<const float*>(c->dSP), c->dLabels,
                                    c->dSubList, c->dFlags + 1 + c->K, c->key0(), c->key1(), it, c->dSubs,
                                    c->dSubsOld, c->dRecheckCount + 1);
}
template <int DPM> void sgm(dpmm_ctx* c, uint32_t it) {
    launch_sub_gauss_mx<DPM>(c->stream, chunks_upper(c, c->K), reinterpret_cast<const float*>(c->dX), c->D, c->K,
                             c->CH, c->dPerm, c->dLabelStart, c->dChunkStart,
                             reinterpret_cast<const unsigned char*>(c->dSPB), c->tau, c->key0(), c->key1(), it,
                             c->dSubs, c->dSubsPerm, c->dRecheck, c->dRecheckCount + 1);
}

template <int DP> void score_f32(dpmm_ctx* c, int grid, int64_t first, int64_t n, int which, void* out) {
    launch_score_gauss<float, DP>(c->stream, grid, reinterpret_cast<const float*>(c->dX), c->D, c->K,
                                  reinterpret_cast<const float*>(c->dP), reinterpret_cast<const float*>(c->dSP),
                                  c->dLabels, first, n, which, reinterpret_cast<float*>(out));
}
template <int DP> void score_f64(dpmm_ctx* c, int grid, int64_t first, int64_t n, int which, void* out) {
    launch_score_gauss<double, DP>(c->stream, grid, reinterpret_cast<const double*>(c->dX), c->D, c->K,
                                   reinterpret_cast<const double*>(c->dP), reinterpret_cast<const double*>(c->dSP),
                                   c->dLabels, first, n, which, reinterpret_cast<double*>(out));
}

// BF16 sub-label screen over the segments, then the FP32 draw of the points it could not decide.
int run_sub_mixed(dpmm_ctx* c, uint32_t iter) {
    int rc = grow(c->dSubsPerm, c->subsPermCap, (size_t)c->N);
    if (rc) return rc;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->dRecheckCount + 1, 1);
    switch (c->DPM) {
        case 32: sgm<32>(c, iter); break;
        case 64: sgm<64>(c, iter); break;
        case 96: sgm<96>(c, iter); break;
        default: sgm<128>(c, iter); break;
    }
    const int grid = (int)std::min<int64_t>(1024, (c->N + kThreads - 1) / kThreads);
    DPMM_DISPATCH_DP(c->DP, sgl_f32, c, iter, grid);
    c->subs_perm_valid = true;   // every position of the partition got its sub-label in both orders
    return DPMM_OK;
}

// BF16 screen over all points, then the FP32 sweep over the points it could not decide.
int run_label_mixed(dpmm_ctx* c, uint32_t iter, int fin) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->dRecheckCount, 1);
    const float* X = reinterpret_cast<const float*>(c->dX);
    const unsigned char* PB = reinterpret_cast<const unsigned char*>(c->dPB);
    switch (c->DPM) {
        case 32: launch_label_gauss_mx<32>(c->stream, c->N, c->D, c->K, X, PB, c->tau, c->key0(), c->key1(), iter, fin,
                                           c->dLabels, c->dRecheck, c->dRecheckCount); break;
        case 64: launch_label_gauss_mx<64>(c->stream, c->N, c->D, c->K, X, PB, c->tau, c->key0(), c->key1(), iter, fin,
                                           c->dLabels, c->dRecheck, c->dRecheckCount); break;
        case 96: launch_label_gauss_mx<96>(c->stream, c->N, c->D, c->K, X, PB, c->tau, c->key0(), c->key1(), iter, fin,
                                           c->dLabels, c->dRecheck, c->dRecheckCount); break;
        case 128: launch_label_gauss_mx<128>(c->stream, c->N, c->D, c->K, X, PB, c->tau, c->key0(), c->key1(), iter,
                                             fin, c->dLabels, c->dRecheck, c->dRecheckCount); break;
        default: return fail(DPMM_EINVAL, "mixed sweep needs D <= 128");
    }
    const int grid = (int)std::min<int64_t>(1024, (c->N + kThreads - 1) / kThreads);
    DPMM_DISPATCH_DP(c->DP, lgl_f32, c, iter, fin, grid);
    return DPMM_OK;
}

int allow_lds(const void* kernel, size_t bytes);

// Pruning bounds of the current FP32 slots (once per parameter set).
int ensure_prune(dpmm_ctx* c) {
    if (c->prune_valid) return DPMM_OK;
    const int K = c->K, D = c->D;
    int rc = grow(c->dPruneM, c->pruneMCap, (size_t)K * K);
    if (rc) return rc;
    if ((rc = grow(c->dPruneSC, c->pruneSCCap, (size_t)2 * K))) return rc;
    if ((rc = grow(c->dPruneV, c->pruneVCap, (size_t)kMaxK * c->DP))) return rc;   // once, at full size
    const size_t lds = (((size_t)D * (D + 1) * 4 + 15) & ~(size_t)15) +
                       ((size_t)16 * std::max(D, 64) + 2 * (size_t)D + (size_t)D * (D + 1) / 2) * 8;
    if ((rc = allow_lds(reinterpret_cast<const void*>(k_prune_prep), lds))) return rc;
    hipLaunchKernelGGL(k_prune_prep, dim3(K), dim3(kThreads), lds, c->stream, reinterpret_cast<const float*>(c->dP),
                       K, D, c->DP, c->dPruneM, c->dPruneSC, c->dPruneV, std::min(c->prune_v_k, K));
    HIPCHK(hipGetLastError());
    c->prune_v_k = std::max(c->prune_v_k, K);
    c->prune_valid = true;
    return DPMM_OK;
}

// Workgroups per partition chunk of the pruned mixed sweep.
static int sweep_split(const dpmm_ctx* c, bool fuse) {
    return c->prune_split > 0 ? c->prune_split : fuse ? c->prune_split_fused : kPruneSplit;
}

// Mixed sweep with exact pruning: partition by the labels before the sweep (labels outside [0, K)
// filed under 0), the pruned BF16 screen over the partition's chunks, then the FP32 sweep over the
// points it could not decide.
int run_label_mixed_pruned(dpmm_ctx* c, uint32_t iter, int fin) {
    int rc = ensure_prune(c);
    if (rc) return rc;
    // the partition the last sub-label sweep used is the one of the current labels unless a relabel
    // came since (which clears partition_valid); that partition files every point under its label
    const bool reuse = c->partition_valid && c->Kpartition == c->K;
    if (!reuse && (rc = build_partition(c, c->K, 1))) return rc;
    if ((rc = grow(c->dEvals, c->evalsCap, 2))) return rc;
    if ((rc = grow(c->dFlags, c->flagsCap, (size_t)3 * kMaxK + 1))) return rc;
    // the sub-label screen rides along (its slots hold this iteration's sub-cluster parameters)
    const bool fuse = c->fuse_mode != 0 && c->DPM <= 64 && c->spb_valid;
    if (fuse) {
        if ((rc = grow(c->dSubsOld, c->subsOldCap, (size_t)c->N))) return rc;
        if ((rc = grow(c->dSubList, c->subListCap, (size_t)c->N))) return rc;
    }
    // dFlags[0]: a label moved; dFlags[1 + k]: a label moved into or out of cluster k; fused, then the
    // lengths of the 2K sub-label lists
    const int nflags = fuse ? 3 * c->K + 1 : c->K + 1;
    hipLaunchKernelGGL(k_zero_int, dim3((nflags + 63) / 64), dim3(64), 0, c->stream, c->dFlags, nflags);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->dRecheckCount, 2);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<int*>(c->dEvals), 4);
    const float* X = reinterpret_cast<const float*>(c->dX);
    const unsigned char* PB = reinterpret_cast<const unsigned char*>(c->dPB);
    MxpExtra sf{};
    sf.skip_same = reuse ? 1 : 0;
    sf.split = sweep_split(c, fuse);
    sf.PBlo = c->refine_mode && c->spb_valid ? reinterpret_cast<const unsigned char*>(c->dPBlo) : nullptr;
    const int grid = chunks_upper(c, c->K) * sf.split;
    if (fuse) {
        sf.SPB = reinterpret_cast<const unsigned char*>(c->dSPB);
        // the sub-labels of this partition, when the last pass left them in its order
        sf.subs_prev = c->subs_perm_valid ? c->dSubsPerm : nullptr;
        sf.subs = c->dSubs;
        sf.subs_old = c->dSubsOld;
        sf.sub_list = c->dSubList;
        sf.seg = c->dFlags + 1 + c->K;
        sf.sub_count = c->dRecheckCount + 1;
    }
#define DPMM_MXP(DPMV)                                                                                          \
    launch_label_gauss_mxp<DPMV>(c->stream, grid, X, c->D, c->K, c->CH, c->dPerm, c->dLabelStart, c->dChunkStart, \
                                 PB, c->dPruneM, c->dPruneSC, c->tau, c->key0(), c->key1(), iter, fin, c->dLabels, \
                                 c->dRecheck, c->dRecheckCount, c->dEvals, c->dFlags, sf)
    if (!t_capturing) (void)hipEventRecord(c->ev[6], c->stream);
    switch (c->DPM) {
        case 32: DPMM_MXP(32); break;
        case 64: DPMM_MXP(64); break;
        case 96: DPMM_MXP(96); break;
        case 128: DPMM_MXP(128); break;
        default: return fail(DPMM_EINVAL, "mixed sweep needs D <= 128");
    }
#undef DPMM_MXP
    if (!t_capturing) (void)hipEventRecord(c->ev[7], c->stream);
    c->ev_main = !t_capturing;
    const int grid2 = (int)std::min<int64_t>(1024, (c->N + kThreads - 1) / kThreads);
    DPMM_DISPATCH_DP(c->DP, lglp_f32, c, iter, fin, grid2);   // the recheck pruned with the same bounds
    // fused: the FP32 sub-label draw of the listed positions, by their final labels (after the recheck)
    if (fuse) DPMM_DISPATCH_DP(c->DP, sgsf_f32, c, iter);
    c->fused_subs = fuse;
    c->fused_iter = iter;
    c->last_pruned_K = c->K;
    c->label_flag = true;   // dFlags[0]: whether the partition of the labels before the sweep still stands
    return DPMM_OK;
}

template <int KPT>
void launch_label_mult_k(dpmm_ctx* c, uint32_t iter, int fin) {
    const int grid = (int)((c->N + 63) / 64);
    const size_t smem = (size_t)(32 * 65 + c->K * 33) * sizeof(float);
    hipLaunchKernelGGL((k_label_mult<KPT>), dim3(grid), dim3(kThreads), smem, c->stream,
                       reinterpret_cast<const int32_t*>(c->dX), c->N, c->D, c->K,
                       reinterpret_cast<const float*>(c->dP), c->stride, c->key0(), c->key1(), iter, fin,
                       c->dLabels);
}

template <int NCB>
void launch_label_mult_mx_k(dpmm_ctx* c, uint32_t iter, int fin) {
    const int grid = (int)((c->N + 127) / 128);
    const int vec = (c->D % 4 == 0 && reinterpret_cast<uintptr_t>(c->dX) % 16 == 0) ? 1 : 0;
    hipLaunchKernelGGL((k_label_mult_mx<NCB>), dim3(grid), dim3(kThreads), 0, c->stream,
                       reinterpret_cast<const int32_t*>(c->dX), c->N, c->D, vec, c->dMB, c->key0(), c->key1(), iter,
                       fin, c->dLabels, c->dRecheck, c->dRecheckCount);
}

// Split-BF16 MFMA screen over all documents, then the FP32 sweep over the ones it could not decide.
int run_label_mult_mixed(dpmm_ctx* c, uint32_t iter, int fin) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, c->dRecheckCount, 1);
    switch (c->mm_ncb) {
        case 1: launch_label_mult_mx_k<1>(c, iter, fin); break;
        case 2: launch_label_mult_mx_k<2>(c, iter, fin); break;
        case 3: launch_label_mult_mx_k<3>(c, iter, fin); break;
        case 4: launch_label_mult_mx_k<4>(c, iter, fin); break;
        case 5: launch_label_mult_mx_k<5>(c, iter, fin); break;
        case 6: launch_label_mult_mx_k<6>(c, iter, fin); break;
        case 7: launch_label_mult_mx_k<7>(c, iter, fin); break;
        default: launch_label_mult_mx_k<8>(c, iter, fin); break;
    }
    const int grid = (int)std::min<int64_t>(1024, (c->N + 3) / 4);
    hipLaunchKernelGGL(k_label_mult_list, dim3(grid), dim3(kThreads), 0, c->stream,
                       reinterpret_cast<const int32_t*>(c->dX), c->D, c->K, reinterpret_cast<const float*>(c->dP),
                       c->stride, c->key0(), c->key1(), iter, fin, c->dRecheck, c->dRecheckCount, c->dLabels);
    return DPMM_OK;
}

int launch_label_mult(dpmm_ctx* c, uint32_t iter, int fin) {
    if (c->K > kMaxKMult) return fail(DPMM_EINVAL, "multinomial K > 433");
    const int kpt = (c->K + 3) / 4;
    if (kpt <= 4) launch_label_mult_k<4>(c, iter, fin);
    else if (kpt <= 8) launch_label_mult_k<8>(c, iter, fin);
    else if (kpt <= 16) launch_label_mult_k<16>(c, iter, fin);
    else if (kpt <= 32) launch_label_mult_k<32>(c, iter, fin);
    else if (kpt <= 64) launch_label_mult_k<64>(c, iter, fin);
    else if (kpt <= 128) launch_label_mult_k<128>(c, iter, fin);
    else return fail(DPMM_EINVAL, "multinomial K > 512 unsupported");
    return DPMM_OK;
}

}  // namespace

// ============================================================================ C ABI ========
extern "C" {

const char* dpmm_last_error(void) { return g_err.c_str(); }

int dpmm_debug_fail_allocations(int64_t after) {
    g_alloc_budget.store(after < 0 ? -1 : after);
    return DPMM_OK;
}

int dpmm_device_count(int* count) {
    if (!count) return fail(DPMM_EINVAL, "null");
    HIPCHK(hipGetDeviceCount(count));
    return DPMM_OK;
}

int dpmm_create(const void* X, int64_t N, int D, int dtype, int prior, uint64_t seed, dpmm_ctx** out) {
    return create_common(X, N, D, dtype, prior, seed, false, out);
}

int dpmm_create_device(const void* dX, int64_t N, int D, int dtype, int prior, uint64_t seed, dpmm_ctx** out) {
    return create_common(dX, N, D, dtype, prior, seed, true, out);
}

int dpmm_destroy(dpmm_ctx* c) {
    if (!c) return DPMM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->owns_X && c->dX) (void)hipFree(c->dX);
    for (void* p : {(void*)c->dLabels, (void*)c->dSubs, (void*)c->dPerm, c->dP, c->dSP, (void*)c->dHist,
                    (void*)c->dRowtot, (void*)c->dLabelStart, (void*)c->dChunkStart, (void*)c->dTable,
                    (void*)c->dPartials, (void*)c->dStats, c->dPB, (void*)c->dRecheck, (void*)c->dRecheckCount,
                    (void*)c->dPrior, (void*)c->dMembers, (void*)c->dLM, (void*)c->dDraws, (void*)c->dDrawStatus,
                    (void*)c->dLogW, (void*)c->dFeat, (void*)c->dMB, c->dSPB, (void*)c->dPruneM,
                    (void*)c->dPruneSC, (void*)c->dEvals, (void*)c->dDirty, (void*)c->dClasses, (void*)c->dPruneV,
                    (void*)c->dPrior1, (void*)c->dPsel, (void*)c->dCtl, (void*)c->dDrawsPrev, (void*)c->dSubsPerm,
                    (void*)c->dStepCl, (void*)c->dStepClTmp, (void*)c->dStepHist, (void*)c->dStepHistTmp,
                    (void*)c->dStepW, (void*)c->dStepMem, (void*)c->dStepCand, (void*)c->dStepCounts,
                    (void*)c->dStepTab, (void*)c->dStepJobs, (void*)c->dStepAcc, (void*)c->dStepSum,
                    (void*)c->dFlags, (void*)c->dSubsPermPrev, (void*)c->dPrevStart, (void*)c->dSubsOld,
                    (void*)c->dSubList, c->dPBlo, (void*)c->dStepMap, (void*)c->dStepBnd})
        if (p) (void)hipFree(p);
    if (c->hStepSum) (void)hipHostFree(c->hStepSum);
    if (c->hStats) (void)hipHostFree(c->hStats);
    if (c->hEvals) (void)hipHostFree(c->hEvals);
    for (void* p : {c->pinMembers, c->pinLogW, c->pinOut})
        if (p) (void)hipHostFree(p);
    if (c->evEvals) (void)hipEventDestroy(c->evEvals);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DPMM_OK;
}

int dpmm_max_clusters(dpmm_ctx* c, int* kmax) {
    if (!c || !kmax) return fail(DPMM_EINVAL, "bad args");
    const int hw = c->prior == DPMM_PRIOR_GAUSSIAN ? kMaxK : kMaxKMult;
    *kmax = c->kcap_limit > 0 ? std::min(c->kcap_limit, hw) : hw;
    return DPMM_OK;
}

int dpmm_limit_clusters(dpmm_ctx* c, int kmax) {
    if (!c || kmax < 0) return fail(DPMM_EINVAL, "bad args");
    c->kcap_limit = kmax;
    return DPMM_OK;
}

int dpmm_set_precision(dpmm_ctx* c, int precision) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    if (precision == DPMM_PREC_FP64 && c->dtype != DPMM_DTYPE_F64)
        return fail(DPMM_EINVAL, "FP64 sweep needs f64 data");
    if (precision != DPMM_PREC_FP64 && c->dtype == DPMM_DTYPE_F64)
        return fail(DPMM_EINVAL, "f64 data sweeps in FP64");
    if (precision == DPMM_PREC_MIXED && c->dtype == DPMM_DTYPE_I32) {
        // counts enter the MFMA as BF16, exact up to 256
        if (c->max_count < 0) {
            int* dm = nullptr;
            if (dev_malloc(&dm, sizeof(int)) != hipSuccess) return fail(DPMM_ENOMEM, "max count");
            hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, dm, 1);
            hipLaunchKernelGGL(k_max_i32, dim3(1024), dim3(kThreads), 0, c->stream,
                               reinterpret_cast<const int32_t*>(c->dX), c->N * (int64_t)c->D, dm);
            int m = 0;
            const hipError_t e1 = hipMemcpyAsync(&m, dm, sizeof(int), hipMemcpyDeviceToHost, c->stream);
            const hipError_t e2 = hipStreamSynchronize(c->stream);
            (void)hipFree(dm);
            if (e1 != hipSuccess || e2 != hipSuccess) return fail(DPMM_EHIP, "max count copy");
            c->max_count = m;
        }
        if (c->max_count > 256) return fail(DPMM_EINVAL, "mixed multinomial sweep needs counts <= 256");
    }
    c->precision = precision;
    c->K = 0;
    return DPMM_OK;
}

int dpmm_sample_initial_labels(dpmm_ctx* c, int K0) {
    if (!c || K0 <= 0) return fail(DPMM_EINVAL, "bad args");
    const int grid = (int)((c->N + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_init_labels, dim3(grid), dim3(kThreads), 0, c->stream, c->dLabels, c->dSubs, c->N, K0,
                       c->key0(), c->key1());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->partition_valid = false;
    c->fused_subs = false;
    c->subs_perm_valid = false;
    c->stats_cached = false;
    c->inc_base = false;
    c->reset_update_ok = false;
    c->partials_valid = false;
    return DPMM_OK;
}

int dpmm_set_labels(dpmm_ctx* c, const int32_t* labels, const int32_t* subs) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    if (labels) HIPCHK(hipMemcpy(c->dLabels, labels, c->N * sizeof(int32_t), hipMemcpyHostToDevice));
    if (subs) HIPCHK(hipMemcpy(c->dSubs, subs, c->N * sizeof(int32_t), hipMemcpyHostToDevice));
    c->partition_valid = false;
    c->fused_subs = false;
    c->subs_perm_valid = false;
    c->stats_cached = false;
    c->inc_base = false;
    c->reset_update_ok = false;
    c->partials_valid = false;
    return DPMM_OK;
}

int dpmm_get_labels(dpmm_ctx* c, int32_t* labels, int32_t* subs) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (labels) HIPCHK(hipMemcpy(labels, c->dLabels, c->N * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (subs) HIPCHK(hipMemcpy(subs, c->dSubs, c->N * sizeof(int32_t), hipMemcpyDeviceToHost));
    return DPMM_OK;
}

int dpmm_set_gaussian_params(dpmm_ctx* c, int K, const double* mu, const double* inv_sigma,
                             const double* logdet_sigma, const double* weights, const double* sub_mu,
                             const double* sub_inv_sigma, const double* sub_logdet_sigma, const double* lr_weights) {
    return set_gauss(c, K, mu, inv_sigma, logdet_sigma, weights, sub_mu, sub_inv_sigma, sub_logdet_sigma,
                     lr_weights, false);
}

int dpmm_set_gaussian_factors(dpmm_ctx* c, int K, const double* mu, const double* chol, const double* logdet_sigma,
                              const double* weights, const double* sub_mu, const double* sub_chol,
                              const double* sub_logdet_sigma, const double* lr_weights) {
    return set_gauss(c, K, mu, chol, logdet_sigma, weights, sub_mu, sub_chol, sub_logdet_sigma, lr_weights, true);
}

int dpmm_set_multinomial_params(dpmm_ctx* c, int K, const double* logtheta, const double* weights,
                                const double* sub_logtheta, const double* lr_weights, int use_weight) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    if (c->prior != DPMM_PRIOR_MULTINOMIAL) return fail(DPMM_EINVAL, "context is not multinomial");
    if (K <= 0 || !logtheta || !sub_logtheta || !lr_weights || (use_weight && !weights))
        return fail(DPMM_EINVAL, "bad multinomial params");
    const int D = c->D;
    const int stride = (D + 1 + 3) & ~3;
    c->params_f64 = false;
    c->stride = stride;
    int rc = ensure_param_buffers(c, K, (size_t)stride * 4);
    if (rc) return rc;
    c->hP.assign((size_t)K * stride * 4, 0);
    c->hSP.assign((size_t)2 * K * stride * 4, 0);
    float* P = reinterpret_cast<float*>(c->hP.data());
    float* SP = reinterpret_cast<float*>(c->hSP.data());
    for (int k = 0; k < K; ++k) {
        for (int d = 0; d < D; ++d) P[(size_t)k * stride + d] = (float)logtheta[(size_t)k * D + d];
        P[(size_t)k * stride + D] = use_weight ? (float)std::log(weights[k]) : 0.f;
    }
    for (int q = 0; q < 2 * K; ++q) {
        for (int d = 0; d < D; ++d) SP[(size_t)q * stride + d] = (float)sub_logtheta[(size_t)q * D + d];
        SP[(size_t)q * stride + D] = (float)std::log(lr_weights[q]);
    }
    HIPCHK(hipMemcpyAsync(c->dP, c->hP.data(), c->hP.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dSP, c->hSP.data(), c->hSP.size(), hipMemcpyHostToDevice, c->stream));
    c->mm_ncb = 0;
    if (c->precision == DPMM_PREC_MIXED) {   // split-BF16 tiles from the same log theta (scratch: dLM)
        rc = grow(c->dLM, c->lmCap, (size_t)K * D + K);
        if (rc) return rc;
        std::vector<double> lw(K, 0.0);
        if (use_weight)
            for (int k = 0; k < K; ++k) lw[k] = std::log(weights[k]);
        HIPCHK(hipMemcpyAsync(c->dLM, logtheta, (size_t)K * D * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->dLM + (size_t)K * D, lw.data(), K * sizeof(double), hipMemcpyHostToDevice, c->stream));
        rc = pack_mult_mixed(c, K, c->dLM, (size_t)D, c->dLM + (size_t)K * D, use_weight);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->K = K;
    c->prune_valid = false;
    return DPMM_OK;
}

// The adaptive pruning's decision input: the (wave, cluster) scoring counters of the last pruned
// sweep, copied to pinned memory behind it.
static int resolve_evals(dpmm_ctx* c) {
    if (!c->evals_pending) return DPMM_OK;
    HIPCHK(hipEventSynchronize(c->evEvals));
    c->evals_pending = false;
    const double f = c->hEvals[1] ? (double)c->hEvals[0] / ((double)c->hEvals[1] * c->evals_K) : 0.0;
    if (f > 0.25) c->prune_skip = 4;
    return DPMM_OK;
}

static int copy_evals_back(dpmm_ctx* c) {
    if (!c->hEvals && hipHostMalloc((void**)&c->hEvals, 2 * sizeof(unsigned long long)) != hipSuccess)
        return fail(DPMM_ENOMEM, "pinned");
    if (!c->evEvals) HIPCHK(hipEventCreateWithFlags(&c->evEvals, hipEventDisableTiming));
    HIPCHK(hipMemcpyAsync(c->hEvals, c->dEvals, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipEventRecord(c->evEvals, c->stream));
    c->evals_pending = true;
    c->evals_copy_due = false;
    return DPMM_OK;
}

int dpmm_sample_cluster_labels(dpmm_ctx* c, uint32_t iter, int fin) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    if (c->K <= 0) return fail(DPMM_ESTATE, "parameters not set");
    if (!t_capturing) (void)hipEventRecord(c->ev[0], c->stream);
    c->last_pruned_K = 0;
    c->label_flag = false;
    c->fused_subs = false;
    c->ev_main = false;
    if (c->prior == DPMM_PRIOR_GAUSSIAN && c->precision == DPMM_PREC_MIXED) {
        bool prune = c->prune_mode != 0 && c->K >= 2 && c->D % 4 == 0;   // 16-byte rows for the gathers
        // adaptive: small sweeps (N K DP^2 below 4e11 multiply-adds) stay unpruned, the fixed cost
        // of the bounds and the partition would exceed what pruning saves
        if (prune && c->prune_mode == 1 && (double)c->N * c->K * c->DP * c->DP < 4e11) prune = false;
        if (prune && c->prune_mode == 1) {
            // adaptive: when the previous pruned sweep still scored more than a quarter of the
            // (wave, cluster) pairs, the next four sweeps run unpruned
            if (!t_capturing) {   // a capture resolves them before it starts
                int rc = resolve_evals(c);
                if (rc) return rc;
            }
            if (c->prune_skip > 0) {
                --c->prune_skip;
                prune = false;
            }
        }
        int rc = prune ? run_label_mixed_pruned(c, iter, fin) : run_label_mixed(c, iter, fin);
        if (rc) return rc;
        if (prune && c->prune_mode == 1) {
            c->evals_K = c->K;
            if (t_capturing) c->evals_copy_due = true;   // copied once the graph has been launched
            else if ((rc = copy_evals_back(c))) return rc;
        }
    } else if (c->prior == DPMM_PRIOR_GAUSSIAN) {
        if (c->params_f64) { DPMM_DISPATCH_DP64(c->DP, lg_f64, c, iter, fin); }
        else { DPMM_DISPATCH_DP(c->DP, lg_f32, c, iter, fin); }
    } else {
        int rc = (c->precision == DPMM_PREC_MIXED && c->mm_ncb > 0) ? run_label_mult_mixed(c, iter, fin)
                                                                     : launch_label_mult(c, iter, fin);
        if (rc) return rc;
    }
    if (!t_capturing) (void)hipEventRecord(c->ev[1], c->stream);
    HIPCHK(hipGetLastError());
    c->ev_label = !t_capturing;
    c->partition_valid = false;
    c->stats_cached = false;
    c->partials_valid = false;
    return DPMM_OK;
}

int dpmm_sample_sub_labels(dpmm_ctx* c, uint32_t iter) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    if (c->K <= 0) return fail(DPMM_ESTATE, "parameters not set");
    int rc;
    const bool mixed = c->prior == DPMM_PRIOR_GAUSSIAN && c->precision == DPMM_PREC_MIXED && c->spb_valid;
    // incremental statistics: dStats holds the statistics of the partition the label sweep started from
    // with the sub-labels in dSubsPerm; they become the baseline (dSubsPermPrev) of this pass, which updates
    // them from the points whose sub-label moved when no label moved (decided on the device)
    const bool baseline = mixed && c->label_flag && c->inc_base && c->subs_perm_valid && c->Kpartition == c->K &&
                          c->Kstats == c->K && !c->dirty_pending && c->stats_refresh > 1 &&
                          c->inc_passes + 1 < c->stats_refresh;
    c->inc_ready = false;
    c->inc_base = false;
    if (baseline) {
        if ((rc = grow(c->dSubsPermPrev, c->subsPermPrevCap, (size_t)c->N))) return rc;
        if ((rc = grow(c->dSubsPerm, c->subsPermCap, (size_t)c->N))) return rc;
        if ((rc = grow(c->dPrevStart, c->prevStartCap, (size_t)kMaxK + 1))) return rc;
        std::swap(c->dSubsPerm, c->dSubsPermPrev);
        std::swap(c->subsPermCap, c->subsPermPrevCap);
        c->inc_ready = true;
    }
    // fused: the sub-labels were drawn by the label sweep, in the order of the partition it swept
    const bool fused = mixed && c->fused_subs && c->fused_iter == iter && c->label_flag && c->Kpartition == c->K;
    if (baseline || fused) {
        if ((rc = grow(c->dPrevStart, c->prevStartCap, (size_t)kMaxK + 1))) return rc;
        // the segment starts of the partition swept by the label sweep (that of the baseline sub-labels)
        hipLaunchKernelGGL(k_copy_int, dim3((c->K + 64) / 64), dim3(64), 0, c->stream, c->dLabelStart, c->dPrevStart,
                           c->K + 1);
    }
    if (c->label_flag && c->Kpartition == c->K) {
        // the partition of the labels before the sweep, rebuilt on the device only if a label moved
        if ((rc = build_partition(c, c->K, 0, c->dFlags))) return rc;
    } else if (!c->partition_valid || c->Kpartition != c->K) {
        if ((rc = build_partition(c, c->K, 0))) return rc;
    }
    if (!t_capturing) (void)hipEventRecord(c->ev[2], c->stream);
    c->stats_cached = false;
    c->partials_valid = false;
    c->subs_perm_valid = false;
    if (fused) {
        if ((rc = grow(c->dSubsPerm, c->subsPermCap, (size_t)c->N))) return rc;
        hipLaunchKernelGGL(k_subs_to_partition, dim3(chunks_upper(c, c->K)), dim3(kThreads), 0, c->stream, c->K, c->CH,
                           c->dPerm, c->dLabelStart, c->dChunkStart, c->dPrevStart, c->dFlags + 1, c->dSubsOld,
                           c->dSubs, c->dSubsPerm);
        c->subs_perm_valid = true;
    } else if (mixed) {
        if ((rc = run_sub_mixed(c, iter))) return rc;
    } else if (c->prior == DPMM_PRIOR_GAUSSIAN) {
        if (c->params_f64) { DPMM_DISPATCH_DP64(c->DP, sg_f64, c, iter); }
        else { DPMM_DISPATCH_DP(c->DP, sg_f32, c, iter); }
    } else {
        hipLaunchKernelGGL(c->D % 4 == 0 ? k_sub_mult<4> : k_sub_mult<1>, dim3(chunks_upper(c, c->K)),
                           dim3(kThreads), (size_t)2 * c->stride * 4, c->stream, reinterpret_cast<const int32_t*>(c->dX), c->D, c->K, c->CH, c->dPerm,
                           c->dLabelStart, c->dChunkStart, reinterpret_cast<const float*>(c->dSP), c->stride,
                           c->key0(), c->key1(), iter, c->dSubs);
    }
    if (!t_capturing) (void)hipEventRecord(c->ev[3], c->stream);
    HIPCHK(hipGetLastError());
    c->ev_sub = !t_capturing;
    return DPMM_OK;
}

int dpmm_log_likelihoods(dpmm_ctx* c, int K, int which, int64_t first, int64_t n, double* out) {
    if (!c || !out || (which != 0 && which != 1)) return fail(DPMM_EINVAL, "bad args");
    if (c->K <= 0) return fail(DPMM_ESTATE, "parameters not set");
    if (K != c->K) return fail(DPMM_EINVAL, "K differs from the cluster count of the parameters set");
    if (first < 0 || n <= 0 || first + n > c->N) return fail(DPMM_EINVAL, "point range outside [0, N)");
    const int cols = which == 0 ? c->K : 2;
    if (which == 1) {   // the sub-cluster slots are indexed by each point's label: check them first
        std::vector<int32_t> lab((size_t)n);
        HIPCHK(hipMemcpyAsync(lab.data(), c->dLabels + first, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int32_t l : lab)
            if (l < 0 || l >= c->K) return fail(DPMM_ESTATE, "a label is outside the clusters whose parameters are set");
    }
    const bool f64 = c->prior == DPMM_PRIOR_GAUSSIAN && c->params_f64;
    const size_t esz = f64 ? 8 : 4;
    void* dOut = nullptr;
    HIPCHK(dev_malloc(&dOut, (size_t)n * cols * esz));
    if (c->prior == DPMM_PRIOR_GAUSSIAN) {
        const int64_t ppb = (c->DP * (int)esz > 256) ? kThreads / 2 : kThreads;
        const int grid = (int)std::min<int64_t>(4096, (n + ppb - 1) / ppb);
        if (f64) { DPMM_DISPATCH_DP64(c->DP, score_f64, c, grid, first, n, which, dOut); }
        else { DPMM_DISPATCH_DP(c->DP, score_f32, c, grid, first, n, which, dOut); }
    } else {
        const int grid = (int)std::min<int64_t>(4096, (n + 3) / 4);
        hipLaunchKernelGGL(c->D % 4 == 0 ? k_score_mult<4> : k_score_mult<1>, dim3(grid), dim3(kThreads), 0,
                           c->stream, reinterpret_cast<const int32_t*>(c->dX), c->D, c->K,
                           reinterpret_cast<const float*>(c->dP), reinterpret_cast<const float*>(c->dSP), c->stride,
                           c->dLabels, first, n, which, reinterpret_cast<float*>(dOut));
    }
    hipError_t e = hipGetLastError();
    std::vector<unsigned char> h((size_t)n * cols * esz);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dOut, h.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dOut);
    if (e != hipSuccess) return fail(DPMM_EHIP, std::string("log likelihoods: ") + hipGetErrorString(e));
    const size_t m = (size_t)n * cols;
    if (f64) std::memcpy(out, h.data(), m * 8);
    else for (size_t j = 0; j < m; ++j) out[j] = (double)reinterpret_cast<const float*>(h.data())[j];
    if (c->prior == DPMM_PRIOR_GAUSSIAN) {   // the constant the slots leave out (common to every score)
        const double common = -0.5 * (double)c->D * c->D * std::log(2.0 * kPiF64);
        for (size_t j = 0; j < m; ++j) out[j] += common;
    }
    return DPMM_OK;
}

// Statistics kernels for K clusters (partition first if needed), nothing read back.  dev_mask (K ints
// on the device, or null): after a relabel that only redrew the sub-labels of the clusters flagged
// nonzero, the pass recomputes those clusters only when the chunk partials of the others are current.
static int enqueue_stats(dpmm_ctx* c, int K, const int32_t* dev_mask = nullptr) {
    if (!c->partition_valid || c->Kpartition != K) {
        int rc = build_partition(c, K, 0);
        if (rc) return rc;
    }
    const size_t E = c->statEntries();
    const int nch = chunks_upper(c, K);
    if (!(c->stats_cached && c->Kpartition == K)) {
        const double* partials_before = c->dPartials;
        int rc = grow(c->dPartials, c->partialsCap, (size_t)2 * nch * E);
        if (c->dPartials != partials_before) c->partials_valid = false;   // a new buffer holds no partials
        if (rc) return rc;
        const double* stats_before = c->dStats;
        rc = grow(c->dStats, c->statsCap, (size_t)2 * K * E);
        if (c->dStats != stats_before) c->Kstats = 0;   // nor statistics
        if (rc) return rc;
        if (c->reset_update_ok && (dev_mask || c->dirty_pending) && c->Kstats == K && c->partition_valid &&
            c->Kpartition == K && c->prior == DPMM_PRIOR_GAUSSIAN && !t_capturing &&
            (dev_mask || (int)c->dirty_host.size() == K)) {
            // resets only: the reset clusters' statistics updated from their points whose sub-label moved
            const int32_t* mask = dev_mask;
            if (!mask) {
                if ((rc = grow(c->dDirty, c->dirtyCap, (size_t)K))) return rc;
                HIPCHK(hipMemcpyAsync(c->dDirty, c->dirty_host.data(), (size_t)K * 4, hipMemcpyHostToDevice, c->stream));
                mask = c->dDirty;
            }
            if ((rc = grow(c->dSubsPermPrev, c->subsPermPrevCap, (size_t)c->N))) return rc;
            if ((rc = grow(c->dPrevStart, c->prevStartCap, (size_t)kMaxK + 1))) return rc;
            if ((rc = grow(c->dFlags, c->flagsCap, (size_t)3 * kMaxK + 1))) return rc;
            hipLaunchKernelGGL(k_refresh_subs_perm, dim3(nch), dim3(kThreads), 0, c->stream, K, c->CH, c->dPerm,
                               c->dLabelStart, c->dChunkStart, c->dSubs, mask, c->dSubsPerm, c->dSubsPermPrev);
            hipLaunchKernelGGL(k_copy_int, dim3((K + 64) / 64), dim3(64), 0, c->stream, c->dLabelStart, c->dPrevStart,
                               K + 1);
            hipLaunchKernelGGL(k_zero_int, dim3((K + 63) / 64), dim3(64), 0, c->stream, c->dFlags + 1, K);
            c->subs_perm_valid = true;
            c->stats_mask = mask;
            c->stats_gate = StatsGate{c->dSubsPermPrev, c->dFlags + 1, c->dPrevStart};
            if (c->dtype == DPMM_DTYPE_F64) { DPMM_DISPATCH_DP64(c->DP, st_f64, c); }
            else { DPMM_DISPATCH_DP(c->DP, st_f32, c); }
            hipLaunchKernelGGL(k_stats_update, dim3(K, (unsigned)((E + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                               c->stream, c->dPartials, c->dChunkStart, K, (int)E, c->dStats, c->dFlags + 1,
                               (c->DP % 16 == 0 && c->DP <= 64) ? 1 : 0, mask);
            c->stats_gate = StatsGate{nullptr, nullptr, nullptr};
            c->stats_mask = nullptr;
            c->stats_cached = true;
            c->partials_valid = false;
            c->dirty_pending = false;
            c->reset_update_ok = false;
            c->inc_base = true;
            c->last_inc = false;
            HIPCHK(hipGetLastError());
            return DPMM_OK;
        }
        c->reset_update_ok = false;
        const bool partial_dev = dev_mask && c->partials_valid && c->partialsK == K && c->Kstats == K && !t_capturing;
        const bool partial = partial_dev || (c->dirty_pending && c->partials_valid && c->partialsK == K &&
                                             c->Kstats == K && (int)c->dirty_host.size() == K && !t_capturing);
        c->stats_mask = nullptr;
        if (partial_dev) {
            c->stats_mask = dev_mask;
        } else if (partial) {
            if ((rc = grow(c->dDirty, c->dirtyCap, (size_t)K))) return rc;
            HIPCHK(hipMemcpyAsync(c->dDirty, c->dirty_host.data(), (size_t)K * 4, hipMemcpyHostToDevice, c->stream));
            c->stats_mask = c->dDirty;
        } else if (!t_capturing) {
            (void)hipEventRecord(c->ev[4], c->stream);   // the phase timing covers full passes only
        }
        // update mode: the clusters whose points kept their labels (per the label sweep's flags, on the
        // device) are updated from the points whose sub-label moved, the others recomputed
        const bool inc = c->inc_ready && !partial && c->Kstats == K && c->subs_perm_valid &&
                         c->prior == DPMM_PRIOR_GAUSSIAN;
        const unsigned ey = (unsigned)((E + kThreads - 1) / kThreads);
        c->stats_gate = inc ? StatsGate{c->dSubsPermPrev, c->dFlags + 1, c->dPrevStart} : StatsGate{nullptr, nullptr, nullptr};
        if (c->prior == DPMM_PRIOR_GAUSSIAN) {
            if (c->dtype == DPMM_DTYPE_F64) { DPMM_DISPATCH_DP64(c->DP, st_f64, c); }
            else { DPMM_DISPATCH_DP(c->DP, st_f32, c); }
        } else {
            hipLaunchKernelGGL(c->D % 4 == 0 ? k_stats_mult<4> : k_stats_mult<1>, dim3(nch), dim3(kThreads), 0,
                               c->stream,
                               reinterpret_cast<const int32_t*>(c->dX), c->D, K, c->CH, c->dPerm, c->dLabelStart,
                               c->dChunkStart, c->dSubs, (int)E, c->dPartials, c->stats_mask);
        }
        if (inc)   // the wave kernel (DP = 16 .. 64) writes the signed change in slot 0
            hipLaunchKernelGGL(k_stats_update, dim3(K, ey), dim3(kThreads), 0, c->stream, c->dPartials, c->dChunkStart, K,
                               (int)E, c->dStats, c->dFlags + 1, (c->DP % 16 == 0 && c->DP <= 64) ? 1 : 0,
                               nullptr);
        else
            hipLaunchKernelGGL(k_stats_reduce, dim3(2 * K, ey), dim3(kThreads), 0, c->stream, c->dPartials,
                               c->dChunkStart, K, (int)E, c->dStats, c->stats_mask);
        c->stats_gate = StatsGate{nullptr, nullptr, nullptr};
        if (!partial && !t_capturing) {
            (void)hipEventRecord(c->ev[5], c->stream);
            c->ev_stats = true;
            c->timed_inc = inc;
        }
        c->stats_mask = nullptr;
        c->stats_cached = true;
        c->Kstats = K;
        c->partials_valid = !inc;   // after an update from the moved points the partials are differences
        c->partialsK = K;
        c->dirty_pending = false;
        c->inc_passes = inc ? c->inc_passes + 1 : 0;
        c->last_inc = inc;
        c->inc_ready = false;
        // the baseline of the next pass: these statistics with the sub-labels in partition order
        c->inc_base = !partial && c->subs_perm_valid && c->prior == DPMM_PRIOR_GAUSSIAN;
    }
    HIPCHK(hipGetLastError());
    return DPMM_OK;
}

int dpmm_sufficient_statistics(dpmm_ctx* c, int K, int64_t* Nsub, double* sum, double* S) {
    if (!c || !Nsub) return fail(DPMM_EINVAL, "bad args");
    if (K <= 0 || K > kMaxK) return fail(DPMM_EINVAL, "bad K");
    int rc = enqueue_stats(c, K);
    if (rc) return rc;
    const size_t E = c->statEntries();
    const size_t n = sum ? (size_t)2 * K * E : (size_t)2 * K;
    if (c->hStatsCap < n) {
        if (c->hStats) (void)hipHostFree(c->hStats);
        c->hStats = nullptr;
        if (hipHostMalloc((void**)&c->hStats, n * sizeof(double)) != hipSuccess) return fail(DPMM_ENOMEM, "pinned");
        c->hStatsCap = n;
    }
    if (!sum) {   // counts only: the first double of every entry
        HIPCHK(hipMemcpy2DAsync(c->hStats, sizeof(double), c->dStats, E * sizeof(double), sizeof(double),
                                (size_t)2 * K, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int q = 0; q < 2 * K; ++q) Nsub[q] = (int64_t)c->hStats[q];
        return DPMM_OK;
    }
    HIPCHK(hipMemcpyAsync(c->hStats, c->dStats, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int D = c->D, DP = c->DP;
    for (int q = 0; q < 2 * K; ++q) {
        const double* src = c->hStats + (size_t)q * E;
        Nsub[q] = (int64_t)src[0];
        for (int d = 0; d < D; ++d) sum[(size_t)q * D + d] = src[1 + d];
        if (S && c->prior == DPMM_PRIOR_GAUSSIAN)
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) S[((size_t)q * D + a) * D + b] = src[1 + DP + (size_t)a * DP + b];
    }
    return DPMM_OK;
}

// Every buffer the sweep segment (label sweep, sub-label sweep, statistics) can grow for the
// current K, grown before a capture so that nothing allocates while it records.
static int ensure_sweep_buffers(dpmm_ctx* c) {
    const int K = c->K;
    int rc = ensure_partition(c, K);
    if (rc) return rc;
    c->CH = chunk_size(c);
    const size_t E = c->statEntries();
    if ((rc = grow(c->dPartials, c->partialsCap, (size_t)2 * chunks_upper(c, K) * E))) return rc;
    if ((rc = grow(c->dStats, c->statsCap, (size_t)2 * K * E))) return rc;
    if (c->prior == DPMM_PRIOR_GAUSSIAN && c->precision == DPMM_PREC_MIXED) {
        if ((rc = grow(c->dPruneM, c->pruneMCap, (size_t)K * K))) return rc;
        if ((rc = grow(c->dPruneSC, c->pruneSCCap, (size_t)2 * K))) return rc;
        if ((rc = grow(c->dPruneV, c->pruneVCap, (size_t)kMaxK * c->DP))) return rc;
        if ((rc = grow(c->dEvals, c->evalsCap, 2))) return rc;
        if ((rc = grow(c->dSubsPerm, c->subsPermCap, (size_t)c->N))) return rc;
        if ((rc = grow(c->dSubsPermPrev, c->subsPermPrevCap, (size_t)c->N))) return rc;
        if ((rc = grow(c->dFlags, c->flagsCap, (size_t)3 * kMaxK + 1))) return rc;
        if ((rc = grow(c->dPrevStart, c->prevStartCap, (size_t)kMaxK + 1))) return rc;
        if (c->fuse_mode != 0 && c->DPM <= 64) {
            if ((rc = grow(c->dSubsOld, c->subsOldCap, (size_t)c->N))) return rc;
            if ((rc = grow(c->dSubList, c->subListCap, (size_t)c->N))) return rc;
        }
    }
    return DPMM_OK;
}

// Records what `record` enqueues into a hipGraph and launches it through the context's executable
// graph, updated in place while the recorded topology stays the same.  Nothing may allocate,
// synchronise or record timing events meanwhile.
static int run_captured(dpmm_ctx* c, const std::function<int()>& record) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    t_capturing = true;
    int rc = record();
    t_capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &g);
    // Nothing of the recorded work runs when capture or instantiation fails: every flag that the
    // recording set on the promise of that work (partition, statistics, pruning bounds, the pruning
    // counters' copy) is cleared again.
    auto drop_recorded = [c]() {
        c->partition_valid = false;
        c->subs_perm_valid = false;
        c->stats_cached = false;
        c->partials_valid = false;
        c->prune_valid = false;
        c->evals_copy_due = false;
        c->last_pruned_K = 0;
        c->label_flag = false;
        c->fused_subs = false;
        c->inc_base = false;
        c->inc_ready = false;
        c->reset_update_ok = false;
    };
    if (rc || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        drop_recorded();
        return rc ? rc : fail(DPMM_EHIP, std::string("graph capture: ") + hipGetErrorString(ec));
    }
    bool updated = false;
    if (c->gexec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(c->gexec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            updated = true;
        } else {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(c->gexec);
            c->gexec = nullptr;
        }
    }
    hipError_t e = hipSuccess;
    if (!updated) e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->gexec = nullptr;
        drop_recorded();
        return fail(DPMM_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    if (updated) ++c->graph_updates;
    else ++c->graph_instantiations;
    HIPCHK(hipGraphLaunch(c->gexec, c->stream));
    if (c->evals_copy_due && (rc = copy_evals_back(c))) return rc;
    return DPMM_OK;
}

int dpmm_set_graph_mode(dpmm_ctx* c, int enable) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    c->graph_mode = enable != 0;
    return DPMM_OK;
}

int dpmm_graph_counts(dpmm_ctx* c, int64_t* instantiations, int64_t* updates) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    if (instantiations) *instantiations = c->graph_instantiations;
    if (updates) *updates = c->graph_updates;
    return DPMM_OK;
}

int dpmm_sweep_iteration(dpmm_ctx* c, uint32_t iter, int final_iter, int64_t* Nsub) {
    if (!c || !Nsub) return fail(DPMM_EINVAL, "bad args");
    if (c->K <= 0) return fail(DPMM_ESTATE, "parameters not set");
    int rc;
    if (!c->graph_mode) {
        if ((rc = dpmm_sample_cluster_labels(c, iter, final_iter))) return rc;
        if ((rc = dpmm_sample_sub_labels(c, iter))) return rc;
        return dpmm_sufficient_statistics(c, c->K, Nsub, nullptr, nullptr);
    }
    if ((rc = resolve_evals(c))) return rc;   // the adaptive pruning decides from these, on the host
    if ((rc = ensure_sweep_buffers(c))) return rc;
    rc = run_captured(c, [&]() {
        int r = dpmm_sample_cluster_labels(c, iter, final_iter);
        if (!r) r = dpmm_sample_sub_labels(c, iter);
        if (!r) r = enqueue_stats(c, c->K);
        return r;
    });
    if (rc) return rc;
    return dpmm_sufficient_statistics(c, c->K, Nsub, nullptr, nullptr);   // the counts (statistics are current)
}

int dpmm_set_ground_truth(dpmm_ctx* c, const int32_t* classes, int num_classes) {
    if (!c || !classes || num_classes <= 0) return fail(DPMM_EINVAL, "bad args");
    for (int64_t i = 0; i < c->N; ++i)
        if (classes[i] < 0 || classes[i] >= num_classes) return fail(DPMM_EINVAL, "a class outside [0, num_classes)");
    if (!c->dClasses && dev_malloc(&c->dClasses, c->N * sizeof(int32_t)) != hipSuccess)
        return fail(DPMM_ENOMEM, "hipMalloc failed");
    HIPCHK(hipMemcpyAsync(c->dClasses, classes, c->N * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->nClasses = num_classes;
    return DPMM_OK;
}

int dpmm_label_contingency(dpmm_ctx* c, int K, int64_t* table) {
    if (!c || !table || K <= 0) return fail(DPMM_EINVAL, "bad args");
    if (!c->dClasses) return fail(DPMM_ESTATE, "no ground truth set");
    const int G = c->nClasses, KG = K * G;
    if ((int64_t)K * G > 16384) return fail(DPMM_EINVAL, "clusters x classes above 16384");
    const int rows = (int)std::min<int64_t>(256, (c->N + kThreads - 1) / kThreads);
    int32_t* dPart = nullptr;
    int64_t* dTab = nullptr;
    if (dev_malloc(&dPart, (size_t)rows * KG * sizeof(int32_t)) != hipSuccess ||
        dev_malloc(&dTab, (size_t)KG * sizeof(int64_t)) != hipSuccess) {
        if (dPart) (void)hipFree(dPart);
        return fail(DPMM_ENOMEM, "hipMalloc failed");
    }
    hipLaunchKernelGGL(k_contingency, dim3(rows), dim3(kThreads), (size_t)KG * sizeof(int), c->stream, c->dLabels,
                       c->dClasses, c->N, K, G, dPart);
    hipLaunchKernelGGL(k_table_reduce, dim3((KG + kThreads - 1) / kThreads), dim3(kThreads), 0, c->stream, dPart, rows,
                       KG, dTab);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(table, dTab, (size_t)KG * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dPart);
    (void)hipFree(dTab);
    if (e != hipSuccess) return fail(DPMM_EHIP, std::string("contingency: ") + hipGetErrorString(e));
    return DPMM_OK;
}

// ---- the allocation-free contingency table -------------------------------------------------------
// Scratch of dpmm_label_contingency_static: the per-workgroup partial tables of k_contingency, in a
// file-scope device array (4 MiB per GPU, held until the process ends).  Host code never passes the
// array's name to a kernel: on each device, on the first call made with it current, its address is
// taken once with hipGetSymbolAddress and kept per device.
constexpr size_t kStaticScratchInts = (size_t)1 << 20;
static __device__ __align__(256) int32_t scratch[kStaticScratchInts];
constexpr int kMaxDevices = 64;
static int32_t* g_scratch_ptr[kMaxDevices];
static hipError_t g_scratch_err[kMaxDevices];
static std::once_flag g_scratch_once[kMaxDevices];

static int scratch_of_current_device(int32_t** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(DPMM_EINVAL, "device index above 63");
    std::call_once(g_scratch_once[dev], [dev]() {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = static_cast<int32_t*>(p);
    });
    if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
        return fail(DPMM_EHIP, std::string("scratch address: ") + hipGetErrorString(g_scratch_err[dev]));
    *out = g_scratch_ptr[dev];
    return DPMM_OK;
}

int dpmm_label_contingency_static(const int32_t* d_labels, const int32_t* d_classes, int64_t N, int K, int C,
                                  int64_t* d_table, void* stream) {
    if (!d_labels || !d_classes || !d_table || N <= 0 || K <= 0 || C <= 0)
        return fail(DPMM_EINVAL, "bad args");
    if ((int64_t)K * C > 16384) return fail(DPMM_EINVAL, "clusters x classes above 16384");
    int32_t* part = nullptr;
    int rc = scratch_of_current_device(&part);
    if (rc) return rc;
    const int KG = K * C;
    // every partial row is written in full, so nothing left by an earlier call is read
    const int rows = (int)std::min<int64_t>({256, (N + kThreads - 1) / kThreads, (int64_t)(kStaticScratchInts / KG)});
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_contingency, dim3(rows), dim3(kThreads), (size_t)KG * sizeof(int), st, d_labels, d_classes, N,
                       K, C, part);
    hipLaunchKernelGGL(k_table_reduce, dim3((KG + kThreads - 1) / kThreads), dim3(kThreads), 0, st, part, rows, KG,
                       d_table);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(DPMM_EHIP, std::string("contingency: ") + hipGetErrorString(e));
    return DPMM_OK;
}

int dpmm_label_nmi(dpmm_ctx* c, int K, double* nmi) {
    if (!c || !nmi) return fail(DPMM_EINVAL, "bad args");
    if (!c->dClasses) return fail(DPMM_ESTATE, "no ground truth set");
    const int G = c->nClasses;
    std::vector<int64_t> t((size_t)K * G);
    int rc = dpmm_label_contingency(c, K, t.data());
    if (rc) return rc;
    // 2 I(X;Y) / (H(X) + H(Y)), natural logarithms, over the points whose label lies in [0, K)
    std::vector<double> pr(K, 0.0), pc(G, 0.0);
    double n = 0.0;
    for (int k = 0; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            const double v = (double)t[(size_t)k * G + g];
            pr[k] += v;
            pc[g] += v;
            n += v;
        }
    if (n <= 0.0) return fail(DPMM_ESTATE, "no labelled points");
    double I = 0.0, hr = 0.0, hc = 0.0;
    for (int k = 0; k < K; ++k)
        for (int g = 0; g < G; ++g) {
            const double v = (double)t[(size_t)k * G + g];
            if (v > 0.0) I += v / n * std::log(v * n / (pr[k] * pc[g]));
        }
    for (int k = 0; k < K; ++k)
        if (pr[k] > 0.0) hr -= pr[k] / n * std::log(pr[k] / n);
    for (int g = 0; g < G; ++g)
        if (pc[g] > 0.0) hc -= pc[g] / n * std::log(pc[g] / n);
    *nmi = hr + hc <= 0.0 ? 1.0 : 2.0 * I / (hr + hc);
    return DPMM_OK;
}

int dpmm_relabel(dpmm_ctx* c, uint32_t iter, int K_old, const int32_t* new_label, const int32_t* sub_action) {
    if (!c || !new_label || !sub_action || K_old <= 0 || K_old > kMaxK) return fail(DPMM_EINVAL, "bad args");
    std::vector<int32_t> table(3 * (size_t)K_old);
    bool labels_change = false;
    for (int q = 0; q < 2 * K_old; ++q) {
        table[q] = new_label[q];
        if (new_label[q] != q / 2) labels_change = true;
    }
    for (int k = 0; k < K_old; ++k) table[2 * K_old + k] = sub_action[k];
    int rc = grow(c->dTable, c->tableCap, table.size());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->dTable, table.data(), table.size() * 4, hipMemcpyHostToDevice, c->stream));
    const int grid = (int)((c->N + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_relabel, dim3(grid), dim3(kThreads), 0, c->stream, c->dLabels, c->dSubs, c->N,
                       c->dTable, K_old, c->key0(), c->key1(), iter);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    bool resets_only = !labels_change;
    for (int k = 0; k < K_old; ++k) resets_only = resets_only && (sub_action[k] == 0 || sub_action[k] == 1);
    // resets after a pass that left the statistics of the partition with the sub-labels in partition order:
    // the next pass updates the reset clusters from their moved sub-labels
    c->reset_update_ok = resets_only && c->inc_base && c->subs_perm_valid && c->partition_valid &&
                         c->Kpartition == K_old && c->Kstats == K_old;
    c->inc_base = false;
    if (labels_change) {
        c->partition_valid = false;
        c->fused_subs = false;
    }
    c->subs_perm_valid = false;   // sub-labels were rewritten by index
    c->stats_cached = false;
    if (resets_only && ((c->partials_valid && c->partialsK == K_old) || c->reset_update_ok)) {
        if (!c->dirty_pending) c->dirty_host.assign(K_old, 0);
        for (int k = 0; k < K_old; ++k) c->dirty_host[k] |= sub_action[k] != 0 ? 1 : 0;
        c->dirty_pending = true;
    } else {
        c->partials_valid = false;
        c->dirty_pending = false;
    }
    return DPMM_OK;
}

int dpmm_last_kernel_ms(dpmm_ctx* c, float* a, float* b, float* s) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (a) { *a = 0; if (c->ev_label) (void)hipEventElapsedTime(a, c->ev[0], c->ev[1]); }
    if (b) { *b = 0; if (c->ev_sub) (void)hipEventElapsedTime(b, c->ev[2], c->ev[3]); }
    if (s) { *s = 0; if (c->ev_stats) (void)hipEventElapsedTime(s, c->ev[4], c->ev[5]); }
    return DPMM_OK;
}

int dpmm_last_sweep_kernel_ms(dpmm_ctx* c, float* ms) {
    if (!c || !ms) return fail(DPMM_EINVAL, "bad args");
    HIPCHK(hipStreamSynchronize(c->stream));
    *ms = 0;
    if (c->ev_main) (void)hipEventElapsedTime(ms, c->ev[6], c->ev[7]);
    return DPMM_OK;
}

int dpmm_set_mixed_tolerance(dpmm_ctx* c, float tau) {
    if (!c || !(tau >= 0.0f) || !(tau <= 16.0f)) return fail(DPMM_EINVAL, "bound scale must be in [0, 16]");
    c->tau = tau;
    return DPMM_OK;
}

int dpmm_screen_error_factors(int DP, float* out) {
    if (!out || (DP != 32 && DP != 64 && DP != 96 && DP != 128)) return fail(DPMM_EINVAL, "DP must be 32, 64, 96 or 128");
    out[0] = label_screen_eps_of(DP);
    out[1] = split_screen_eps(DP);
    out[2] = sub_rel_eps_a(DP);
    out[3] = sub_rel_eps_g(DP);
    return DPMM_OK;
}

int dpmm_last_recheck_count(dpmm_ctx* c, int64_t* n) {
    if (!c || !n) return fail(DPMM_EINVAL, "bad args");
    *n = -1;
    if (!c->dRecheckCount) return DPMM_OK;
    int v = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&v, c->dRecheckCount, sizeof(int), hipMemcpyDeviceToHost));
    *n = v;
    return DPMM_OK;
}

int dpmm_set_statistics_refresh(dpmm_ctx* c, int every) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    c->stats_refresh = every;
    c->inc_passes = 0;
    return DPMM_OK;
}

int dpmm_last_statistics_update(dpmm_ctx* c, int* moved_only) {
    if (!c || !moved_only) return fail(DPMM_EINVAL, "bad args");
    *moved_only = 0;
    if (!c->timed_inc) return DPMM_OK;
    int flag = 1;
    HIPCHK(hipMemcpyAsync(&flag, c->dFlags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *moved_only = flag == 0 ? 1 : 2;
    return DPMM_OK;
}

int dpmm_set_label_pruning(dpmm_ctx* c, int mode) {
    if (!c || mode < 0 || mode > 2) return fail(DPMM_EINVAL, "pruning mode is 0, 1 or 2");
    c->prune_mode = mode;
    c->prune_skip = 0;
    return DPMM_OK;
}

int dpmm_set_label_refinement(dpmm_ctx* c, int enable) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    c->refine_mode = enable ? 1 : 0;
    return DPMM_OK;
}

int dpmm_set_sweep_workgroups(dpmm_ctx* c, int per_chunk) {
    if (!c || per_chunk < 0 || per_chunk > 64) return fail(DPMM_EINVAL, "workgroups per chunk is 0 .. 64");
    c->prune_split = per_chunk;
    return DPMM_OK;
}

int dpmm_set_fused_sub_labels(dpmm_ctx* c, int enable) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    c->fuse_mode = enable ? 1 : 0;
    return DPMM_OK;
}

int dpmm_last_label_evaluations(dpmm_ctx* c, double* fraction) {
    if (!c || !fraction) return fail(DPMM_EINVAL, "bad args");
    *fraction = -1.0;
    if (c->last_pruned_K <= 0 || !c->dEvals) return DPMM_OK;
    unsigned long long v[2] = {0, 0};
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(v, c->dEvals, sizeof(v), hipMemcpyDeviceToHost));
    if (v[1] > 0) *fraction = (double)v[0] / ((double)v[1] * c->last_pruned_K);
    return DPMM_OK;
}

int dpmm_last_sub_recheck_count(dpmm_ctx* c, int64_t* n) {
    if (!c || !n) return fail(DPMM_EINVAL, "bad args");
    *n = -1;
    if (!c->dRecheckCount) return DPMM_OK;
    int v = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&v, c->dRecheckCount + 1, sizeof(int), hipMemcpyDeviceToHost));
    *n = v;
    return DPMM_OK;
}

int dpmm_synchronize(dpmm_ctx* c) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    HIPCHK(hipStreamSynchronize(c->stream));
    return DPMM_OK;
}

// ---- batched conjugate work ---------------------------------------------------------------------
static int make_prior(dpmm_ctx* c, double k0, const double* m0, double v0, const double* psi0, const double* alpha0,
                      PriorArgs& out, double*& dbuf) {
    const int D = c->D;
    PriorArgs h{};
    std::vector<double> host;
    if (c->prior == DPMM_PRIOR_GAUSSIAN) {
        if (!m0 || !psi0 || !(k0 > 0.0) || !(v0 > D - 1.0)) return fail(DPMM_EINVAL, "bad NIW hyperparameters");
        std::vector<double> L;
        if (!chol_lower(D, psi0, L)) return fail(DPMM_ENOTSPD, "prior psi is not SPD");
        double ld = 0.0;
        for (int i = 0; i < D; ++i) ld += std::log(L[(size_t)i * D + i]);
        h.k0 = k0;
        h.v0 = v0;
        h.logdet_psi0 = 2.0 * ld;
        double lg = (double)D * (D - 1) / 4.0 * std::log(kPiF64);
        for (int d = 1; d <= D; ++d) lg += std::lgamma(v0 / 2.0 + (1 - d) / 2.0);
        h.lmvg0 = lg;
        double tp = 0.0;
        for (int d = 0; d < D; ++d) tp += psi0[(size_t)d * D + d];
        h.trpsi0 = v0 * tp;
        host.assign(m0, m0 + D);
        host.insert(host.end(), psi0, psi0 + (size_t)D * D);
    } else {
        if (!alpha0) return fail(DPMM_EINVAL, "alpha0 is null");
        double s = 0.0;
        for (int d = 0; d < D; ++d) {
            if (!(alpha0[d] > 0.0)) return fail(DPMM_EINVAL, "alpha0 must be positive");
            s += alpha0[d];
        }
        h.alpha0_sum = s;
        h.lg_alpha0_sum = std::lgamma(s);
        host.assign(alpha0, alpha0 + D);
    }
    // the new prior's buffer first: a failure leaves the previous prior in place
    double* nb = nullptr;
    if (dev_malloc(&nb, host.size() * sizeof(double)) != hipSuccess) return fail(DPMM_ENOMEM, "prior");
    if (hipMemcpy(nb, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(nb);
        return fail(DPMM_EHIP, "prior copy");
    }
    if (dbuf) (void)hipFree(dbuf);
    dbuf = nb;
    if (c->prior == DPMM_PRIOR_GAUSSIAN) {
        h.m0 = dbuf;
        h.psi0 = dbuf + D;
    } else {
        h.alpha0 = dbuf;
    }
    out = h;
    return DPMM_OK;
}

int dpmm_set_prior(dpmm_ctx* c, double k0, const double* m0, double v0, const double* psi0, const double* alpha0) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    const int rc = make_prior(c, k0, m0, v0, psi0, alpha0, c->prior_args, c->dPrior);
    if (rc) return rc;
    c->has_prior = true;
    return DPMM_OK;
}

int dpmm_set_outlier_prior(dpmm_ctx* c, double k0, const double* m0, double v0, const double* psi0,
                           const double* alpha0) {
    if (!c) return fail(DPMM_EINVAL, "null ctx");
    const int rc = make_prior(c, k0, m0, v0, psi0, alpha0, c->prior_args1, c->dPrior1);
    if (rc) return rc;
    c->has_prior1 = true;
    return DPMM_OK;
}

namespace {
// Pinned host buffer of at least `bytes`; a larger one replaces it (never while a sweep is being
// captured).  Callers synchronise the stream before returning, so no copy uses the old buffer.
int pin_reserve(void*& p, size_t& cap, size_t bytes) {
    if (cap >= bytes) return DPMM_OK;
    if (t_capturing) return fail(DPMM_ESTATE, "buffer growth while a sweep is being captured");
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    if (hipHostMalloc(&p, bytes) != hipSuccess) return fail(DPMM_ENOMEM, "pinned host buffer");
    cap = bytes;
    return DPMM_OK;
}

int upload_members(dpmm_ctx* c, int n, const int32_t* members) {
    const int lim = 2 * c->Kstats;
    for (size_t e = 0; e < (size_t)4 * n; ++e)
        if (members[e] < -1 || members[e] >= lim) return fail(DPMM_EINVAL, "member outside the current statistics");
    int rc = grow(c->dMembers, c->membersCap, (size_t)4 * n);
    if (rc) return rc;
    const size_t bytes = (size_t)4 * n * sizeof(int32_t);
    if ((rc = pin_reserve(c->pinMembers, c->pinMembersCap, bytes))) return rc;
    std::memcpy(c->pinMembers, members, bytes);
    HIPCHK(hipMemcpyAsync(c->dMembers, c->pinMembers, bytes, hipMemcpyHostToDevice, c->stream));
    return DPMM_OK;
}

size_t lds_bytes(size_t doubles) { return std::max<size_t>(doubles, 1) * sizeof(double); }

// Opt a kernel in to more than the default dynamic LDS when a large D needs it.
int allow_lds(const void* kernel, size_t bytes) {
    if (bytes <= 65536) return DPMM_OK;
    HIPCHK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return DPMM_OK;
}
}  // namespace

int dpmm_set_statistics(dpmm_ctx* c, int K, const int64_t* Nsub, const double* sum, const double* S) {
    if (!c || K <= 0 || K > kMaxK || !Nsub || !sum) return fail(DPMM_EINVAL, "bad args");
    const bool gauss = c->prior == DPMM_PRIOR_GAUSSIAN;
    if (gauss && !S) return fail(DPMM_EINVAL, "Gaussian statistics need S");
    const size_t E = c->statEntries();
    const int D = c->D, DP = c->DP;
    std::vector<double> h((size_t)2 * K * E, 0.0);
    for (int q = 0; q < 2 * K; ++q) {
        double* dst = h.data() + (size_t)q * E;
        if (Nsub[q] < 0) return fail(DPMM_EINVAL, "negative count");
        dst[0] = (double)Nsub[q];
        for (int d = 0; d < D; ++d) dst[1 + d] = sum[(size_t)q * D + d];
        if (gauss)
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) dst[1 + DP + (size_t)a * DP + b] = S[((size_t)q * D + a) * D + b];
    }
    int rc = grow(c->dStats, c->statsCap, h.size());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->dStats, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->Kstats = K;
    c->stats_cached = false;
    c->partials_valid = false;   // not the statistics of the current labels
    c->inc_base = false;
    return DPMM_OK;
}

int dpmm_log_marginal_batch(dpmm_ctx* c, int n, const int32_t* members, double* logml, double* logdet_psi) {
    return dpmm_log_marginal_batch_ex(c, n, members, nullptr, logml, logdet_psi);
}

int dpmm_log_marginal_batch_ex(dpmm_ctx* c, int n, const int32_t* members, const int32_t* prior_sel, double* logml,
                               double* logdet_psi) {
    if (!c || n < 0 || (n > 0 && (!members || !logml))) return fail(DPMM_EINVAL, "bad args");
    if (!c->has_prior) return fail(DPMM_ESTATE, "prior not set");
    if (c->Kstats <= 0) return fail(DPMM_ESTATE, "no statistics computed yet");
    if (n == 0) return DPMM_OK;
    int rc = upload_members(c, n, members);
    if (rc) return rc;
    const int32_t* dsel = nullptr;
    if (prior_sel) {
        bool any = false;
        for (int j = 0; j < n; ++j) {
            if (prior_sel[j] < 0 || prior_sel[j] > 3) return fail(DPMM_EINVAL, "prior selector outside 0..3");
            any |= prior_sel[j] != 0;
        }
        if (any && !c->has_prior1) return fail(DPMM_ESTATE, "outlier prior not set");
        if (any) {
            if ((rc = grow(c->dPsel, c->pselCap, (size_t)n))) return rc;
            HIPCHK(hipMemcpyAsync(c->dPsel, prior_sel, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
            dsel = c->dPsel;
        }
    }
    rc = grow(c->dLM, c->lmCap, (size_t)2 * n);
    if (rc) return rc;
    const int D = c->D;
    const size_t lds = c->prior == DPMM_PRIOR_GAUSSIAN ? lds_bytes((size_t)D * (D + 1) / 2 + D) : lds_bytes(0);
    rc = allow_lds(reinterpret_cast<const void*>(k_log_marginal), lds);
    if (rc) return rc;
    hipLaunchKernelGGL(k_log_marginal, dim3(n), dim3(kThreads), lds, c->stream, c->dStats, (int)c->statEntries(),
                       c->DP, D, c->prior, c->dMembers, dsel, c->prior_args, dsel ? c->prior_args1 : c->prior_args,
                       c->dLM, nullptr, nullptr);
    HIPCHK(hipGetLastError());
    if ((rc = pin_reserve(c->pinOut, c->pinOutCap, (size_t)2 * n * sizeof(double)))) return rc;
    const double* out = static_cast<const double*>(c->pinOut);
    HIPCHK(hipMemcpyAsync(c->pinOut, c->dLM, (size_t)2 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < n; ++j) {
        logml[j] = out[2 * j];
        if (logdet_psi) logdet_psi[j] = out[2 * j + 1];
    }
    return DPMM_OK;
}

// The parameter draw of 3K parts and the packing of the sweeps' parameter slots, enqueued without
// a synchronisation.  dmem (4 statistics entries per part), the draw control dctl (or null) and the
// log weights at c->dLogW (K mixture weights, then 2K sub-cluster weights) are on the device.
static int enqueue_draws(dpmm_ctx* c, uint32_t iter, int K, const int32_t* dmem, const int32_t* dctl,
                         int full_mu_cov, int use_weight) {
    const int D = c->D, np = 3 * K;
    const int dstride = c->drawStride();
    int rc;
    if ((rc = grow(c->dDraws, c->drawsCap, (size_t)np * dstride))) return rc;
    if ((rc = grow(c->dDrawStatus, c->drawStatusCap, (size_t)np))) return rc;
    const bool gauss = c->prior == DPMM_PRIOR_GAUSSIAN;
    const size_t lds = gauss ? lds_bytes((size_t)D * (D + 1) + 3 * D) : lds_bytes(D);
    if ((rc = allow_lds(reinterpret_cast<const void*>(k_draw_params), lds))) return rc;
    hipLaunchKernelGGL(k_draw_params, dim3(np), dim3(kThreads), lds, c->stream, c->dStats, (int)c->statEntries(),
                       c->DP, D, c->prior, dmem, dctl, c->dDrawsPrev, c->prior_args,
                       c->has_prior1 ? c->prior_args1 : c->prior_args, iter, c->key0(), c->key1(), full_mu_cov,
                       dstride, c->dDraws, c->dDrawStatus);
    HIPCHK(hipGetLastError());
    // parameter slots of the sweeps
    if (gauss) {
        const bool f64 = c->precision == DPMM_PREC_FP64;
        const int stride = gauss_slot_stride(c->DP);
        const size_t esz = f64 ? 8 : 4;
        c->params_f64 = f64;
        c->stride = stride;
        if ((rc = ensure_param_buffers(c, K, (size_t)stride * esz))) return rc;
        unsigned char* pb = nullptr;
        unsigned char* spb = nullptr;
        unsigned char* plo = nullptr;
        c->spb_valid = false;
        c->fused_subs = false;
        if (c->precision == DPMM_PREC_MIXED) {
            if ((rc = ensure_mixed_buffers(c, K))) return rc;
            pb = reinterpret_cast<unsigned char*>(c->dPB);
            spb = reinterpret_cast<unsigned char*>(c->dSPB);
            plo = reinterpret_cast<unsigned char*>(c->dPBlo);
            c->spb_valid = true;
        }
        if (f64)
            hipLaunchKernelGGL(k_pack_gauss<double>, dim3(np), dim3(kThreads), 0, c->stream, c->dDraws, dstride, D,
                               c->DP, c->DPM, c->dLogW, K, reinterpret_cast<double*>(c->dP),
                               reinterpret_cast<double*>(c->dSP), pb, spb, plo);
        else
            hipLaunchKernelGGL(k_pack_gauss<float>, dim3(np), dim3(kThreads), 0, c->stream, c->dDraws, dstride, D,
                               c->DP, c->DPM, c->dLogW, K, reinterpret_cast<float*>(c->dP),
                               reinterpret_cast<float*>(c->dSP), pb, spb, plo);
    } else {
        const int stride = (D + 1 + 3) & ~3;
        c->params_f64 = false;
        c->stride = stride;
        if ((rc = ensure_param_buffers(c, K, (size_t)stride * 4))) return rc;
        hipLaunchKernelGGL(k_pack_mult, dim3(np), dim3(kThreads), 0, c->stream, c->dDraws, dstride, D, stride,
                           c->dLogW, K, use_weight, reinterpret_cast<float*>(c->dP), reinterpret_cast<float*>(c->dSP));
        c->mm_ncb = 0;
        if (c->precision == DPMM_PREC_MIXED && (rc = pack_mult_mixed(c, K, c->dDraws, (size_t)3 * dstride, c->dLogW, use_weight)))
            return rc;
    }
    HIPCHK(hipGetLastError());
    return DPMM_OK;
}

int dpmm_sample_params(dpmm_ctx* c, uint32_t iter, int K, const int32_t* members, const double* weights,
                       const double* lr_weights, int full_mu_cov, int use_weight) {
    return dpmm_sample_params_ex(c, iter, K, members, nullptr, weights, lr_weights, full_mu_cov, use_weight);
}

int dpmm_sample_params_ex(dpmm_ctx* c, uint32_t iter, int K, const int32_t* members, const int32_t* ctl,
                          const double* weights, const double* lr_weights, int full_mu_cov, int use_weight) {
    if (!c || K <= 0 || !members || !weights || !lr_weights) return fail(DPMM_EINVAL, "bad args");
    if (K > kMaxK) return fail(DPMM_EINVAL, "K exceeds 2048");
    if (!c->has_prior) return fail(DPMM_ESTATE, "prior not set");
    if (c->Kstats <= 0) return fail(DPMM_ESTATE, "no statistics computed yet");
    const int np = 3 * K;
    int rc = upload_members(c, np, members);
    if (rc) return rc;
    const int dstride0 = c->drawStride();
    const int32_t* dctl = nullptr;
    if (ctl) {
        bool keep = false;
        for (int p = 0; p < np; ++p) {
            const int32_t* q = ctl + 4 * p;
            if (q[0] >= 3 * c->Kdraw || q[0] < -1) return fail(DPMM_EINVAL, "kept row outside the previous draws");
            if (q[1] < -1 || q[2] < -1 || q[3] < 0 || q[3] > 1) return fail(DPMM_EINVAL, "bad draw control");
            if (q[0] < 0 && q[3] == 1 && !c->has_prior1) return fail(DPMM_ESTATE, "outlier prior not set");
            keep |= q[0] >= 0;
        }
        if (keep) {   // the previous rows, before the draw buffer is reused (or reallocated) below
            if ((rc = grow(c->dDrawsPrev, c->drawsPrevCap, (size_t)3 * c->Kdraw * dstride0))) return rc;
            HIPCHK(hipMemcpyAsync(c->dDrawsPrev, c->dDraws, (size_t)3 * c->Kdraw * dstride0 * sizeof(double),
                                  hipMemcpyDeviceToDevice, c->stream));
        }
        if ((rc = grow(c->dCtl, c->ctlCap, (size_t)4 * np))) return rc;
        HIPCHK(hipMemcpyAsync(c->dCtl, ctl, (size_t)4 * np * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        dctl = c->dCtl;
    }
    if ((rc = pin_reserve(c->pinLogW, c->pinLogWCap, (size_t)3 * K * sizeof(double)))) return rc;
    double* logw = static_cast<double*>(c->pinLogW);
    for (int k = 0; k < K; ++k) {
        if (!(weights[k] >= 0.0) || !(lr_weights[2 * k] >= 0.0) || !(lr_weights[2 * k + 1] >= 0.0))
            return fail(DPMM_EINVAL, "weights must be non-negative");
        logw[k] = std::log(weights[k]);
        logw[K + 2 * k] = std::log(lr_weights[2 * k]);
        logw[K + 2 * k + 1] = std::log(lr_weights[2 * k + 1]);
    }
    if ((rc = grow(c->dLogW, c->logwCap, (size_t)3 * K))) return rc;
    HIPCHK(hipMemcpyAsync(c->dLogW, logw, (size_t)3 * K * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = enqueue_draws(c, iter, K, c->dMembers, dctl, full_mu_cov, use_weight))) return rc;
    if ((rc = pin_reserve(c->pinOut, c->pinOutCap, np * sizeof(int32_t)))) return rc;
    const int32_t* st = static_cast<const int32_t*>(c->pinOut);
    HIPCHK(hipMemcpyAsync(c->pinOut, c->dDrawStatus, np * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->Kdraw = K;
    for (int p = 0; p < np; ++p)
        if (st[p]) return fail(DPMM_ENOTSPD, "posterior draw failed: a factor is not SPD");
    c->K = K;
    c->prune_valid = false;
    return DPMM_OK;
}

int dpmm_pair_features(dpmm_ctx* c, int K, const double* scal, int npairs, const int32_t* pi, const int32_t* pj,
                       float* features) {
    if (!c || K <= 0 || !scal || npairs < 0 || (npairs > 0 && (!pi || !pj || !features)))
        return fail(DPMM_EINVAL, "bad args");
    if (c->prior != DPMM_PRIOR_GAUSSIAN) return fail(DPMM_EINVAL, "pair features need a Gaussian context");
    if (!c->has_prior) return fail(DPMM_ESTATE, "prior not set");
    if (K > c->Kstats) return fail(DPMM_ESTATE, "statistics hold fewer clusters");
    if (npairs == 0) return DPMM_OK;
    for (int p = 0; p < npairs; ++p)
        if (pi[p] < 0 || pi[p] >= K || pj[p] < 0 || pj[p] >= K) return fail(DPMM_EINVAL, "pair outside 0..K-1");
    const int D = c->D;
    // device scratch: summaries (K x 2D), scalars (K x 7), pairs (2 x npairs), features (npairs x 16)
    const size_t nd = (size_t)K * 2 * D + (size_t)K * 7;
    int rc = grow(c->dLM, c->lmCap, std::max<size_t>(nd, (size_t)2 * K));
    if (rc) return rc;
    rc = grow(c->dMembers, c->membersCap, (size_t)2 * npairs);
    if (rc) return rc;
    rc = grow(c->dFeat, c->featCap, (size_t)16 * npairs);
    if (rc) return rc;
    double* dSumm = c->dLM;
    double* dScal = c->dLM + (size_t)K * 2 * D;
    HIPCHK(hipMemcpyAsync(dScal, scal, (size_t)K * 7 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dMembers, pi, (size_t)npairs * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dMembers + npairs, pj, (size_t)npairs * 4, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_cluster_summary, dim3(K), dim3(kThreads), 0, c->stream, c->dStats, (int)c->statEntries(),
                       c->DP, D, c->prior_args, dSumm);
    hipLaunchKernelGGL(k_pair_features, dim3((npairs + 3) / 4), dim3(kThreads), 0, c->stream, dSumm, dScal, D,
                       npairs, c->dMembers, c->dMembers + npairs, c->dFeat);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(features, c->dFeat, (size_t)16 * npairs * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DPMM_OK;
}

int dpmm_get_drawn_params(dpmm_ctx* c, double* mu, double* chol_inv_sigma, double* logdet_sigma) {
    if (!c || !mu) return fail(DPMM_EINVAL, "bad args");
    if (c->Kdraw <= 0) return fail(DPMM_ESTATE, "no draws yet");
    const int D = c->D, np = 3 * c->Kdraw, dstride = c->drawStride();
    std::vector<double> h((size_t)np * dstride);
    HIPCHK(hipMemcpyAsync(h.data(), c->dDraws, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int T = D * (D + 1) / 2;
    for (int p = 0; p < np; ++p) {
        const double* src = h.data() + (size_t)p * dstride;
        std::copy(src, src + D, mu + (size_t)p * D);
        if (c->prior != DPMM_PRIOR_GAUSSIAN) continue;
        if (chol_inv_sigma) {
            double* R = chol_inv_sigma + (size_t)p * D * D;
            std::fill(R, R + (size_t)D * D, 0.0);
            for (int i = 0; i < D; ++i)
                for (int j = 0; j <= i; ++j) R[(size_t)i * D + j] = src[D + i * (i + 1) / 2 + j];
        }
        if (logdet_sigma) logdet_sigma[p] = src[D + T];
    }
    return DPMM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- device-resident iteration ----
namespace {

int grid_relabel(const dpmm_ctx* c) { return (int)std::min<int64_t>((c->N + kThreads - 1) / kThreads, 1024); }

// Every cache derived from the labels or sub-labels, after a relabel pass on the device.
void labels_rewritten(dpmm_ctx* c) {
    c->partition_valid = false;
    c->fused_subs = false;
    c->subs_perm_valid = false;
    c->stats_cached = false;
    c->partials_valid = false;
    c->dirty_pending = false;
    c->inc_base = false;
    c->reset_update_ok = false;
}

// k_log_marginal over the device-built job list jobs (4 statistics entries each), the grid sized for
// `grid` jobs and the count read from njobs; results (log ML, log det psi_n) at c->dLM, in row
// out_map[j] of job j when out_map is given.
int launch_log_marginal_dev(dpmm_ctx* c, int grid, const int32_t* jobs, const int32_t* njobs,
                            const int32_t* out_map = nullptr) {
    const int D = c->D;
    const size_t lds = c->prior == DPMM_PRIOR_GAUSSIAN ? lds_bytes((size_t)D * (D + 1) / 2 + D) : lds_bytes(0);
    int rc = allow_lds(reinterpret_cast<const void*>(k_log_marginal), lds);
    if (rc) return rc;
    if (grid <= 0) return DPMM_OK;
    hipLaunchKernelGGL(k_log_marginal, dim3(grid), dim3(kThreads), lds, c->stream, c->dStats, (int)c->statEntries(),
                       c->DP, D, c->prior, jobs, nullptr, c->prior_args, c->prior_args, c->dLM, njobs, out_map);
    HIPCHK(hipGetLastError());
    return DPMM_OK;
}

// What enqueue_draws grows for K clusters, grown ahead of a capture.
int ensure_draw_buffers(dpmm_ctx* c, int K) {
    const int np = 3 * K, D = c->D;
    int rc;
    if ((rc = grow(c->dDraws, c->drawsCap, (size_t)np * c->drawStride()))) return rc;
    if ((rc = grow(c->dDrawStatus, c->drawStatusCap, (size_t)np))) return rc;
    if (c->prior == DPMM_PRIOR_GAUSSIAN) {
        const size_t esz = c->precision == DPMM_PREC_FP64 ? 8 : 4;
        if ((rc = ensure_param_buffers(c, K, (size_t)gauss_slot_stride(c->DP) * esz))) return rc;
        if (c->precision == DPMM_PREC_MIXED && (rc = ensure_mixed_buffers(c, K))) return rc;
    } else {
        if ((rc = ensure_param_buffers(c, K, (size_t)((D + 1 + 3) & ~3) * 4))) return rc;
        if (c->precision == DPMM_PREC_MIXED && K <= 256) {
            if ((rc = grow(c->dMB, c->mbCap, mm_block_bytes(D, (K + 31) / 32)))) return rc;
            if ((rc = ensure_recheck(c))) return rc;
        }
    }
    return DPMM_OK;
}

// One iteration from the parameter draw to the summary copy, enqueued without a synchronisation.
int enqueue_step(dpmm_ctx* c, const StepArgs& a0, int sweep_final, int full_mu_cov, int use_weight, int grid_jobs,
                 int32_t* trace_labels, int32_t* trace_subs) {
    const int K = a0.K;
    // the merge bound applies when there are candidate pairs to test
    StepArgs a = a0;
    a.merge_bound = a0.merge_bound && grid_jobs > 3 * K ? 1 : 0;
    hipLaunchKernelGGL(k_step_sample, dim3(1), dim3(kStepThreads), 0, c->stream, a, c->dStepCl, c->dStepHist,
                       c->dStepW, c->dStepMem, c->dLogW);
    HIPCHK(hipGetLastError());
    int rc = enqueue_draws(c, a.t, K, c->dStepMem, nullptr, full_mu_cov, use_weight);
    if (rc) return rc;
    c->Kdraw = K;
    c->K = K;
    c->prune_valid = false;
    if ((rc = dpmm_sample_cluster_labels(c, a.t, sweep_final))) return rc;
    if ((rc = dpmm_sample_sub_labels(c, a.t))) return rc;
    if ((rc = enqueue_stats(c, K))) return rc;
    if (trace_labels)
        HIPCHK(hipMemcpyAsync(trace_labels, c->dLabels, (size_t)c->N * 4, hipMemcpyDeviceToHost, c->stream));
    if (trace_subs) HIPCHK(hipMemcpyAsync(trace_subs, c->dSubs, (size_t)c->N * 4, hipMemcpyDeviceToHost, c->stream));
    const int npairs_max = grid_jobs - 3 * K;
    const bool bound = a.merge_bound != 0;
    // with the merge bound: the pairs' output map, then the map and prior selectors of the 4K jobs of
    // the first launch; W_k's results in log-ML rows wrow + k, past the pairs'
    int32_t* pmap = c->dStepMap;
    int32_t* omap = pmap + npairs_max;
    int32_t* psel = omap + 4 * K;
    const int wrow = grid_jobs;
    hipLaunchKernelGGL(k_step_jobs, dim3(1), dim3(kStepThreads), 0, c->stream, a, c->dStepCl, c->dStats,
                       c->dStepCand, c->dStepJobs, c->dStepCounts, bound ? psel : nullptr, bound ? omap : nullptr,
                       wrow);
    if (bound) {
        // the parts' log MLs with log det W_k of every cluster in one launch, the merge bound of every
        // candidate pair, then the log MLs of the pairs it keeps
        PriorArgs hw = c->prior_args;
        hw.k0 = 0.0;
        hw.v0 = 0.5 * c->prior_args.v0;
        const size_t lds = lds_bytes((size_t)c->D * (c->D + 1) / 2 + c->D);
        if ((rc = allow_lds(reinterpret_cast<const void*>(k_log_marginal), lds))) return rc;
        hipLaunchKernelGGL(k_log_marginal, dim3(4 * K), dim3(kThreads), lds, c->stream, c->dStats,
                           (int)c->statEntries(), c->DP, c->D, c->prior, c->dStepJobs, psel, c->prior_args, hw, c->dLM,
                           nullptr, omap);
        double* vec = c->dStepBnd;
        double* sc = vec + (size_t)c->D * K;
        hipLaunchKernelGGL(k_step_bound_prep, dim3(K), dim3(64), 0, c->stream, a, c->dStats, c->dLM,
                           c->dLM + (size_t)2 * wrow, c->prior_args, c->D, c->DP, vec, sc);
        int32_t* pjobs = c->dStepJobs + (size_t)4 * 4 * K;
        hipLaunchKernelGGL(k_step_pair_bound, dim3((npairs_max + kThreads / 64 - 1) / (kThreads / 64)), dim3(kThreads),
                           0, c->stream, a, c->dStepCand, c->dStepCounts, vec, sc, c->prior_args, c->D,
                           c->prior_args.trpsi0, pjobs, pmap, c->dLM);
        if ((rc = launch_log_marginal_dev(c, npairs_max, pjobs, c->dStepCounts + 2, pmap))) return rc;
    } else if ((rc = launch_log_marginal_dev(c, grid_jobs, c->dStepJobs, c->dStepCounts))) {
        return rc;
    }
    int32_t* tab1 = c->dStepTab;
    int32_t* tab2 = tab1 + 3 * kMaxK;
    int32_t* aux = tab2 + 3 * kMaxK;
    hipLaunchKernelGGL(k_step_decide, dim3(1), dim3(kStepThreads), 0, c->stream, a, c->dStepCl, c->dLM, c->dStepCand,
                       c->dStepCounts, c->dDrawStatus, 3 * K, c->dStepAcc, tab1, tab2, aux, c->dStepSum);
    hipLaunchKernelGGL(k_step_relabel, dim3(grid_relabel(c)), dim3(kThreads), 0, c->stream, c->dLabels, c->dSubs,
                       c->N, tab1, K, c->dStepSum, 1, c->key0(), c->key1(), a.t);
    hipLaunchKernelGGL(k_step_relabel, dim3(grid_relabel(c)), dim3(kThreads), 0, c->stream, c->dLabels, c->dSubs,
                       c->N, tab2, 0, c->dStepSum, 2, c->key0(), c->key1(), a.t);
    hipLaunchKernelGGL(k_step_compact, dim3(1), dim3(kStepThreads), 0, c->stream, c->dStepCl, c->dStepHist,
                       c->dStepClTmp, c->dStepHistTmp, aux, c->dStepSum, 2, a.B);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->hStepSum, c->dStepSum, sizeof(StepSummary), hipMemcpyDeviceToHost, c->stream));
    return DPMM_OK;
}

int step_error(const StepSummary& s) {
    if (s.error & 1) return fail(DPMM_ENOTSPD, "posterior draw failed: a factor is not SPD");
    if (s.error & 2) return fail(DPMM_ENOTSPD, "a log marginal likelihood is not finite");
    return DPMM_OK;
}

}  // namespace

namespace dpmm {

int step_upload(dpmm_ctx* c, int B, int K, const StepCluster* cl, const double* hist) {
    if (!c || B < 1 || K <= 0 || K > kMaxK || !cl || !hist) return fail(DPMM_EINVAL, "bad args");
    size_t cap;
    int rc;
    // fixed-size buffers (grow keeps a buffer that exists); the history rings follow the burn-out length
    cap = c->dStepCl ? (size_t)kMaxK : 0;
    if ((rc = grow(c->dStepCl, cap, (size_t)kMaxK))) return rc;
    cap = c->dStepClTmp ? (size_t)kMaxK : 0;
    if ((rc = grow(c->dStepClTmp, cap, (size_t)kMaxK))) return rc;
    cap = c->dStepW ? (size_t)kMaxK : 0;
    if ((rc = grow(c->dStepW, cap, (size_t)kMaxK))) return rc;
    cap = c->dStepMem ? (size_t)12 * kMaxK : 0;
    if ((rc = grow(c->dStepMem, cap, (size_t)12 * kMaxK))) return rc;
    cap = c->dStepCand ? (size_t)kMaxK : 0;
    if ((rc = grow(c->dStepCand, cap, (size_t)kMaxK))) return rc;
    cap = c->dStepCounts ? (size_t)4 : 0;
    if ((rc = grow(c->dStepCounts, cap, (size_t)4))) return rc;
    cap = c->dStepTab ? (size_t)8 * kMaxK : 0;
    if ((rc = grow(c->dStepTab, cap, (size_t)8 * kMaxK))) return rc;
    cap = c->dStepSum ? (size_t)1 : 0;
    if ((rc = grow(c->dStepSum, cap, (size_t)1))) return rc;
    if (c->step_B != B) {
        cap = 0;
        double* h = nullptr;
        double* ht = nullptr;
        if ((rc = grow(h, cap, (size_t)kMaxK * B))) return rc;
        cap = 0;
        if ((rc = grow(ht, cap, (size_t)kMaxK * B))) {
            (void)hipFree(h);
            return rc;
        }
        if (c->dStepHist) (void)hipFree(c->dStepHist);
        if (c->dStepHistTmp) (void)hipFree(c->dStepHistTmp);
        c->dStepHist = h;
        c->dStepHistTmp = ht;
        c->step_B = B;
    }
    if (!c->hStepSum && hipHostMalloc((void**)&c->hStepSum, sizeof(StepSummary)) != hipSuccess) {
        c->hStepSum = nullptr;
        return fail(DPMM_ENOMEM, "pinned");
    }
    HIPCHK(hipMemcpyAsync(c->dStepCl, cl, (size_t)K * sizeof(StepCluster), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dStepHist, hist, (size_t)K * B * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DPMM_OK;
}

int step_download(dpmm_ctx* c, int K, int Kw, StepCluster* cl, double* hist, double* weights) {
    if (!c || K <= 0 || K > kMaxK || Kw < 0 || Kw > kMaxK) return fail(DPMM_EINVAL, "bad args");
    if (!c->dStepCl) return fail(DPMM_ESTATE, "no cluster table on the device");
    if (cl) HIPCHK(hipMemcpyAsync(cl, c->dStepCl, (size_t)K * sizeof(StepCluster), hipMemcpyDeviceToHost, c->stream));
    if (hist)
        HIPCHK(hipMemcpyAsync(hist, c->dStepHist, (size_t)K * c->step_B * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    if (weights && Kw > 0)
        HIPCHK(hipMemcpyAsync(weights, c->dStepW, (size_t)Kw * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return DPMM_OK;
}

int step_iteration(dpmm_ctx* c, const StepArgs& args, int sweep_final, int full_mu_cov, int use_weight,
                   int32_t* trace_labels, int32_t* trace_subs, StepSummary* out) {
    if (!c || !out) return fail(DPMM_EINVAL, "bad args");
    if (!c->dStepCl || !c->hStepSum) return fail(DPMM_ESTATE, "no cluster table on the device");
    if (!c->has_prior) return fail(DPMM_ESTATE, "prior not set");
    StepArgs a = args;
    a.E = (int)c->statEntries();
    a.B = c->step_B;
    const int K = a.K;
    if (K <= 0 || K > kMaxK || a.kcap < 1 || a.kcap > kMaxK) return fail(DPMM_EINVAL, "bad cluster count");
    const int64_t maxpairs = a.no_more_splits ? 0 : (int64_t)K * (K - 1) / 2;
    const int64_t maxjobs = 3 * (int64_t)K + maxpairs;
    int rc;
    if ((rc = grow(c->dStepJobs, c->stepJobsCap, (size_t)4 * maxjobs))) return rc;
    if (a.merge_bound && maxpairs > 0) {
        if ((rc = grow(c->dStepJobs, c->stepJobsCap, (size_t)4 * (maxjobs + K)))) return rc;
        if ((rc = grow(c->dStepMap, c->stepMapCap, (size_t)maxpairs + 8 * K))) return rc;
        if ((rc = grow(c->dStepBnd, c->stepBndCap, (size_t)K * (c->D + 6)))) return rc;
        if ((rc = grow(c->dLM, c->lmCap, (size_t)2 * (maxjobs + K)))) return rc;
    }
    if ((rc = grow(c->dStepAcc, c->stepAccCap, (size_t)std::max<int64_t>(maxpairs, 1)))) return rc;
    if ((rc = grow(c->dLM, c->lmCap, (size_t)2 * maxjobs))) return rc;
    if ((rc = grow(c->dLogW, c->logwCap, (size_t)3 * K))) return rc;
    if (c->graph_mode && !trace_labels && !trace_subs) {
        // the whole iteration recorded into one hipGraph
        if ((rc = resolve_evals(c))) return rc;
        if ((rc = ensure_draw_buffers(c, K))) return rc;
        c->K = K;
        if ((rc = ensure_sweep_buffers(c))) return rc;
        rc = run_captured(c, [&]() {
            return enqueue_step(c, a, sweep_final, full_mu_cov, use_weight, (int)maxjobs, nullptr, nullptr);
        });
    } else {
        rc = enqueue_step(c, a, sweep_final, full_mu_cov, use_weight, (int)maxjobs, trace_labels, trace_subs);
    }
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    StepSummary s = *c->hStepSum;
    if (s.followup && s.nsplit == 0) {
        // resets only: the labels and the partition stand, the reset clusters' sub-labels changed
        c->reset_update_ok = c->inc_base && c->subs_perm_valid && c->partition_valid && c->Kpartition == K &&
                             c->Kstats == K;
        c->inc_base = false;
        c->subs_perm_valid = false;
        c->stats_cached = false;
        c->dirty_pending = false;
    } else if (s.followup || s.relabel2) {
        labels_rewritten(c);
    }
    if ((rc = step_error(s))) return rc;
    if (s.followup) {
        // the statistics and log ML of the reset and split clusters, then the merges and removals; with
        // resets only, the pass recomputes the reset clusters (their relabel actions are the mask)
        const int Ks = s.K_split;
        if ((rc = grow(c->dStepJobs, c->stepJobsCap, (size_t)12 * Ks))) return rc;
        if ((rc = grow(c->dLM, c->lmCap, (size_t)6 * Ks))) return rc;
        if ((rc = enqueue_stats(c, Ks, s.nsplit == 0 ? c->dStepTab + 2 * K : nullptr))) return rc;
        hipLaunchKernelGGL(k_step_touched_jobs, dim3(1), dim3(kStepThreads), 0, c->stream, Ks, a.E, c->dStepCl,
                           c->dStats, c->dStepCand, c->dStepJobs, c->dStepCounts);
        if ((rc = launch_log_marginal_dev(c, 3 * Ks, c->dStepJobs, c->dStepCounts))) return rc;
        hipLaunchKernelGGL(k_step_touched_apply, dim3(1), dim3(kStepThreads), 0, c->stream, c->dStepCl, c->dLM,
                           c->dStepCand, c->dStepCounts, c->dStepSum);
        int32_t* tab2 = c->dStepTab + 3 * kMaxK;
        int32_t* aux = tab2 + 3 * kMaxK;
        hipLaunchKernelGGL(k_step_relabel, dim3(grid_relabel(c)), dim3(kThreads), 0, c->stream, c->dLabels, c->dSubs,
                           c->N, tab2, 0, c->dStepSum, 3, c->key0(), c->key1(), a.t);
        hipLaunchKernelGGL(k_step_compact, dim3(1), dim3(kStepThreads), 0, c->stream, c->dStepCl, c->dStepHist,
                           c->dStepClTmp, c->dStepHistTmp, aux, c->dStepSum, 3, a.B);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c->hStepSum, c->dStepSum, sizeof(StepSummary), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        s = *c->hStepSum;
        if (s.relabel2) labels_rewritten(c);
        if ((rc = step_error(s))) return rc;
    }
    *out = s;
    return DPMM_OK;
}

}  // namespace dpmm

// Internal accessors for the host driver (same library).
namespace dpmm {
int set_error(int code, const char* msg) { return fail(code, msg); }

bool ctx_has_outlier_prior(const dpmm_ctx* c) { return c && c->has_prior1; }

int ctx_dims(const dpmm_ctx* c, int64_t* N, int* D, int* prior, int* precision) {
    if (!c) return DPMM_EINVAL;
    *N = c->N; *D = c->D; *prior = c->prior; *precision = c->precision;
    return DPMM_OK;
}
}  // namespace dpmm
