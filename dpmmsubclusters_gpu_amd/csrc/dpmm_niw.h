// Batched FP64 conjugate work of the sampler on the device.  One workgroup per job:
//   k_log_marginal  posterior of a sum of sub-cluster statistics (niw::calc_posterior,
//                   priors/niw.cpp:23-49; multinomial_prior::calc_posterior, 12-22), its
//                   Cholesky factor and log marginal likelihood (niw.cpp:122-138;
//                   multinomial_prior.cpp:56-67).  Every cluster, sub-cluster and merge pair of
//                   an iteration goes through one launch.
//   k_draw_params   parameter draw from such a posterior (niw::sample_distribution,
//                   niw.cpp:82-107, Bartlett form of rinvwish.ipp:35-58; Dirichlet draw of
//                   multinomial_prior.cpp:24-49) on the part's Philox stream.
//   k_pack_params   the draws, with log mixture / sub-cluster weights, into the label and
//                   sub-label sweeps' parameter slots (FP32 or FP64, and the BF16 screen).
// Matrices live in LDS in packed lower-triangular form, element (i, j), j <= i, at
// i (i + 1) / 2 + j, so two D = 128 factors fit in one workgroup's LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpmm_kernels.h"
#include "dpmm_mixed.h"
#include "dpmm_rng.h"

namespace dpmm {

constexpr double kPiF64 = 3.14159265358979323846;

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// (i, j) of packed index e.
__device__ __forceinline__ void tri_ij(int e, int& i, int& j) {
    int r = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
    while (r * (r + 1) / 2 > e) --r;
    while ((r + 1) * (r + 2) / 2 <= e) ++r;
    i = r;
    j = e - r * (r + 1) / 2;
}

struct PriorArgs {
    double k0, v0, logdet_psi0, lmvg0;   // lmvg0 = log Gamma_D(v0 / 2)
    double trpsi0;                       // v0 tr psi0
    double alpha0_sum, lg_alpha0_sum;    // multinomial: sum alpha0, lgamma(sum alpha0)
    const double* m0;                    // D
    const double* psi0;                  // D x D row-major (the reference's psi = Psi / v0)
    const double* alpha0;                // D
};

__device__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    __syncthreads();
    return s;
}

// Sequential stream (c0, n, t, tag) from word position w: the words of Philox blocks n = w / 4, ...
struct DevStream {
    uint32_t c0, t, tag, k0, k1, w, n;
    Ctr4 blk;
    __device__ DevStream(uint32_t c0_, uint32_t t_, uint32_t tag_, uint32_t k0_, uint32_t k1_, uint32_t w0)
        : c0(c0_), t(t_), tag(tag_), k0(k0_), k1(k1_), w(w0), n(0xFFFFFFFFu) {}
    __device__ uint32_t word() {
        const uint32_t b = w >> 2;
        if (b != n) { blk = philox4x32_10(c0, b, t, tag, k0, k1); n = b; }
        const uint32_t p = w & 3u;
        ++w;
        return p == 0 ? blk.v[0] : p == 1 ? blk.v[1] : p == 2 ? blk.v[2] : blk.v[3];
    }
    __device__ double uniform() { return u01_f64(word()); }
    __device__ double normal() {   // Box-Muller, cosine branch
        const double a = uniform();
        const double b = uniform();
        return sqrt(-2.0 * log(a)) * cos(2.0 * kPiF64 * b);
    }
    // Marsaglia & Tsang (2000); a < 1 through Gamma(a + 1) U^(1/a).
    __device__ double gamma(double a) {
        if (!(a > 0.0)) return 0.0;
        const bool boost = a < 1.0;
        const double ab = boost ? a + 1.0 : a;
        const double d = ab - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
        double g;
        for (;;) {
            const double z = normal();
            double v = 1.0 + c * z;
            if (v <= 0.0) continue;
            v = v * v * v;
            if (log(uniform()) < 0.5 * z * z + d - d * v + d * log(v)) { g = d * v; break; }
        }
        if (boost) g *= pow(uniform(), 1.0 / a);
        return g;
    }
};

// Normal number m of a stream whose first normals start at word 0: words 2m, 2m + 1.
__device__ __forceinline__ double stream_normal_at(uint32_t c0, uint32_t t, uint32_t k0, uint32_t k1, uint32_t w) {
    DevStream s(c0, t, TAG_PARAM, k0, k1, w);
    return s.normal();
}

// ---- posterior of one statistics sum ----------------------------------------------------------
// stats: entries of E doubles [N, sum (DP), S (DP x DP)]; mem: up to 4 entry indices (-1 unused).
// Fills P (packed lower posterior psi, scaled by `scale`), m (D) and scal = {N, kappa, nu}.
__device__ void niw_posterior(const double* __restrict__ stats, size_t E, int DP, int D, const int* mem,
                              const PriorArgs& h, double scale_by_nu, double* P, double* m, double* scal) {
    if (threadIdx.x == 0) {
        // the members' counts loaded together (an unused member reads entry 0, which exists when any
        // member does, and is not added)
        const bool any = mem[0] >= 0 || mem[1] >= 0 || mem[2] >= 0 || mem[3] >= 0;
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        if (any) {
#pragma unroll
            for (int t = 0; t < 4; ++t) c[t] = stats[(size_t)(mem[t] >= 0 ? mem[t] : 0) * E];
        }
        double n = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (mem[t] >= 0) n += c[t];
        scal[0] = n;
        scal[1] = n == 0.0 ? h.k0 : h.k0 + n;
        scal[2] = n == 0.0 ? h.v0 : h.v0 + n;
    }
    __syncthreads();
    const double N = scal[0], kn = scal[1], vn = scal[2];
    const int T = D * (D + 1) / 2;
    if (N == 0.0) {   // no points: the posterior is the prior itself
        for (int d = threadIdx.x; d < D; d += blockDim.x) m[d] = h.m0[d];
        for (int e = threadIdx.x; e < T; e += blockDim.x) {
            int a, b;
            tri_ij(e, a, b);
            const double p = h.psi0[(size_t)a * D + b];
            P[e] = scale_by_nu != 0.0 ? p * vn : p;
        }
        __syncthreads();
        return;
    }
    // every member's value is loaded before any is summed (an unused member reads entry 0 and is not
    // added), so the loads of a thread are in flight together rather than one after another
    const size_t base[4] = {(size_t)(mem[0] >= 0 ? mem[0] : 0) * E, (size_t)(mem[1] >= 0 ? mem[1] : 0) * E,
                            (size_t)(mem[2] >= 0 ? mem[2] : 0) * E, (size_t)(mem[3] >= 0 ? mem[3] : 0) * E};
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
        double x[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) x[t] = stats[base[t] + 1 + d];
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (mem[t] >= 0) s += x[t];
        m[d] = (h.k0 * h.m0[d] + s) / kn;
    }
    __syncthreads();
    const double iv = 1.0 / vn;
    // entry (a, b), a >= b, reads element b * DP + a of the scatter sums: wave w takes rows
    // b = w, w + W, ... (W waves), lane l the columns a = b + l, b + l + 64, ..., so a wave's loads are
    // contiguous; the loads of 8 rows are issued before any is summed
    const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    constexpr int kRows = 8;
    for (int b0 = threadIdx.x >> 6; b0 < D; b0 += kRows * nw) {
        for (int a0 = 0; a0 < D; a0 += 64) {
            double x[kRows][4], ps[kRows];
#pragma unroll
            for (int u = 0; u < kRows; ++u) {
                const int b = b0 + u * nw, a = b + a0 + lane;
                const bool in = b < D && a < D;
                const size_t off = in ? (size_t)b * DP + a : 0;
#pragma unroll
                for (int t = 0; t < 4; ++t) x[u][t] = stats[base[t] + 1 + DP + off];
                ps[u] = h.psi0[in ? (size_t)b * D + a : 0];
            }
#pragma unroll
            for (int u = 0; u < kRows; ++u) {
                const int b = b0 + u * nw, a = b + a0 + lane;
                if (b >= D || a >= D) continue;
                double s = 0.0;
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if (mem[t] >= 0) s += x[u][t];
                const double p = (h.v0 * ps[u] + h.k0 * h.m0[b] * h.m0[a] + s - kn * m[b] * m[a]) * iv;
                P[tri(a, b)] = scale_by_nu != 0.0 ? p * vn : p;
            }
        }
    }
    __syncthreads();
}

// ---- register forms of the triangular factorisations (D <= 64 RPL, blockDim = kThreads = 4 waves) --
// Lane r of every wave holds the rows r + 64 h (h < RPL) of the matrix; each wave holds a quarter of
// the columns in its registers (16 RPL slots per row) and does that quarter of every step's updates.
// Every matrix-index loop is unrolled so every register index is static; entries above the diagonal
// that a static index bound shows are never read (row block 0 against columns >= 64) are skipped.
// A step's cost is the latency of its chain (pivot, quotient, update of the next column), not its
// arithmetic.
static_assert(kThreads == 256, "the register forms assume four waves per workgroup");

// Lane `lane` (a constant or wave-uniform index) of x, read by every lane.
__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// LDS of the register Cholesky, shared by its instantiations: two panels of four 128-entry columns,
// then the 128 pivots.
__device__ __forceinline__ double* chol_wave_lds() {
    __shared__ double buf[2 * 4 * 128 + 128];
    return buf;
}

// Cholesky in panels of four columns, panel b held by wave b % 4 (wave w holds the columns
// j = 16 (s / 4) + 4 w + s % 4 in its slots s).  The panel's wave runs its four steps on the
// panel's own columns in-wave (pivot and a_jk read across lanes with v_readlane), publishes the four
// columns as each step left them in one of two LDS panels, one barrier, then every wave applies the
// four steps to its columns right of the panel: row i forms f = a_ik / a_kk and subtracts f a_jk.
// Every entry sees the steps in order k = 0, 1, ... with the same operands and operations as in the
// block-wide LDS form below, so both give the same bits.  The subtraction also reaches entries above
// the diagonal (never read) or subtracts an exact zero (rows past D hold zeros), so it needs no mask.
// The panel after next overwrites the LDS panel a step read, after the barrier between them.  The
// columns are scaled by 1 / sqrt(a_kk) at the end (the LDS form's lazy scaling, the same product).
template <int RPL>
__device__ bool chol_packed_wave(double* P, int D) {
    constexpr int NS = 16 * RPL, NC = 64 * RPL;
    __shared__ int ok_w;
    double* lds = chol_wave_lds();
    double* piv = lds + 2 * 4 * 128;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), r = threadIdx.x & 63;
    double v[RPL][NS];
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
        const int i = r + 64 * h;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const int j = 16 * (s >> 2) + 4 * w + (s & 3);
            v[h][s] = (i < D && j <= i) ? P[i * (i + 1) / 2 + j] : 0.0;
        }
    }
    bool ok = true;
#pragma unroll
    for (int b = 0; b < NS; ++b) {
        if (4 * b < D && ok) {   // uniform in the workgroup
            const int q = b >> 2;
            double* c = lds + (b & 1) * 4 * 128;   // column t of the panel at c + 128 t
            if (w == (b & 3)) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int k = 4 * b + t;
                    if (k < D && ok) {   // uniform
                        const double akk = readlane_f64(v[k >> 6][4 * q + t], k & 63);
                        ok = akk > 0.0 && isfinite(akk);
                        if (r == 0) piv[k] = akk;
#pragma unroll
                        for (int h = 0; h < RPL; ++h) {
                            if (64 * h + 63 <= k) continue;   // rows above the pivot
                            const double f = v[h][4 * q + t] / akk;
#pragma unroll
                            for (int u = t + 1; u < 4; ++u)
                                v[h][4 * q + u] -= f * readlane_f64(v[(4 * b + u) >> 6][4 * q + t], (4 * b + u) & 63);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int h = 0; h < RPL; ++h) c[128 * t + r + 64 * h] = v[h][4 * q + t];
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int k = 4 * b + t;
                if (k < D) {   // uniform
                    const double akk = c[128 * t + k];
                    ok = ok && akk > 0.0 && isfinite(akk);   // every thread reads the same pivots
#pragma unroll
                    for (int h = 0; h < RPL; ++h) {
                        if (64 * h + 63 <= k) continue;   // rows above the pivot: nothing below the diagonal
                        const double f = c[128 * t + r + 64 * h] / akk;   // garbage after a failed pivot
#pragma unroll
                        for (int s = 4 * (q + 1); s < NS; ++s) {
                            if (64 * h + 63 < 16 * (s >> 2)) continue;   // above the diagonal
                            v[h][s] -= f * c[128 * t + 16 * (s >> 2) + 4 * w + (s & 3)];
                        }
                        if (w > (b & 3) && 64 * h + 63 >= 16 * q) {   // this wave's panel 4 q + w is right of b
#pragma unroll
                            for (int u = 0; u < 4; ++u) v[h][4 * q + u] -= f * c[128 * t + 16 * q + 4 * w + u];
                        }
                    }
                }
            }
        }
    }
    if (ok) {
#pragma unroll
        for (int h = 0; h < RPL; ++h) {
            const int i = r + 64 * h;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int j = 16 * (s >> 2) + 4 * w + (s & 3);
                if (64 * h + 63 < 16 * (s >> 2)) continue;
                if (i < D && j <= i) {
                    const double psq = sqrt(piv[j]);
                    P[i * (i + 1) / 2 + j] = j == i ? psq : v[h][s] * (1.0 / psq);
                }
            }
        }
    }
    static_assert(NC <= 128, "the LDS panels hold 128-entry columns");
    if (threadIdx.x == 0) ok_w = ok;
    __syncthreads();
    const bool res = ok_w != 0;
    __syncthreads();   // ok_w and the pivots are read before a later call writes them
    return res;
}

// X = L^-1 by forward substitution: step k subtracts (L_ik / L_kk) X_kj from the entries X_ij,
// j <= k, of every row i > k (X unscaled); every row is scaled by 1 / L_ii at the end.  Wave w holds
// the columns j = 4 s + w in its slots s; row k of a wave's columns is read from the lane holding it
// (v_readlane), so the waves never wait for one another and the only barrier is the final one.  Rows
// i <= k take a zero factor (an exact no-op).  L is not modified.
template <int RPL>
__device__ void lower_inverse_wave(const double* L, double* X, int D) {
    constexpr int NS = 16 * RPL;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), r = threadIdx.x & 63;
    double x[RPL][NS];
#pragma unroll
    for (int h = 0; h < RPL; ++h)
#pragma unroll
        for (int s = 0; s < NS; ++s) x[h][s] = 4 * s + w == r + 64 * h ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 64 * RPL; ++k) {
        if (k < D) {   // uniform
            const double lkk = L[k * (k + 1) / 2 + k];
#pragma unroll
            for (int h = 0; h < RPL; ++h) {
                if (64 * h + 63 <= k) continue;   // every row of the block is at or above k
                const int i = r + 64 * h;
                const double f = (i > k && i < D) ? L[i * (i + 1) / 2 + (i > k ? k : 0)] / lkk : 0.0;
#pragma unroll
                for (int s = 0; s < (k >> 2); ++s) x[h][s] -= f * readlane_f64(x[k >> 6][s], k & 63);
                if (w <= (k & 3)) x[h][k >> 2] -= f * readlane_f64(x[k >> 6][k >> 2], k & 63);
            }
        }
    }
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
        const int i = r + 64 * h;
        if (i < D) {
            const double inv = 1.0 / L[i * (i + 1) / 2 + i];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int j = 4 * s + w;
                if (j <= i) X[i * (i + 1) / 2 + j] = x[h][s] * inv;
            }
        }
    }
    __syncthreads();
}

// Row-parallel layout of the triangular routines below (blockDim = kThreads, D <= 128): thread t
// owns row t / LN and, within it, the columns congruent to t % LN modulo LN.
__device__ __forceinline__ int row_lanes(int D) { return D <= 64 ? 4 : 2; }

// ---- FP64 MFMA products of packed lower-triangular D x D matrices (D <= 128) --------------------
// v_mfma_f64_16x16x4f64 with lane l = 16 kp + q: the first operand is element (q, kp) of a 16 x 4
// block, the second element (kp, q) of a 4 x 16 block, and accumulator v element (4 v + kp, q) of the
// 16 x 16 product.  Entries outside the triangle or past D read as zero.  The sums run in the
// hardware's order, not the LDS forms' (the results agree to rounding).
typedef double f64x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double packed_at(const double* M, int i, int j, int D) {
    return (j <= i && i < D) ? M[i * (i + 1) / 2 + j] : 0.0;
}

// R = B A, written over A.  The ten 16 x 16 tiles (I >= J) of the lower triangle, tile (I, J)
// summing over k in [16 J, 16 I + 16), are shared among the waves with their MFMA counts balanced.
__device__ void lower_product_mfma(const double* B, double* A, int D) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kp = lane >> 4, q = lane & 15;
    // tiles as 4 I + J, 4 bits each, in the order dealt to waves w, w + 4, w + 8:
    // (3,0) (3,1) (2,0) (2,1) (0,0) (1,0) (3,2) (1,1) (2,2) (3,3)
    constexpr uint64_t kTiles = 12ull | 13ull << 4 | 8ull << 8 | 9ull << 12 | 0ull << 16 | 4ull << 20 | 14ull << 24 |
                                5ull << 28 | 10ull << 32 | 15ull << 36;
    f64x4_t acc[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int t = wv + 4 * u;
        const int tile = (int)((kTiles >> (4 * (t < 10 ? t : 0))) & 15);
        const int I = tile >> 2, J = tile & 3;
        f64x4_t c = {0.0, 0.0, 0.0, 0.0};
        if (t < 10 && 16 * I < D) {
            for (int s = 4 * J; s < 4 * I + 4; ++s) {
                const double x = packed_at(B, 16 * I + q, 4 * s + kp, D);
                const double y = packed_at(A, 4 * s + kp, 16 * J + q, D);
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
            }
        }
        acc[u] = c;
    }
    __syncthreads();   // every wave has read A
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        const int t = wv + 4 * u;
        if (t >= 10) continue;
        const int tile = (int)((kTiles >> (4 * t)) & 15);
        const int I = tile >> 2, J = tile & 3;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * I + 4 * v + kp, j = 16 * J + q;
            if (j <= i && i < D) A[i * (i + 1) / 2 + j] = acc[u][v];
        }
    }
    __syncthreads();
}

// G = J (R R^T) J: tile (I, J) of S = R R^T (I >= J) sums over k < 16 J + 16; entry (i, j) of S,
// i >= j, goes to G's entry (D-1-j, D-1-i).
__device__ void gram_rev_mfma(const double* R, double* G, int D) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kp = lane >> 4, q = lane & 15;
    // (3,3) (2,2) (3,2) (1,1) (2,1) (3,1) (0,0) (1,0) (2,0) (3,0), 4 bits each as 4 I + J
    constexpr uint64_t kTiles = 15ull | 10ull << 4 | 14ull << 8 | 5ull << 12 | 9ull << 16 | 13ull << 20 | 0ull << 24 |
                                4ull << 28 | 8ull << 32 | 12ull << 36;
    for (int t = wv; t < 10; t += 4) {
        const int tile = (int)((kTiles >> (4 * t)) & 15);
        const int I = tile >> 2, J = tile & 3;
        if (16 * I >= D) continue;
        f64x4_t c = {0.0, 0.0, 0.0, 0.0};
        for (int s = 0; s < 4 * J + 4; ++s) {
            const double x = packed_at(R, 16 * I + q, 4 * s + kp, D);
            const double y = packed_at(R, 16 * J + q, 4 * s + kp, D);
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * I + 4 * v + kp, j = 16 * J + q;
            if (j <= i && i < D) G[tri(D - 1 - j, D - 1 - i)] = c[v];
        }
    }
    __syncthreads();
}

// Tile t of the lower 16 x 16 block triangle of an 8 x 8 block matrix (64 < D <= 128), dealt by
// decreasing block distance I - J (the MFMA count of a product tile grows with it), so the waves
// taking t = w, w + 4, ... get similar shares.
__device__ __forceinline__ void lower_tile8(int t, int& I, int& J) {
    int d = 7;
    while (t >= 8 - d) {
        t -= 8 - d;
        --d;
    }
    I = t + d;
    J = t;
}

// R = B A written over A for 64 < D <= 128: the 36 lower tiles, nine per wave, held until every wave
// has read A.  Tile (I, J) sums over k in [16 J, 16 I + 16), as for D <= 64.
__device__ void lower_product_mfma8(const double* B, double* A, int D) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kp = lane >> 4, q = lane & 15;
    f64x4_t acc[9];
#pragma unroll
    for (int u = 0; u < 9; ++u) {
        int I, J;
        lower_tile8(wv + 4 * u, I, J);
        f64x4_t c = {0.0, 0.0, 0.0, 0.0};
        if (16 * I < D) {
            for (int s = 4 * J; s < 4 * I + 4; ++s) {
                const double x = packed_at(B, 16 * I + q, 4 * s + kp, D);
                const double y = packed_at(A, 4 * s + kp, 16 * J + q, D);
                c = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
            }
        }
        acc[u] = c;
    }
    __syncthreads();   // every wave has read A
#pragma unroll
    for (int u = 0; u < 9; ++u) {
        int I, J;
        lower_tile8(wv + 4 * u, I, J);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * I + 4 * v + kp, j = 16 * J + q;
            if (j <= i && i < D) A[i * (i + 1) / 2 + j] = acc[u][v];
        }
    }
    __syncthreads();
}

// G = J (R R^T) J for 64 < D <= 128, tile by tile as for D <= 64.
__device__ void gram_rev_mfma8(const double* R, double* G, int D) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kp = lane >> 4, q = lane & 15;
    for (int t = wv; t < 36; t += 4) {
        int I, J;
        lower_tile8(t, I, J);
        if (16 * I >= D) continue;
        f64x4_t c = {0.0, 0.0, 0.0, 0.0};
        for (int s = 0; s < 4 * J + 4; ++s) {
            const double x = packed_at(R, 16 * I + q, 4 * s + kp, D);
            const double y = packed_at(R, 16 * J + q, 4 * s + kp, D);
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * I + 4 * v + kp, j = 16 * J + q;
            if (j <= i && i < D) G[tri(D - 1 - j, D - 1 - i)] = c[v];
        }
    }
    __syncthreads();
}

// In-place Cholesky of packed lower P, right-looking with a lazily scaled pivot column: step k
// subtracts (a_ik / a_kk) a_jk from the trailing entries (k < j <= i) with column k as the
// earlier steps left it, while the owners of column k - 1, which nothing reads any more, scale it
// by 1 / sqrt(a_{k-1,k-1}).  One barrier per column.  Uniform result: false when a pivot is not
// positive and finite (every thread reads the same pivot).  D <= 128 takes the register form.
__device__ bool chol_packed(double* P, int D, double* /*unused*/) {
    if (D <= 64) return chol_packed_wave<1>(P, D);
    if (D <= 128) return chol_packed_wave<2>(P, D);
    const int LN = row_lanes(D), r = threadIdx.x / LN, q = threadIdx.x % LN;
    const int rb = r * (r + 1) / 2;
    const bool own = r < D;
    double psq = 0.0, pinv = 0.0;   // sqrt and reciprocal of the previous pivot
    for (int k = 0; k < D; ++k) {
        const double akk = P[tri(k, k)];
        if (!(akk > 0.0) || !isfinite(akk)) return false;
        if (k > 0 && own && q == 0 && r >= k - 1) {
            if (r == k - 1) P[rb + r] = psq;
            else P[rb + k - 1] *= pinv;
        }
        if (own && r > k) {
            const double f = P[rb + k] / akk;
            for (int j = k + 1 + q; j <= r; j += LN) P[rb + j] -= f * P[tri(j, k)];
        }
        psq = sqrt(akk);
        pinv = 1.0 / psq;
        __syncthreads();
    }
    if (own && q == 0 && r == D - 1) P[rb + r] = psq;
    __syncthreads();
    return true;
}

// X = L^-1 for packed lower L, by forward substitution with rows in parallel: step k subtracts
// (L_ik / L_kk) X_kj (j <= k) from every row i > k, row k as the earlier steps left it, while
// row k - 1, which nothing reads any more, is scaled by 1 / L_{k-1,k-1}.  One barrier per row.
__device__ void lower_inverse_packed(const double* L, double* X, int D) {
    if (D <= 64) return lower_inverse_wave<1>(L, X, D);
    if (D <= 128) return lower_inverse_wave<2>(L, X, D);
    const int LN = row_lanes(D), r = threadIdx.x / LN, q = threadIdx.x % LN;
    const int rb = r * (r + 1) / 2;
    const bool own = r < D;
    if (own)
        for (int j = q; j <= r; j += LN) X[rb + j] = j == r ? 1.0 : 0.0;
    __syncthreads();
    for (int k = 0; k < D; ++k) {
        if (own && r > k) {
            const double f = L[rb + k] / L[tri(k, k)];
            const int kb = tri(k, 0);
            for (int j = q; j <= k; j += LN) X[rb + j] -= f * X[kb + j];
        }
        if (own && r == k - 1) {
            const double inv = 1.0 / L[rb + r];
            for (int j = q; j <= r; j += LN) X[rb + j] *= inv;
        }
        __syncthreads();
    }
    if (own && r == D - 1) {
        const double inv = 1.0 / L[rb + r];
        for (int j = q; j <= r; j += LN) X[rb + j] *= inv;
    }
    __syncthreads();
}

// B A for packed lower B and A (D <= 128), written over A: every thread forms its entries
// (k from j to i, in order) in registers, then all are stored after a barrier.
__device__ void lower_product_packed(const double* B, double* A, int D) {
    if (D <= 64) return lower_product_mfma(B, A, D);
    if (D <= 128) return lower_product_mfma8(B, A, D);
    constexpr int kMaxE = (128 * 129 / 2 + kThreads - 1) / kThreads;
    const int T = D * (D + 1) / 2;
    double acc[kMaxE];
#pragma unroll
    for (int m = 0; m < kMaxE; ++m) {
        const int e = threadIdx.x + m * kThreads;
        double s = 0.0;
        if (e < T) {
            int i, j;
            tri_ij(e, i, j);
            const int ib = tri(i, 0);
            for (int k = j; k <= i; ++k) s += B[ib + k] * A[tri(k, j)];
        }
        acc[m] = s;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kMaxE; ++m) {
        const int e = threadIdx.x + m * kThreads;
        if (e < T) A[e] = acc[m];
    }
    __syncthreads();
}

// G = X^T X (lower part) for packed lower X, each entry divided by `div`.
__device__ void gram_t_packed(const double* X, double* G, int D, double div) {
    const int T = D * (D + 1) / 2;
    for (int e = threadIdx.x; e < T; e += blockDim.x) {
        int a, b;
        tri_ij(e, a, b);
        double s = 0.0;
        for (int k = a; k < D; ++k) s += X[tri(k, a)] * X[tri(k, b)];
        G[e] = div != 1.0 ? s / div : s;
    }
    __syncthreads();
}

// In place, packed lower X becomes J X^T J (J the reversal permutation): element (i, j) takes
// X[D-1-j][D-1-i].  For a symmetric X that is J X J.  The map is an involution on the packed
// triangle, so every pair is swapped by the thread owning its smaller index.
__device__ void reverse_packed(double* X, int D) {
    const int T = D * (D + 1) / 2;
    for (int e = threadIdx.x; e < T; e += blockDim.x) {
        int i, j;
        tri_ij(e, i, j);
        const int e2 = tri(D - 1 - j, D - 1 - i);
        if (e < e2) {
            const double a = X[e];
            X[e] = X[e2];
            X[e2] = a;
        }
    }
    __syncthreads();
}

// G = J (R R^T) J for packed lower R: entry (i, j) is sum_{k <= a} R[a][k] R[b][k] with
// a = D-1-i <= b = D-1-j (k in increasing order).
__device__ void gram_rev_packed(const double* R, double* G, int D) {
    if (D <= 64) return gram_rev_mfma(R, G, D);
    if (D <= 128) return gram_rev_mfma8(R, G, D);
    const int T = D * (D + 1) / 2;
    for (int e = threadIdx.x; e < T; e += blockDim.x) {
        int i, j;
        tri_ij(e, i, j);
        const int a = D - 1 - i, b = D - 1 - j;
        const int ab = tri(a, 0), bb = tri(b, 0);
        double s = 0.0;
        for (int k = 0; k <= a; ++k) s += R[ab + k] * R[bb + k];
        G[e] = s;
    }
    __syncthreads();
}

// ---- kernels -------------------------------------------------------------------------------------

// out[2 j] = log marginal likelihood, out[2 j + 1] = log det psi_n (NaN when psi_n is not SPD); with
// out_map, row out_map[j] instead of j.
// Two priors: h (the model's) and h1 (the outlier cluster's); psel[j] (NULL: 0) picks, bit 0, the prior
// the posterior is formed from and, bit 1, the prior of the log marginal likelihood's prior terms (the
// reference's log_marginal_likelihood(prior, posterior, stats) takes them separately, niw.cpp:122-138,
// and its split test forms sub-cluster posteriors from the cluster's prior,
// local_clusters_actions.cpp:221-228).
__global__ void __launch_bounds__(kThreads) k_log_marginal(const double* __restrict__ stats, int E, int DP, int D,
                                                           int prior, const int32_t* __restrict__ members,
                                                           const int32_t* __restrict__ psel, PriorArgs h,
                                                           PriorArgs h1, double* __restrict__ out,
                                                           const int32_t* __restrict__ njobs,
                                                           const int32_t* __restrict__ out_map) {
    extern __shared__ double lds[];
    __shared__ double red[8];
    __shared__ double scal[4];
    __shared__ double sh[2];
    __shared__ int mem[4];
    const int job = blockIdx.x;
    if (njobs && job >= *njobs) return;   // a grid sized for the most jobs a device-built list can hold
    if (threadIdx.x < 4) mem[threadIdx.x] = members[4 * job + threadIdx.x];
    __syncthreads();
    const int sel = psel ? psel[job] : 0;
    const PriorArgs& hp = (sel & 1) ? h1 : h;
    const PriorArgs& hm = (sel & 2) ? h1 : h;
    if (prior == DPMM_PRIOR_MULTINOMIAL) {
        // lgamma(sum a0) - lgamma(sum an) + sum_d [lgamma(an_d) - lgamma(a0_d)]
        double part = 0.0, an_part = 0.0;
        for (int d = threadIdx.x; d < D; d += blockDim.x) {
            double s = 0.0;
            for (int t = 0; t < 4; ++t)
                if (mem[t] >= 0) s += stats[(size_t)mem[t] * E + 1 + d];
            const double an = hp.alpha0[d] + s;
            an_part += an;
            part += lgamma(an) - lgamma(hm.alpha0[d]);
        }
        const double sum_an = block_sum(an_part, red);
        const double tot = block_sum(part, red);
        if (threadIdx.x == 0) {
            out[2 * job] = hm.lg_alpha0_sum - lgamma(sum_an) + tot;
            out[2 * job + 1] = 0.0;
        }
        return;
    }
    double* P = lds;
    double* m = lds + D * (D + 1) / 2;
    niw_posterior(stats, (size_t)E, DP, D, mem, hp, 0.0, P, m, scal);
    const bool ok = chol_packed(P, D, sh);
    double ld = 0.0, lg = 0.0;
    const double vn = scal[2];
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
        if (ok) ld += log(P[tri(d, d)]);
        lg += lgamma(vn / 2.0 + (double)(1 - (d + 1)) / 2.0);
    }
    ld = 2.0 * block_sum(ld, red);
    lg = block_sum(lg, red) + (double)D * (D - 1) / 4.0 * log(kPiF64);
    if (threadIdx.x == 0) {
        const double N = scal[0], kn = scal[1];
        double L;
        if (!ok) {
            L = __builtin_nan("");
            ld = __builtin_nan("");
        } else if (N == 0.0 && (sel & 1) == ((sel >> 1) & 1)) {
            L = 0.0;   // the posterior is the prior: every term cancels
            ld = hp.logdet_psi0;
        } else {
            L = -N * (double)D * 0.5 * log(kPiF64) + lg - hm.lmvg0 +
                (hm.v0 / 2.0) * (D * log(hm.v0) + hm.logdet_psi0) - (vn / 2.0) * (D * log(vn) + ld) +
                (D / 2.0) * log(hm.k0 / kn);
        }
        const size_t o = out_map ? (size_t)out_map[job] : (size_t)job;   // out_map: the job's row of out
        out[2 * o] = L;
        out[2 * o + 1] = ld;
    }
}

// Part p = 3k + r of the draw (r: 0 cluster, 1 left, 2 right): members[4p..], Philox stream
// (seed, p, iter, TAG_PARAM); output row p of `draws`
// (Gaussian: [mu (D) | R packed lower (D (D + 1) / 2) | logdet Sigma]; multinomial: log theta (D)).
// status[p] = 0 on success, 1 if a factorisation failed.  ctl (NULL: every part drawn as above) has 4
// entries per part: a row of `prev` (the previous call's draws) whose parameters the part keeps (-1:
// draw), the Philox stream's slot and iteration (-1: p and iter), and the prior the posterior is formed
// from (0: h, 1: h1).
__global__ void __launch_bounds__(kThreads) k_draw_params(const double* __restrict__ stats, int E, int DP, int D,
                                                          int prior, const int32_t* __restrict__ members,
                                                          const int32_t* __restrict__ ctl,
                                                          const double* __restrict__ prev, PriorArgs h0,
                                                          PriorArgs h1, uint32_t iter0, uint32_t key0,
                                                          uint32_t key1, int full_mu_cov, int draw_stride,
                                                          double* __restrict__ draws, int32_t* __restrict__ status) {
    extern __shared__ double lds[];
    __shared__ double red[8];
    __shared__ double scal[4];
    __shared__ double sh[2];
    __shared__ int mem[4];
    const int p = blockIdx.x;
    double* out = draws + (size_t)p * draw_stride;
    uint32_t c0 = (uint32_t)p, iter = iter0;
    int sel = 0;
    if (ctl) {
        const int src = ctl[4 * p];
        if (src >= 0) {   // the part keeps parameters drawn before
            const double* in = prev + (size_t)src * draw_stride;
            for (int e = threadIdx.x; e < draw_stride; e += blockDim.x) out[e] = in[e];
            if (threadIdx.x == 0) status[p] = 0;
            return;
        }
        if (ctl[4 * p + 1] >= 0) c0 = (uint32_t)ctl[4 * p + 1];
        if (ctl[4 * p + 2] >= 0) iter = (uint32_t)ctl[4 * p + 2];
        sel = ctl[4 * p + 3] & 1;
    }
    const PriorArgs& h = sel ? h1 : h0;
    if (threadIdx.x < 4) mem[threadIdx.x] = members[4 * p + threadIdx.x];
    __syncthreads();
    if (prior == DPMM_PRIOR_MULTINOMIAL) {
        // Dirichlet(alpha0 + sum): coordinate d's gamma from the sub-stream at word 4096 d (block
        // 1024 d), all coordinates in parallel; normalised in a fixed order, logged
        double* g = lds;
        for (int d = threadIdx.x; d < D; d += blockDim.x) {
            double s = 0.0;
            for (int t = 0; t < 4; ++t)
                if (mem[t] >= 0) s += stats[(size_t)mem[t] * E + 1 + d];
            DevStream rs(c0, iter, TAG_PARAM, key0, key1, 4096u * (uint32_t)d);
            g[d] = rs.gamma(h.alpha0[d] + s);
        }
        __syncthreads();
        if (threadIdx.x == 0) {   // the oracle's summation order
            double s = 0.0;
            for (int d = 0; d < D; ++d) s += g[d];
            sh[0] = s;
        }
        __syncthreads();
        const double s = sh[0];
        for (int d = threadIdx.x; d < D; d += blockDim.x) out[d] = log(g[d] / s);
        if (threadIdx.x == 0) status[p] = 0;
        return;
    }
    const int T = D * (D + 1) / 2;
    double* P = lds;
    double* Q = lds + T;
    double* m = lds + 2 * T;
    double* z = m + D;
    double* gd = z + D;
    // B = chol((nu psi_n)^-1).  With Cr = chol(J nu psi_n J) (J reverses the index order),
    // nu psi_n = U U^T for the upper U = J Cr J, so (nu psi_n)^-1 = U^-T U^-1 and, U^-T being lower
    // with a positive diagonal, B = U^-T = J Cr^-T J: one Cholesky and one triangular inverse.
    niw_posterior(stats, (size_t)E, DP, D, mem, h, 1.0, P, m, scal);
    const double kn = scal[1], vn = scal[2];
    reverse_packed(P, D);
    bool ok = chol_packed(P, D, sh);
    if (!ok) {
        if (threadIdx.x == 0) status[p] = 1;
        return;
    }
    lower_inverse_packed(P, Q, D);   // Cr^-1
    reverse_packed(Q, D);            // B
    // Bartlett factor A (into P): normals below the diagonal in row order (words 2m, 2m + 1 of
    // normal m), then the D normals of mu on the same stream; the D diagonal chi variates
    // sqrt(2 Gamma((nu - i) / 2)) from the streams starting at block 1024 (16 + i), all in parallel.
    for (int e = threadIdx.x; e < T; e += blockDim.x) {
        int i, j;
        tri_ij(e, i, j);
        if (j < i) P[e] = stream_normal_at(c0, iter, key0, key1, 2u * (uint32_t)(i * (i - 1) / 2 + j));
    }
    for (int i = threadIdx.x; i < D; i += blockDim.x) {
        DevStream rs(c0, iter, TAG_PARAM, key0, key1, 4096u * (16u + (uint32_t)i));
        gd[i] = sqrt(2.0 * rs.gamma((vn - i) / 2.0));
        z[i] = stream_normal_at(c0, iter, key0, key1, (uint32_t)(D * (D - 1)) + 2u * (uint32_t)i);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < D; i += blockDim.x) P[tri(i, i)] = gd[i];
    __syncthreads();
    lower_product_packed(Q, P, D);   // R = B A, into P
    double ld = 0.0;
    for (int i = threadIdx.x; i < D; i += blockDim.x) ld += log(P[tri(i, i)]);
    ld = block_sum(ld, red);
    for (int e = threadIdx.x; e < T; e += blockDim.x) out[D + e] = P[e];
    if (threadIdx.x == 0) out[D + T] = -2.0 * ld;
    // mu = m_n + chol(Sigma / kappa) z with Sigma = (R R^T)^-1 (diagonal only unless full_mu_cov, the
    // reference's pdflib.cpp:4007-4014).  With Cm = chol(J R R^T J), R R^T = U U^T for U = J Cm J, and
    // chol(Sigma / kappa) = U^-T / sqrt(kappa), whose diagonal is 1 / (Cm[D-1-i][D-1-i] sqrt(kappa)).
    if (!full_mu_cov) {
        gram_rev_packed(P, Q, D);
        ok = chol_packed(Q, D, sh);
        if (!ok) {
            if (threadIdx.x == 0) status[p] = 1;
            return;
        }
        const double rk = 1.0 / sqrt(kn);
        for (int i = threadIdx.x; i < D; i += blockDim.x)
            out[i] = m[i] + z[i] * (rk / Q[tri(D - 1 - i, D - 1 - i)]);
    } else {
        lower_inverse_packed(P, Q, D);   // R^-1
        gram_t_packed(Q, P, D, kn);      // R^-T R^-1 / kappa
        ok = chol_packed(P, D, sh);
        if (!ok) {
            if (threadIdx.x == 0) status[p] = 1;
            return;
        }
        for (int i = threadIdx.x; i < D; i += blockDim.x) {
            double x = m[i];
            for (int j = 0; j <= i; ++j) x += P[tri(i, j)] * z[j];
            out[i] = x;
        }
    }
    if (threadIdx.x == 0) status[p] = 0;
}

// Round-to-nearest-even FP32 -> BF16 bits (NaN kept quiet).
__device__ __forceinline__ uint16_t bf16_bits_rne(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// One block per part p = 3k + r: r = 0 writes cluster slot k of P (and the BF16 screen slot k of
// PB when non-null), r = 1, 2 sub-cluster slot 2k + r - 1 of SP (and of SPB when non-null).
// logw: K cluster log weights then 2K sub-cluster log weights.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_pack_gauss(const double* __restrict__ draws, int draw_stride, int D,
                                                         int DP, int DPM, const double* __restrict__ logw, int K,
                                                         T* __restrict__ P, T* __restrict__ SP,
                                                         unsigned char* __restrict__ PB,
                                                         unsigned char* __restrict__ SPB,
                                                         unsigned char* __restrict__ PBlo) {
    const int p = blockIdx.x, k = p / 3, r = p - 3 * k;
    const double* src = draws + (size_t)p * draw_stride;
    const int stride = gauss_slot_stride(DP);
    T* dst = r == 0 ? P + (size_t)k * stride : SP + (size_t)(2 * k + r - 1) * stride;
    const double lw = r == 0 ? logw[k] : logw[K + 2 * k + r - 1];
    const int TR = D * (D + 1) / 2;
    const double cst = -0.5 * src[D + TR] + lw;
    for (int e = threadIdx.x; e < stride; e += blockDim.x) dst[e] = T(0);
    // BF16 slots: a cluster's screen slot (tiles | mu | c | F) with its lo tiles in PBlo, or a
    // sub-cluster's split slot (hi tiles | lo tiles | mu | c | -c_rel | g_rel)
    uint16_t* tiles = nullptr;
    uint16_t* lo_tiles = nullptr;
    unsigned char* mslot_base = nullptr;
    int mslot = 0;
    if (r == 0 && PB) { mslot = mx_slot_bytes(DPM); mslot_base = PB + (size_t)k * mslot; }
    if (r > 0 && SPB) { mslot = mx2_slot_bytes(DPM); mslot_base = SPB + (size_t)(2 * k + r - 1) * mslot; }
    const int mu_off = (r > 0 ? 2 : 1) * mx_tiles_bytes(DPM);
    if (mslot_base) {
        tiles = reinterpret_cast<uint16_t*>(mslot_base);
        if (r > 0) lo_tiles = reinterpret_cast<uint16_t*>(mslot_base + mx_tiles_bytes(DPM));
        else if (PBlo) lo_tiles = reinterpret_cast<uint16_t*>(PBlo + (size_t)k * mx_tiles_bytes(DPM));
        for (int e = threadIdx.x; e < mslot / 2; e += blockDim.x) tiles[e] = 0;
        if (r == 0 && lo_tiles)
            for (int e = threadIdx.x; e < mx_tiles_bytes(DPM) / 2; e += blockDim.x) lo_tiles[e] = 0;
    }
    __syncthreads();
    // W = R^T: W[j][l] = R[l][j] for l >= j
    for (int e = threadIdx.x; e < TR; e += blockDim.x) {
        int l, j;
        tri_ij(e, l, j);
        const double v = src[D + e];
        dst[tri_row_offset(DP, j) - (j & ~3) + l] = (T)v;
        if (tiles) {
            const float w = (float)v;
            const uint16_t hb = bf16_bits_rne(w);
            tiles[mx_tile_index(DPM, j, l)] = hb;
            if (lo_tiles) lo_tiles[mx_tile_index(DPM, j, l)] = bf16_bits_rne(w - __uint_as_float((uint32_t)hb << 16));
        }
    }
    for (int l = threadIdx.x; l < D; l += blockDim.x) {
        dst[gauss_mu_offset(DP) + l] = (T)src[l];
        if (tiles) reinterpret_cast<float*>(mslot_base + mu_off)[l] = (float)src[l];
    }
    if (threadIdx.x == 0) {
        dst[gauss_c_offset(DP)] = (T)cst;
        if (tiles) reinterpret_cast<float*>(mslot_base + mu_off)[DPM] = (float)cst;
    }
    if (!tiles) return;
    // W[j][l] of the FP32 value, row j by thread j
    auto wf = [&](int j, int l) -> double { return (double)(float)src[D + l * (l + 1) / 2 + j]; };
    if (r == 0) {
        // F >= ||W||_F^2 (rounded up), summed in a fixed order
        __shared__ double part[kThreads];
        double f = 0.0;
        for (int j = threadIdx.x; j < D; j += blockDim.x)
            for (int l = j; l < D; ++l) f += wf(j, l) * wf(j, l);
        part[threadIdx.x] = f;
        __syncthreads();
        if (threadIdx.x == 0) {
            double F = 0.0;
            for (int t = 0; t < (int)blockDim.x; ++t) F += part[t];
            reinterpret_cast<float*>(mslot_base + mu_off)[DPM + 1] = (float)(F * (1.0 + 0x1p-20));
        }
    } else {
        // the screen relative to the parent cluster's mean (draws of part 3k): -c and g
        const double* par = draws + (size_t)(3 * k) * draw_stride;
        float* rel = reinterpret_cast<float*>(mslot_base + mx2_rel_offset(DPM));
        const double kappa = (double)sub_rel_eps_g(DPM) / (double)sub_rel_eps_a(DPM);
        __shared__ double part[2][kThreads];
        double g2 = 0.0, h2 = 0.0;
        for (int j = threadIdx.x; j < D; j += blockDim.x) {
            double c = 0.0, b = 0.0;
            for (int l = j; l < D; ++l) {
                const double w = wf(j, l);
                const double dl = (double)(float)src[l] - (double)(float)par[l];
                c += w * dl;
                b += fabs(w) * fabs(dl);
                const double wh = (double)__uint_as_float((uint32_t)bf16_bits_rne((float)w) << 16);
                h2 += wh * wh;
            }
            rel[j] = (float)(-c);
            const float g = (float)(kappa * (b + fabs(c)) * (1.0 + 0x1p-20) + 0x1p-120);
            rel[DPM + j] = g;
            g2 += (double)g * (double)g;
        }
        part[0][threadIdx.x] = g2;
        part[1][threadIdx.x] = h2;
        __syncthreads();
        if (threadIdx.x == 0) {   // ||g|| and ||W_hi||_F, rounded up, in a fixed order
            double G = 0.0, H = 0.0;
            for (int t = 0; t < (int)blockDim.x; ++t) {
                G += part[0][t];
                H += part[1][t];
            }
            float* m = reinterpret_cast<float*>(mslot_base + mu_off);
            m[DPM + 1] = (float)(sqrt(G) * (1.0 + 0x1p-20));
            m[DPM + 2] = (float)(sqrt(H) * (1.0 + 0x1p-20));
        }
    }
}

// Multinomial slots: [log theta (D) | log weight | pad] in FP32; clusters carry log pi_k only when
// use_weight (the reference's label score leaves it out, cudaKernel_multinomial.cu:60-76).
__global__ void __launch_bounds__(kThreads) k_pack_mult(const double* __restrict__ draws, int draw_stride, int D,
                                                        int stride, const double* __restrict__ logw, int K,
                                                        int use_weight, float* __restrict__ P,
                                                        float* __restrict__ SP) {
    const int p = blockIdx.x, k = p / 3, r = p - 3 * k;
    const double* src = draws + (size_t)p * draw_stride;
    float* dst = r == 0 ? P + (size_t)k * stride : SP + (size_t)(2 * k + r - 1) * stride;
    for (int e = threadIdx.x; e < stride; e += blockDim.x) {
        float v = 0.f;
        if (e < D) v = (float)src[e];
        else if (e == D) v = r == 0 ? (use_weight ? (float)logw[k] : 0.f) : (float)logw[K + 2 * k + r - 1];
        dst[e] = v;
    }
}

// ---- screener pair features ----------------------------------------------------------------------
// Cluster k (entries 2k, 2k + 1): posterior mean m (D) and diagonal of psi_n (D), into out[k][0..2D).
__global__ void __launch_bounds__(kThreads) k_cluster_summary(const double* __restrict__ stats, int E, int DP, int D,
                                                              PriorArgs h, double* __restrict__ out) {
    const int k = blockIdx.x;
    const double* s0 = stats + (size_t)(2 * k) * E;
    const double* s1 = s0 + E;
    const double N = s0[0] + s1[0];
    const double kn = h.k0 + N, vn = h.v0 + N;
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
        double m, dp;
        if (N == 0.0) {
            m = h.m0[d];
            dp = h.psi0[(size_t)d * D + d];
        } else {
            const double sum = s0[1 + d] + s1[1 + d];
            const double S = s0[1 + DP + (size_t)d * DP + d] + s1[1 + DP + (size_t)d * DP + d];
            m = (h.k0 * h.m0[d] + sum) / kn;
            dp = (h.v0 * h.psi0[(size_t)d * D + d] + h.k0 * h.m0[d] * h.m0[d] + S - kn * m * m) * (1.0 / vn);
        }
        out[(size_t)k * 2 * D + d] = m;
        out[(size_t)k * 2 * D + D + d] = dp;
    }
}

// One wave per pair (i, j): the 16 features of the merge screener from the summaries and the
// per-cluster scalars (N, Nl, Nr, L, Ll, Lr, log det psi_n).
__global__ void __launch_bounds__(kThreads) k_pair_features(const double* __restrict__ summ,
                                                            const double* __restrict__ scal, int D, int npairs,
                                                            const int32_t* __restrict__ pi,
                                                            const int32_t* __restrict__ pj, float* __restrict__ out) {
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (p >= npairs) return;
    const int i = pi[p], j = pj[p];
    const double* mi = summ + (size_t)i * 2 * D;
    const double* mj = summ + (size_t)j * 2 * D;
    double di = 0.0, dj = 0.0, dq = 0.0;
    for (int a = lane; a < D; a += 64) {
        const double d = mi[a] - mj[a];
        const double si = mi[D + a], sj = mj[D + a];
        di += d * d / si;
        dj += d * d / sj;
        dq += d * d / (si + sj);
    }
    for (int o = 32; o > 0; o >>= 1) {
        di += __shfl_xor(di, o);
        dj += __shfl_xor(dj, o);
        dq += __shfl_xor(dq, o);
    }
    if (lane < 16) {
        const double* a = scal + (size_t)i * 7;
        const double* b = scal + (size_t)j * 7;
        double f;
        switch (lane) {
            case 0: f = log(a[0]); break;
            case 1: f = log(b[0]); break;
            case 2: f = log(a[0] + b[0]); break;
            case 3: f = a[3] / (a[0] * D); break;
            case 4: f = b[3] / (b[0] * D); break;
            case 5: f = a[6] / D; break;
            case 6: f = b[6] / D; break;
            case 7: f = log1p(di / D); break;
            case 8: f = log1p(dj / D); break;
            case 9: f = log1p(dq / D); break;
            case 10: f = (a[4] + a[5] - a[3]) / (a[0] * D); break;
            case 11: f = (b[4] + b[5] - b[3]) / (b[0] * D); break;
            case 12: f = fabs(log(a[1] + 1.0) - log(a[2] + 1.0)); break;
            case 13: f = fabs(log(b[1] + 1.0) - log(b[2] + 1.0)); break;
            case 14: f = fabs(a[6] / D - b[6] / D); break;
            default: f = fabs(log(a[0]) - log(b[0])); break;
        }
        out[(size_t)p * 16 + lane] = (float)f;
    }
}

}  // namespace dpmm
