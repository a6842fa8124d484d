// Kernels of the device-resident sampler iteration: the scalar work of group_step
// (local_clusters_actions.cpp:460-490) that otherwise runs on the host between device launches.
//   k_step_sample    sample_clusters (local_clusters_actions.cpp:341-371): sub-cluster weights
//                    ~ Dir(N_l + alpha / 2, N_r + alpha / 2), the burn-out history and splittable rule
//                    (shared_actions.cpp:39-54), mixture weights ~ Dir(N_1, ..., N_K, alpha) with the
//                    last entry dropped (357-365); the rows and log weights of the parameter draw.
//   k_step_jobs      each part's count and statistics entries (cluster = l + r, niw.cpp:140-148),
//                    the merge candidates and the log-ML job list: every part, every candidate pair.
//   k_step_decide    reset_bad_clusters (429-447), check_and_split (244-285) within the cluster cap,
//                    check_and_merge (287-321) + should_merge (shared_actions.cpp:88-111) with the
//                    greedy lexicographic resolution, remove_empty_clusters (400-419): the new cluster
//                    records and the two relabel tables.
//   k_step_relabel   one relabel pass: resets and splits, or merges and removals.
//   k_step_compact   the cluster table renumbered after merges and removals.
//   k_step_touched_* statistics entries and log ML of the clusters a reset or a split touched.
// Every cluster loop runs in one workgroup (K <= 2048).  The decisions read the Philox counters the
// host driver reads and evaluate the same expressions in the same order with contraction off, so the
// two take the same decisions (lgamma, log and the gamma draws' transcendentals may differ from the
// host library's in the last bit).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpmm_niw.h"
#include "dpmm_rng.h"
#include "dpmm_step_state.h"

namespace dpmm {

constexpr int kStepThreads = 1024;
constexpr int kStepMaxK = 2048;

// Rank of this thread's flag among the set flags of the block, in thread order, and their total.
// Every thread of the block calls it.
__device__ __forceinline__ int step_rank(bool flag, int* ws, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long b = __ballot(flag);
    const int below = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) ws[w] = __popcll(b);
    __syncthreads();
    int off = 0, tot = 0;
    const int nw = (int)(blockDim.x >> 6);
    for (int i = 0; i < nw; ++i) {
        if (i < w) off += ws[i];
        tot += ws[i];
    }
    __syncthreads();
    total = tot;
    return off + below;
}

__device__ __forceinline__ void step_set_part(StepPart& p, int64_t N, int q0, int q1) {
    p.N = N;
    p.mem[0] = q0;
    p.mem[1] = q1;
    p.mem[2] = p.mem[3] = -1;
}

// Pair p of n candidates in lexicographic order: (ra, rb), ra < rb; row ra starts at ra n - ra (ra + 1) / 2.
__device__ __forceinline__ void step_pair(int64_t p, int n, int& ra, int& rb) {
    const double b = 2.0 * n - 1.0;
    int r = (int)((b - sqrt(fmax(b * b - 8.0 * (double)p, 0.0))) * 0.5);
    r = max(0, min(r, n - 2));
    auto row0 = [n](int q) { return (int64_t)q * n - (int64_t)q * (q + 1) / 2; };
    while (r > 0 && row0(r) > p) --r;
    while (r + 1 < n - 1 && row0(r + 1) <= p) ++r;
    ra = r;
    rb = (int)(p - row0(r)) + r + 1;
}

// Two-component Dirichlet on stream (c0, t, tag): the gammas in order, each divided by their sum.
__device__ __noinline__ void step_dirichlet2(uint32_t c0, uint32_t t, uint32_t tag, uint32_t k0, uint32_t k1,
                                                double a0, double a1, double* out) {
#pragma clang fp contract(off)
    DevStream rs(c0, t, tag, k0, k1, 0);
    const double g0 = rs.gamma(a0);
    const double g1 = rs.gamma(a1);
    const double s = g0 + g1;
    out[0] = g0 / s;
    out[1] = g1 / s;
}

// sample_clusters for clusters 0 .. K-1.  members: 3K rows of 4 statistics entries (cluster, l, r of
// each cluster); logw: log mixture weight of each cluster, then the 2K log sub-cluster weights.
// Mixture component i (clusters, then alpha) draws its gamma from stream (i, t, TAG_MIXW).
__global__ void __launch_bounds__(kStepThreads) k_step_sample(StepArgs a, StepCluster* __restrict__ cl,
                                                             double* __restrict__ hist,
                                                             double* __restrict__ weights,
                                                             int32_t* __restrict__ members,
                                                             double* __restrict__ logw) {
#pragma clang fp contract(off)
    __shared__ double g[kStepMaxK + 1];
    __shared__ double s_sum;
    const int K = a.K, B = a.B;
    // tasks 0 .. K-1: the clusters' sub-cluster weights and histories; K .. 2K: the mixture's gammas
    for (int task = threadIdx.x; task <= 2 * K; task += blockDim.x) {
        if (task >= K) {
            const int i = task - K;
            DevStream rs((uint32_t)i, a.t, TAG_MIXW, a.key0, a.key1, 0);
            g[i] = rs.gamma(i < K ? (double)cl[i].c.N : a.alpha);
            continue;
        }
        const int k = task;
        StepCluster& c = cl[k];
        step_dirichlet2((uint32_t)k, a.t, TAG_LRW, a.key0, a.key1, (double)c.l.N + a.alpha / 2,
                        (double)c.r.N + a.alpha / 2, c.lrw);
        // burn-out history: push L_l + L_r; splittable once B entries average within 1e-2 of the last
        const double L = c.l.lml + c.r.lml;
        double* h = hist + (size_t)k * B;
        h[c.hist_head] = L;
        c.hist_head = c.hist_head + 1 == B ? 0 : c.hist_head + 1;
        if (c.hist_n < B) ++c.hist_n;
        if (c.hist_n == B) {
            const double w = 1 / (B - 0.1);
            double now = 0.0;
            for (int j = 0; j < B; ++j) {   // oldest to newest: the ring is full, the oldest at the head
                int q = c.hist_head + j;
                if (q >= B) q -= B;
                now += h[q] * w;
            }
            if (now - L < 1e-2) c.splittable = 1;
        }
        const StepPart* ps[3] = {&c.c, &c.l, &c.r};
        for (int p = 0; p < 3; ++p)
            for (int e = 0; e < 4; ++e) members[4 * (3 * k + p) + e] = ps[p]->mem[e];
        logw[K + 2 * k] = log(c.lrw[0]);
        logw[K + 2 * k + 1] = log(c.lrw[1]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i <= K; ++i) s += g[i];
        s_sum = s;
    }
    __syncthreads();
    // get_dirichlet_distribution (local_clusters_actions.cpp:323-339): one empty cluster gets weight 0
    const bool degenerate = K == 1 && cl[0].c.N == 0;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const double w = degenerate ? 0.0 : g[k] / s_sum;
        weights[k] = w;
        logw[k] = log(w);
    }
}

// The parts of clusters 0 .. K-1 from the statistics pass (entry q = 2k + s holds sub-cluster s of
// cluster k; its first double is the count), the merge candidates (splittable, non-empty) and the
// log-ML jobs: 3k + {0, 1, 2} the parts of cluster k, then 3K + p the candidate pairs in
// lexicographic order.  counts[0] = jobs, counts[1] = candidates.  With the merge bound the pairs are
// listed later (k_step_pair_bound) and jobs 3K + k are the bound's W_k, with their prior selectors
// (psel) and output rows (omap).
__global__ void __launch_bounds__(kStepThreads) k_step_jobs(StepArgs a, StepCluster* __restrict__ cl,
                                                           const double* __restrict__ stats,
                                                           int32_t* __restrict__ cand, int32_t* __restrict__ jobs,
                                                           int32_t* __restrict__ counts, int32_t* __restrict__ psel,
                                                           int32_t* __restrict__ omap, int wrow) {
    __shared__ int ws[16];
    const int K = a.K;
    int ncand = 0;
    for (int base = 0; base < K; base += blockDim.x) {
        const int k = base + threadIdx.x;
        bool f = false;
        if (k < K) {
            StepCluster& c = cl[k];
            const int64_t Nl = (int64_t)stats[(size_t)(2 * k) * a.E];
            const int64_t Nr = (int64_t)stats[(size_t)(2 * k + 1) * a.E];
            step_set_part(c.l, Nl, 2 * k, -1);
            step_set_part(c.r, Nr, 2 * k + 1, -1);
            step_set_part(c.c, Nl + Nr, 2 * k, 2 * k + 1);
            const StepPart* ps[3] = {&c.c, &c.l, &c.r};
            for (int p = 0; p < 3; ++p)
                for (int e = 0; e < 4; ++e) jobs[4 * (3 * k + p) + e] = ps[p]->mem[e];
            if (a.merge_bound) {
                // the merge bound's W_k (job 3K + k: both sub-clusters under the prior h1 of the launch,
                // result in row wrow + k) beside the parts (rows 3k + p, prior h)
                int32_t* w = jobs + 4 * (3 * K + k);
                w[0] = 2 * k;
                w[1] = 2 * k + 1;
                w[2] = -1;
                w[3] = -1;
                for (int p = 0; p < 3; ++p) {
                    psel[3 * k + p] = 0;
                    omap[3 * k + p] = 3 * k + p;
                }
                psel[3 * K + k] = 3;
                omap[3 * K + k] = wrow + k;
            }
            f = !a.no_more_splits && c.splittable && Nl + Nr > 0;
        }
        int tot;
        const int r = step_rank(f, ws, tot);
        if (f) cand[ncand + r] = k;
        ncand += tot;
    }
    __syncthreads();
    // with the merge bound, k_step_pair_bound lists the pairs that need their log ML
    for (int ra = 0; ra + 1 < ncand && !a.merge_bound; ++ra) {
        const int64_t row0 = (int64_t)ra * ncand - (int64_t)ra * (ra + 1) / 2;
        const int i = cand[ra];
        for (int rb = ra + 1 + (int)threadIdx.x; rb < ncand; rb += blockDim.x) {
            const int j = cand[rb];
            int32_t* m = jobs + 4 * ((int64_t)3 * K + row0 + (rb - ra - 1));
            m[0] = 2 * i;
            m[1] = 2 * i + 1;
            m[2] = 2 * j;
            m[3] = 2 * j + 1;
        }
    }
    if (threadIdx.x == 0) {
        counts[0] = 3 * K + (int)((int64_t)ncand * (ncand - 1) / 2);
        counts[1] = ncand;
        counts[2] = 0;   // pairs listed by the merge bound
        counts[3] = 0;   // pairs it rejected
    }
}

// The merge bound.  For the NIW posterior, Psi_n = v0 psi0 + S + P (the unnormalised scatter, v_n
// times the psi the reference keeps), where S is the scatter of the points about their mean xbar and
// P = (k0 N / kappa_n) (xbar - m0)(xbar - m0)^T.  Merging clusters i and j gives
//   Psi_m = W_i + W_j + c d d^T + P_m,   W_k = v0 psi0 / 2 + S_k,  d = xbar_i - xbar_j,
//   c = N_i N_j / (N_i + N_j),
// with P_m positive semidefinite, so by Minkowski's determinant inequality and the rank-one update
// (log det (A + c d d^T) = log det A + log(1 + c d^T A^-1 d) >= log det A + log(1 + c |d|^2 / tr A)),
//   log det Psi_m >= D log(det W_i^(1/D) + det W_j^(1/D)) + log(1 + c |d|^2 / tr(W_i + W_j)).
// The log ML (niw.cpp:122-138) decreases in log det Psi_m, so its value at the bound bounds the merged
// log ML from above, and with it the merge's Hastings ratio (should_merge, shared_actions.cpp:88-111).
// A pair whose bounded ratio is below -100 is rejected whatever its uniform (the test accepts above
// log u >= log 2^-33, or log 0.1 in the final iterations), which is what its exact test would decide,
// so its log ML is not computed.  log det W_k comes from one factorisation per cluster (the log-ML
// kernel under a prior with k0 = 0 and v0 / 2, whose psi_n is W_k / v_n).
//
// Per cluster k (one wave each): the data mean (vec, D doubles) and, in sc (6 doubles): an upper bound
// of tr S_k (the cancellation of sum x^2 - N |xbar|^2 covered), log det W_k, N_k, the cluster's log ML
// and lgamma(N_k + alpha / 2) - lgamma(N_k).
__global__ void __launch_bounds__(64) k_step_bound_prep(StepArgs a, const double* __restrict__ stats,
                                                        const double* __restrict__ lm,
                                                        const double* __restrict__ lw, PriorArgs h, int D, int DP,
                                                        double* __restrict__ vec, double* __restrict__ sc) {
    const int k = blockIdx.x, lane = threadIdx.x;
    const double* e0 = stats + (size_t)(2 * k) * a.E;
    const double* e1 = stats + (size_t)(2 * k + 1) * a.E;
    const double N = e0[0] + e1[0];
    double m2 = 0.0, trq = 0.0;
    for (int d = lane; d < D; d += 64) {
        const double xb = N > 0.0 ? (e0[1 + d] + e1[1 + d]) / N : 0.0;
        vec[(size_t)k * D + d] = xb;
        m2 += xb * xb;
        trq += e0[1 + DP + (size_t)d * DP + d] + e1[1 + DP + (size_t)d * DP + d];
    }
    for (int o = 32; o > 0; o >>= 1) {
        m2 += __shfl_xor(m2, o);
        trq += __shfl_xor(trq, o);
    }
    if (lane == 0) {
        sc[6 * k] = (trq - N * m2) + 1e-12 * (trq + N * m2);
        sc[6 * k + 1] = (double)D * log(0.5 * h.v0 + N) + lw[2 * k + 1];
        sc[6 * k + 2] = N;
        sc[6 * k + 3] = lm[2 * (3 * k)];
        sc[6 * k + 4] = N > 0.0 ? lgamma(N + 0.5 * a.alpha) - lgamma(N) : 0.0;   // the ratio's terms of k
    }
}

// One wave per candidate pair p (lexicographic over the candidates): the merge bound, then either the
// pair's log-ML job listed (jobs row q, out_map[q] = 3K + p) or its log ML set to -inf.  trpsi0 =
// v0 tr psi0.
__global__ void __launch_bounds__(kThreads) k_step_pair_bound(StepArgs a, const int32_t* __restrict__ cand,
                                                              int32_t* __restrict__ counts,
                                                              const double* __restrict__ vec,
                                                              const double* __restrict__ sc, PriorArgs h, int D,
                                                              double trpsi0, int32_t* __restrict__ jobs,
                                                              int32_t* __restrict__ out_map, double* __restrict__ lm) {
    const int lane = threadIdx.x & 63;
    const int64_t p = (int64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
    const int K = a.K, ncand = counts[1];
    const int64_t np = (int64_t)ncand * (ncand - 1) / 2;
    if (p >= np) return;
    int ra, rb;
    step_pair(p, ncand, ra, rb);
    const int i = cand[ra], j = cand[rb];
    const double* vi = vec + (size_t)i * D;
    const double* vj = vec + (size_t)j * D;
    const double Ni = sc[6 * i + 2], Nj = sc[6 * j + 2], Nm = Ni + Nj, vm = h.v0 + Nm;
    double d2 = 0.0, lg = 0.0;
    for (int d = lane; d < D; d += 64) {
        const double t = vi[d] - vj[d];
        d2 += t * t;
        // log Gamma_D(v_m / 2) from above, term by term: lgamma(x) < (x - 1/2) log x - x + log(2 pi) / 2 +
        // 1 / (12 x) for x > 0 (Stirling's series with Robbins' remainder bound), one log instead of lgamma
        const double x = vm / 2.0 + (double)(1 - (d + 1)) / 2.0;
        lg += (x - 0.5) * log(x) - x + 0.91893853320467274 + 1.0 / (12.0 * x);
    }
    for (int o = 32; o > 0; o >>= 1) {
        d2 += __shfl_xor(d2, o);
        lg += __shfl_xor(lg, o);
    }
    if (lane != 0) return;
    const double tri = sc[6 * i], trj = sc[6 * j], lwi = sc[6 * i + 1], lwj = sc[6 * j + 1];
    const double lmi = sc[6 * i + 3], lmj = sc[6 * j + 3];
    const double trA = trpsi0 + tri + trj;
    bool reject = false;
    if (Ni > 0.0 && Nj > 0.0 && trA > 0.0 && isfinite(lwi) && isfinite(lwj) && isfinite(lmi) && isfinite(lmj)) {
        const double hi = fmax(lwi, lwj) / D, lo = fmin(lwi, lwj) / D;
        const double lb = (double)D * (hi + log1p(exp(lo - hi))) + log1p(Ni * Nj / Nm * d2 / trA);
        const double Lub = -Nm * (double)D * 0.5 * log(kPiF64) + lg + (double)D * (D - 1) / 4.0 * log(kPiF64) -
                           h.lmvg0 + (h.v0 / 2.0) * (D * log(h.v0) + h.logdet_psi0) - (vm / 2.0) * lb +
                           (D / 2.0) * log(h.k0 / (h.k0 + Nm));
        const double al = a.alpha;
        const double ub = -log(al) + lgamma(al) - 2 * lgamma(0.5 * al) + lgamma(Nm) - lgamma(Nm + al) +
                          sc[6 * i + 4] + sc[6 * j + 4] + Lub - lmi - lmj;
        reject = ub < -100.0;
    }
    if (reject) {
        lm[2 * (3 * (int64_t)K + p)] = -__builtin_huge_val();
        lm[2 * (3 * (int64_t)K + p) + 1] = __builtin_nan("");
        atomicAdd(&counts[3], 1);
    } else {
        const int q = atomicAdd(&counts[2], 1);
        int32_t* m = jobs + 4 * (int64_t)q;
        m[0] = 2 * i;
        m[1] = 2 * i + 1;
        m[2] = 2 * j;
        m[3] = 2 * j + 1;
        out_map[q] = 3 * K + (int32_t)p;
    }
}

// split_cluster_local (local_clusters_actions.cpp:183-194): k keeps the left half, nk gets the right
// half; each half is create_splittable_from_params of the old sub-cluster (shared_actions.cpp:57-70)
// with sub-cluster weights ~ Dir(alpha / 2, alpha / 2), its parts recomputed from its new sub-labels.
__device__ __noinline__ void step_split(const StepArgs& a, StepCluster* __restrict__ cl, int k, int nk) {
    StepCluster& c = cl[k];
    StepCluster& R = cl[nk];
    const StepPart pl = c.l, pr = c.r;
    R.c = R.l = R.r = pr;
    c.c = c.l = c.r = pl;
    R.splittable = c.splittable = 0;
    R.hist_n = c.hist_n = 0;
    R.hist_head = c.hist_head = 0;
    R.touched = c.touched = 1;
    step_dirichlet2((uint32_t)k, a.t, TAG_SPLIT_LRW, a.key0, a.key1, a.alpha / 2, a.alpha / 2, c.lrw);
    step_dirichlet2((uint32_t)nk, a.t, TAG_SPLIT_LRW, a.key0, a.key1, a.alpha / 2, a.alpha / 2, R.lrw);
}

// merge_clusters + merge_clusters_to_splittable (local_clusters_actions.cpp:204-213, shared_actions.cpp:
// 72-86): i becomes the merged cluster with log ML L, its sub-clusters the old i (a clone of the merged
// part, as the reference leaves it) and j, sub-cluster weights ~ Dir(N_i + alpha / 2, N_j + alpha / 2);
// j is emptied.
__device__ __noinline__ void step_merge(const StepArgs& a, StepCluster* __restrict__ cl, int i, int j, double L) {
    StepCluster& A = cl[i];
    StepCluster& Bc = cl[j];
    StepPart m;
    m.N = A.c.N + Bc.c.N;
    m.mem[0] = 2 * i;
    m.mem[1] = 2 * i + 1;
    m.mem[2] = 2 * j;
    m.mem[3] = 2 * j + 1;
    m.lml = L;
    m.ld = 0.0;
    step_dirichlet2((uint32_t)i, a.t, TAG_MERGE_LRW, a.key0, a.key1, (double)A.c.N + 0.5 * a.alpha,
                    (double)Bc.c.N + 0.5 * a.alpha, A.lrw);
    A.r = Bc.c;
    A.c = m;
    A.l = m;
    A.splittable = 0;
    A.hist_n = 0;
    A.hist_head = 0;
    Bc.c.N = 0;
    Bc.splittable = 0;
}

// Decisions of the iteration over the jobs' log MLs lm (2 doubles per job: log ML, log det psi_n).
// tab1: the reset + split relabel table over the K clusters (new label of (k, s) at 2k + s, action of
// cluster k at 2K + k: 1 reset, 2 split); tab2: the merge + removal table over the K_split clusters
// (actions 3 merged into, 4 merged away); aux: merge target (K_split) then new index (K_split) of
// every cluster; accl: scratch for the accepted pairs.  draw_status: the parameter draw's per-part
// status (n_status entries).
__global__ void __launch_bounds__(kStepThreads) k_step_decide(StepArgs a, StepCluster* __restrict__ cl,
                                                             const double* __restrict__ lm,
                                                             const int32_t* __restrict__ cand,
                                                             const int32_t* __restrict__ counts,
                                                             const int32_t* __restrict__ draw_status, int n_status,
                                                             int32_t* __restrict__ accl, int32_t* __restrict__ tab1,
                                                             int32_t* __restrict__ tab2, int32_t* __restrict__ aux,
                                                             StepSummary* __restrict__ sum) {
#pragma clang fp contract(off)
    __shared__ int ws[16];
    __shared__ int s_err, s_nreset, s_nmerge;
    __shared__ unsigned char used[kStepMaxK];
    __shared__ double lgn[kStepMaxK], lgnh[kStepMaxK];
    const int K = a.K;
    if (threadIdx.x == 0) {
        s_err = 0;
        s_nreset = 0;
        s_nmerge = 0;
    }
    for (int k = threadIdx.x; k < kStepMaxK; k += blockDim.x) used[k] = 0;
    __syncthreads();
    int err = 0, nreset = 0;
    for (int p = threadIdx.x; p < n_status; p += blockDim.x)
        if (draw_status[p]) err |= 1;
    // the parts' log MLs; reset_bad_clusters: a cluster with an empty sub-cluster gets new sub-labels,
    // an empty history and is not splittable
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        StepCluster& c = cl[k];
        StepPart* ps[3] = {&c.c, &c.l, &c.r};
        for (int p = 0; p < 3; ++p) {
            const double L = lm[2 * (3 * k + p)];
            if (!isfinite(L)) err |= 2;
            ps[p]->lml = L;
            ps[p]->ld = lm[2 * (3 * k + p) + 1];
        }
        c.touched = 0;
        tab1[2 * k] = k;
        tab1[2 * k + 1] = k;
        tab1[2 * K + k] = 0;
        if (c.l.N == 0 || c.r.N == 0) {
            c.hist_n = 0;
            c.hist_head = 0;
            c.splittable = 0;
            c.touched = 1;
            tab1[2 * K + k] = 1;
            ++nreset;
        }
    }
    if (err) atomicOr(&s_err, err);
    if (nreset) atomicAdd(&s_nreset, nreset);
    __syncthreads();
    // check_and_split: log H = log alpha + lgamma(N_l) + L_l + lgamma(N_r) + L_r - lgamma(N) - L against
    // log U; accepted splits beyond the cluster cap are dropped, the highest indices first
    int nsplit = 0;
    if (!a.no_more_splits) {
        const int room = a.kcap - K;
        for (int base = 0; base < K; base += blockDim.x) {
            const int k = base + threadIdx.x;
            bool acc = false;
            if (k < K) {
                const StepCluster& c = cl[k];
                if (c.splittable && c.c.N > 1 && !a.fin && c.l.N != 0 && c.r.N != 0) {
                    const double lhr = log(a.alpha) + lgamma((double)c.l.N) + c.l.lml + lgamma((double)c.r.N) +
                                       c.r.lml - (lgamma((double)c.c.N) + c.c.lml);
                    const double u =
                        u01_f64(philox4x32_10((uint32_t)k, 0, a.t, TAG_SPLIT_U, a.key0, a.key1).v[0]);
                    acc = lhr > log(u);
                }
            }
            int tot;
            const int r = step_rank(acc, ws, tot);
            if (acc && nsplit + r < room) {
                const int nk = K + nsplit + r;
                step_split(a, cl, k, nk);
                tab1[2 * k + 1] = nk;
                tab1[2 * K + k] = 2;
            }
            nsplit = min(nsplit + tot, max(room, 0));
        }
    }
    __syncthreads();
    const int Ks = K + nsplit;
    int32_t* merge_into = aux;
    int32_t* remap = aux + Ks;
    for (int k = threadIdx.x; k < Ks; k += blockDim.x) merge_into[k] = -1;
    __syncthreads();
    // check_and_merge: every candidate pair on the statistics before the splits; accepted pairs are
    // resolved greedily in lexicographic order, a cluster merging at most once (a merged cluster is
    // not splittable), which gives the reference's sequential loop's outcome
    if (!a.no_more_splits) {
        const int ncand = counts[1];
        const int64_t np = (int64_t)ncand * (ncand - 1) / 2;
        const double al = a.alpha;
        const double cst = -log(al) + lgamma(al) - 2 * lgamma(0.5 * al);
        // per-cluster terms of the ratio, once (the candidates' counts are those before the splits)
        for (int k = threadIdx.x; k < K; k += blockDim.x) {
            const double Nk = (double)cl[k].c.N;
            lgn[k] = cl[k].c.N > 0 ? lgamma(Nk) : 0.0;
            lgnh[k] = lgamma(Nk + 0.5 * al);
        }
        __syncthreads();
        int nacc = 0;
        for (int64_t base = 0; base < np; base += blockDim.x) {
            const int64_t p = base + threadIdx.x;
            bool acc = false;
            if (p < np) {
                int ra, rb;
                step_pair(p, ncand, ra, rb);
                const int i = cand[ra], j = cand[rb];
                const StepCluster& ci = cl[i];
                const StepCluster& cj = cl[j];
                const double L = lm[2 * (3 * (int64_t)K + p)];
                if (isfinite(L) && ci.splittable && cj.splittable && ci.c.N > 0 && cj.c.N > 0) {
                    // lgamma(Nm) - lgamma(Nm + alpha) < 0 for Nm >= 2: without it the ratio is an upper
                    // bound, and a pair far below log(2^-32) and log(0.1) is rejected whatever its uniform
                    const double ub = cst + lgnh[i] - lgn[i] - lgn[j] + lgnh[j] + L - ci.c.lml - cj.c.lml;
                    if (!(ub < -100.0)) {
                        const double Nm = (double)ci.c.N + (double)cj.c.N;
                        const double lhr = cst + lgamma(Nm) - lgamma(Nm + al) + lgnh[i] - lgn[i] - lgn[j] + lgnh[j] +
                                           L - ci.c.lml - cj.c.lml;
                        const double u = u01_f64(philox4x32_10(((uint32_t)i << 16) | (uint32_t)j, 0, a.t, TAG_MERGE_U,
                                                               a.key0, a.key1).v[0]);
                        acc = lhr > log(u) || (a.fin && lhr > log(0.1));
                    }
                }
            }
            int tot;
            const int r = step_rank(acc, ws, tot);
            if (acc) accl[nacc + r] = (int32_t)p;
            nacc += tot;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int nm = 0;
            for (int q = 0; q < nacc; ++q) {
                const int64_t p = accl[q];
                int ra, rb;
                step_pair(p, ncand, ra, rb);
                const int i = cand[ra], j = cand[rb];
                if (used[i] || used[j]) continue;
                used[i] = used[j] = 1;
                merge_into[j] = i;
                ++nm;
                step_merge(a, cl, i, j, lm[2 * (3 * (int64_t)K + p)]);
            }
            s_nmerge = nm;
        }
    }
    __syncthreads();
    // remove_empty_clusters: the clusters with points keep their order
    int nk = 0;
    for (int base = 0; base < Ks; base += blockDim.x) {
        const int k = base + threadIdx.x;
        const bool keep = k < Ks && cl[k].c.N > 0;
        int tot;
        const int r = step_rank(keep, ws, tot);
        if (k < Ks) remap[k] = keep ? nk + r : -1;
        nk += tot;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < Ks; k += blockDim.x) {
        const int tgt = merge_into[k] >= 0 ? merge_into[k] : k;
        const int lab = remap[tgt] >= 0 ? remap[tgt] : 0;
        tab2[2 * k] = lab;
        tab2[2 * k + 1] = lab;
        tab2[2 * Ks + k] = merge_into[k] >= 0 ? 4 : 0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < Ks; k += blockDim.x)
        if (merge_into[k] >= 0) tab2[2 * Ks + merge_into[k]] = 3;
    if (threadIdx.x == 0) {
        sum->K_split = Ks;
        sum->K_new = nk;
        sum->nsplit = nsplit;
        sum->nreset = s_nreset;
        sum->nmerge = s_nmerge;
        sum->followup = nsplit + s_nreset > 0;
        sum->relabel2 = s_nmerge > 0 || nk != Ks;
        sum->error = s_err;
        sum->npairs = a.no_more_splits ? 0 : a.merge_bound ? counts[2] : counts[1] * (counts[1] - 1) / 2;
        sum->nbound = a.no_more_splits || !a.merge_bound ? 0 : counts[3];
    }
}

// One relabel pass (the table layout of k_step_decide).  mode 1: the reset + split table over the
// K_arg clusters, when resets or splits were accepted; mode 2: the merge + removal table over K_split
// clusters, when there are merges or removals and no reset or split waits for its statistics; mode 3:
// the same table after that follow-up.
__global__ void __launch_bounds__(kThreads) k_step_relabel(int32_t* __restrict__ labels, int32_t* __restrict__ subs,
                                                           int64_t N, const int32_t* __restrict__ table, int K_arg,
                                                           const StepSummary* __restrict__ sum, int mode,
                                                           uint32_t key0, uint32_t key1, uint32_t iter) {
    if (mode == 1 ? !sum->followup : (!sum->relabel2 || (mode == 2 && sum->followup))) return;
    const int K = mode == 1 ? K_arg : sum->K_split;
    // grid-stride over the points: a pass that is gated off costs a few workgroups only
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < N; i += (int64_t)gridDim.x * kThreads) {
        const int k = labels[i], s = subs[i];
        if ((unsigned)k >= (unsigned)K || (unsigned)s > 1u) continue;
        const int act = table[2 * K + k];
        labels[i] = table[2 * k + s];
        int ns = s;
        if (act == 1 || act == 2)
            ns = (int)(philox4x32_10((uint32_t)i, 0, iter, act == 1 ? TAG_RESET : TAG_SPLITSUB, key0, key1).v[0] & 1u);
        else if (act == 3)
            ns = 0;
        else if (act == 4)
            ns = 1;
        subs[i] = ns;
    }
}

// The cluster records and histories renumbered by aux's new indices (gated as k_step_relabel's
// mode 2 / 3), through tmp.
__global__ void __launch_bounds__(kStepThreads) k_step_compact(StepCluster* __restrict__ cl,
                                                              double* __restrict__ hist,
                                                              StepCluster* __restrict__ tmp,
                                                              double* __restrict__ htmp,
                                                              const int32_t* __restrict__ aux,
                                                              const StepSummary* __restrict__ sum, int mode, int B) {
    if (!sum->relabel2 || (mode == 2 && sum->followup)) return;
    const int Ks = sum->K_split, Kn = sum->K_new;
    const int32_t* remap = aux + Ks;
    for (int k = threadIdx.x; k < Ks; k += blockDim.x)
        if (remap[k] >= 0) tmp[remap[k]] = cl[k];
    for (int e = threadIdx.x; e < Ks * B; e += blockDim.x) {
        const int k = e / B;
        if (remap[k] >= 0) htmp[(size_t)remap[k] * B + (e - k * B)] = hist[e];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < Kn; k += blockDim.x) cl[k] = tmp[k];
    for (int e = threadIdx.x; e < Kn * B; e += blockDim.x) hist[e] = htmp[e];
}

// The parts of the clusters flagged touched (of K_split) from a statistics pass over the relabelled
// points, as log-ML jobs 3r + {0, 1, 2} for the r-th touched cluster (list[r]); counts[0] = jobs.
__global__ void __launch_bounds__(kStepThreads) k_step_touched_jobs(int Ks, int E, StepCluster* __restrict__ cl,
                                                                   const double* __restrict__ stats,
                                                                   int32_t* __restrict__ list,
                                                                   int32_t* __restrict__ jobs,
                                                                   int32_t* __restrict__ counts) {
    __shared__ int ws[16];
    int n = 0;
    for (int base = 0; base < Ks; base += blockDim.x) {
        const int k = base + threadIdx.x;
        const bool f = k < Ks && cl[k].touched;
        int tot;
        const int r = step_rank(f, ws, tot);
        if (f) {
            StepCluster& c = cl[k];
            const int64_t Nl = (int64_t)stats[(size_t)(2 * k) * E];
            const int64_t Nr = (int64_t)stats[(size_t)(2 * k + 1) * E];
            step_set_part(c.l, Nl, 2 * k, -1);
            step_set_part(c.r, Nr, 2 * k + 1, -1);
            step_set_part(c.c, Nl + Nr, 2 * k, 2 * k + 1);
            const StepPart* ps[3] = {&c.c, &c.l, &c.r};
            const int q = n + r;
            list[q] = k;
            for (int p = 0; p < 3; ++p)
                for (int e = 0; e < 4; ++e) jobs[4 * (3 * q + p) + e] = ps[p]->mem[e];
        }
        n += tot;
    }
    if (threadIdx.x == 0) counts[0] = 3 * n;
}

__global__ void __launch_bounds__(kStepThreads) k_step_touched_apply(StepCluster* __restrict__ cl,
                                                                    const double* __restrict__ lm,
                                                                    const int32_t* __restrict__ list,
                                                                    const int32_t* __restrict__ counts,
                                                                    StepSummary* __restrict__ sum) {
    __shared__ int s_err;
    if (threadIdx.x == 0) s_err = 0;
    __syncthreads();
    const int n = counts[0] / 3;
    int err = 0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        StepCluster& c = cl[list[q]];
        StepPart* ps[3] = {&c.c, &c.l, &c.r};
        for (int p = 0; p < 3; ++p) {
            const double L = lm[2 * (3 * q + p)];
            if (!isfinite(L)) err = 2;
            ps[p]->lml = L;
            ps[p]->ld = lm[2 * (3 * q + p) + 1];
        }
        c.touched = 0;
    }
    if (err) atomicOr(&s_err, err);
    __syncthreads();
    if (threadIdx.x == 0) sum->error |= s_err;
}

}  // namespace dpmm
