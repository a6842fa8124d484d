// Counter-based random numbers for the sampler: Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), keyed by the 64-bit seed.
//
// Every stochastic step of an iteration t reads a fixed counter, so a point's noise does not
// depend on launch shape, tile order or thread count:
//   init labels        (i, 0, 0, TAG_INIT)    word0 -> label, word1 & 1 -> sub-label
//   label noise        (i, k/4, t, TAG_LABEL) word k%4 -> Gumbel noise of cluster k
//   sub-label draw     (i, 0, t, TAG_SUB)     word0 -> logistic threshold
//   reset relabel      (i, 0, t, TAG_RESET)   word0 & 1
//   split relabel      (i, 0, t, TAG_SPLITSUB) word0 & 1
//   sequential streams (c0, n, t, tag), n = 0, 1, ... : parameter draws, Dirichlet weights,
//   split / merge test uniforms, the sub-cluster weights a split or a merge sets; a multinomial
//   parameter draw gives coordinate d the stream that starts at n = 1024 d, and the mixture
//   weights' Dirichlet gives component i the stream (i, n, t, TAG_MIXW).
// The categorical draws are Gumbel-max (labels) and a logistic threshold (sub-labels), which
// have the distributions of the reference's softmax + alias sampling (cudaKernel.cu:27-124,
// 352-403).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DPMM_HD __host__ __device__ __forceinline__
#else
#define DPMM_HD inline
#endif

namespace dpmm {

enum RngTag : uint32_t {
    TAG_INIT = 1, TAG_LABEL = 2, TAG_SUB = 3, TAG_RESET = 4, TAG_SPLITSUB = 5,
    TAG_PARAM = 6, TAG_LRW = 7, TAG_MIXW = 8, TAG_SPLIT_U = 9, TAG_MERGE_U = 10,
    TAG_SPLIT_LRW = 11, TAG_MERGE_LRW = 12   // sub-cluster weights set by a split / a merge
};

struct Ctr4 { uint32_t v[4]; };

DPMM_HD Ctr4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    Ctr4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// u = (w + 0.5) 2^-32 in (0, 1).
DPMM_HD double u01_f64(uint32_t w) { return ((double)w + 0.5) * 0x1p-32; }

}  // namespace dpmm
