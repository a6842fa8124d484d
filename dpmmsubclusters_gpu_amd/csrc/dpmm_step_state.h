// State of the device-resident sampler iteration: one record per cluster in device memory, in the
// layout shared by the host driver (which uploads it from and reads it back into its own cluster
// list) and the step kernels.  Plain C++ so that both the host compiler and hipcc read it.
#pragma once
#include <stdint.h>

namespace dpmm {

// A cluster part: point count, the statistics entries whose sum it is (-1 unused), the log
// marginal likelihood of that sum and log det psi_n of its posterior.
struct StepPart {
    int64_t N;
    int32_t mem[4];
    double lml;
    double ld;
};

// One cluster: parts, sub-cluster weights, splittable flag and the burn-out history, a ring of B
// log-likelihood sums (B = the burn-out length) held in a separate array of B doubles per cluster.
struct StepCluster {
    StepPart c, l, r;
    double lrw[2];
    int32_t splittable;
    int32_t hist_n;      // valid history entries, at most B
    int32_t hist_head;   // ring slot of the next entry
    int32_t touched;     // reset or split in this iteration: its statistics are recomputed
};

// What one device iteration decided, copied back to the host at its end.
struct StepSummary {
    int32_t K_split;    // clusters after the splits (the numbering of the relabel tables)
    int32_t K_new;      // clusters after the merges and the removal of empty clusters
    int32_t nsplit, nreset, nmerge;
    int32_t followup;   // resets or splits: their clusters' statistics come before the merges apply
    int32_t relabel2;   // merges or removals: labels and the cluster table are renumbered
    int32_t error;      // bit 0: a parameter draw failed (factor not SPD); bit 1: a log ML not finite
    int32_t npairs;     // candidate pairs given the exact merge test
    int32_t nbound;     // candidate pairs rejected by the merge bound before their log ML
};

// Scalars of one iteration's kernels.
struct StepArgs {
    uint32_t key0, key1, t;
    int K;               // clusters at the start of the iteration
    int fin;             // final iterations: no splits, merges also above log 0.1
    int no_more_splits;  // no split or merge tests (run_model's flag, or the cluster cap reached)
    int kcap;            // largest cluster count the device kernels take
    int B;               // burn-out length
    int E;               // doubles per statistics entry
    int merge_bound;     // NIW: reject hopeless merge pairs by a bound before their log ML
    double alpha;
};

}  // namespace dpmm
