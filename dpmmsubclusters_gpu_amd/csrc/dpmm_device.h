// Internal interface between the device layer (dpmm_device.hip) and the host driver.
#pragma once
#include <stdint.h>

#include "../../include/dpmm_hip.h"
#include "dpmm_step_state.h"

namespace dpmm {
int ctx_dims(const dpmm_ctx* c, int64_t* N, int* D, int* prior, int* precision);
// Whether dpmm_set_outlier_prior has been called on the context.
bool ctx_has_outlier_prior(const dpmm_ctx* c);
// Records msg as dpmm_last_error() and returns code.
int set_error(int code, const char* msg);

// Device-resident sampler iteration (kernels in dpmm_step.h).
// step_upload: K cluster records and their burn-out histories (B doubles each, ring order) to the
// device table.
int step_upload(dpmm_ctx* c, int B, int K, const StepCluster* cl, const double* hist);
// One iteration on the device: parameter draw, label and sub-label sweeps, statistics, every part's
// and candidate pair's log ML, the reset / split / merge / removal decisions and relabel passes, with
// one synchronisation at its end (two when resets or splits need their clusters' statistics first).
// In graph mode everything up to that synchronisation is one hipGraph.  trace_labels / trace_subs
// (host arrays of N, or null) receive the sweeps' output before the relabel passes.
int step_iteration(dpmm_ctx* c, const StepArgs& a, int sweep_final, int full_mu_cov, int use_weight,
                   int32_t* trace_labels, int32_t* trace_subs, StepSummary* out);
// The device table's K records and histories, and the last draw's Kw mixture weights (any may be null).
int step_download(dpmm_ctx* c, int K, int Kw, StepCluster* cl, double* hist, double* weights);
}
