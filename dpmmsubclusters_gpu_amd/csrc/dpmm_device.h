// Internal interface between the device layer (dpmm_device.hip) and the host driver.
#pragma once
#include <stdint.h>

#include "../../include/dpmm_hip.h"

namespace dpmm {
int ctx_dims(const dpmm_ctx* c, int64_t* N, int* D, int* prior, int* precision);
// Whether dpmm_set_outlier_prior has been called on the context.
bool ctx_has_outlier_prior(const dpmm_ctx* c);
// Records msg as dpmm_last_error() and returns code.
int set_error(int code, const char* msg);
}
