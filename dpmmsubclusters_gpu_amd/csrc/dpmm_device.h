// Internal interface between the device layer (dpmm_device.hip) and the host driver.
#pragma once
#include <stdint.h>

#include "../../include/dpmm_hip.h"

namespace dpmm {
int ctx_dims(const dpmm_ctx* c, int64_t* N, int* D, int* prior, int* precision);
}
