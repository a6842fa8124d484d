// Explicit instantiations of the BF16 MFMA label sweep (see dpmm_mixed.h).
#include "dpmm_mixed.h"

namespace dpmm {
DPMM_MX_INSTANTIATE(32)
DPMM_MX_INSTANTIATE(64)
DPMM_MX_INSTANTIATE(96)
DPMM_MX_INSTANTIATE(128)
}  // namespace dpmm
