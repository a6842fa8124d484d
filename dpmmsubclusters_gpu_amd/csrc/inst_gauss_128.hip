// Explicit instantiations of the Gaussian kernels (see dpmm_gauss.h).
#include "dpmm_gauss.h"

namespace dpmm {
DPMM_GAUSS_INSTANTIATE(float, 128)
}  // namespace dpmm
