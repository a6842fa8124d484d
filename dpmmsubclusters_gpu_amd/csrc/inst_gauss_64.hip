// Explicit instantiations of the Gaussian kernels (see dpmm_gauss.h).
#include "dpmm_gauss.h"

namespace dpmm {
DPMM_GAUSS_INSTANTIATE(float, 64)
DPMM_GAUSS_INSTANTIATE(double, 64)
}  // namespace dpmm
