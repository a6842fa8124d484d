// Explicit instantiations of the Gaussian kernels (see dpmm_gauss.h).
#include "dpmm_gauss.h"

namespace dpmm {
DPMM_GAUSS_INSTANTIATE(float, 4)
DPMM_GAUSS_INSTANTIATE(float, 8)
DPMM_GAUSS_INSTANTIATE(float, 16)
DPMM_GAUSS_INSTANTIATE(float, 32)
DPMM_GAUSS_INSTANTIATE(float, 48)
DPMM_GAUSS_INSTANTIATE(double, 4)
DPMM_GAUSS_INSTANTIATE(double, 8)
DPMM_GAUSS_INSTANTIATE(double, 16)
DPMM_GAUSS_INSTANTIATE(double, 32)
DPMM_GAUSS_INSTANTIATE(double, 48)
}  // namespace dpmm
