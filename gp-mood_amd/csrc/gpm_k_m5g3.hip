// Kernel instantiations for Mord = 5, ngp = 3 (all numFlux options).
#include "gpm_kernels_impl.h"

namespace gpm {

GPM_INSTANTIATE_COMBO(5, 3)

}  // namespace gpm
