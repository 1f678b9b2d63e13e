// Kernel instantiations for Mord = 3, ngp = 2 (all numFlux options).
#include "gpm_kernels_impl.h"

namespace gpm {

GPM_INSTANTIATE_COMBO(3, 2)

}  // namespace gpm
