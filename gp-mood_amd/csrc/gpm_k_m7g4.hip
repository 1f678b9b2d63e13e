// Kernel instantiations for Mord = 7, ngp = 4 (all numFlux options).
#include "gpm_kernels_impl.h"

namespace gpm {

GPM_INSTANTIATE_COMBO(7, 4)

}  // namespace gpm
