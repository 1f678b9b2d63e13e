// Kernel instantiations for the first-order Godunov method (FOG: Mord 1 template, one Gauss point).
#include "gpm_kernels_impl.h"

namespace gpm {
GPM_INSTANTIATE_COMBO(1, 1)
}  // namespace gpm
