// Kernel instantiations for Mord = 7, ngp = 4.
#include "gpm_kernels_impl.h"

namespace gpm {

template <>
hipError_t launch_stage_mg<7, 4>(int stage, const StageArgs &a, int nblocks, hipStream_t s)
{
    switch (stage) {
    case 1: hipLaunchKernelGGL((k_stage<7, 4, 1>), dim3(nblocks), dim3(kStageThreads), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_stage<7, 4, 2>), dim3(nblocks), dim3(kStageThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_stage<7, 4, 3>), dim3(nblocks), dim3(kStageThreads), 0, s, a); break;
    }
    return hipGetLastError();
}

template <>
const void *mood_kernel_mg<7, 4>(int stage)
{
    if (stage == 1) return reinterpret_cast<const void *>(&k_mood<7, 4, 1>);
    if (stage == 2) return reinterpret_cast<const void *>(&k_mood<7, 4, 2>);
    return reinterpret_cast<const void *>(&k_mood<7, 4, 3>);
}

}  // namespace gpm
