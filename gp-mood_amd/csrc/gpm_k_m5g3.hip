// Kernel instantiations for Mord = 5, ngp = 3.
#include "gpm_kernels_impl.h"

namespace gpm {

template <>
hipError_t launch_stage_mg<5, 3>(int stage, const StageArgs &a, int nblocks, hipStream_t s)
{
    switch (stage) {
    case 1: hipLaunchKernelGGL((k_stage<5, 3, 1>), dim3(nblocks), dim3(kStageThreads), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_stage<5, 3, 2>), dim3(nblocks), dim3(kStageThreads), 0, s, a); break;
    default: hipLaunchKernelGGL((k_stage<5, 3, 3>), dim3(nblocks), dim3(kStageThreads), 0, s, a); break;
    }
    return hipGetLastError();
}

template <>
const void *mood_kernel_mg<5, 3>(int stage)
{
    if (stage == 1) return reinterpret_cast<const void *>(&k_mood<5, 3, 1>);
    if (stage == 2) return reinterpret_cast<const void *>(&k_mood<5, 3, 2>);
    return reinterpret_cast<const void *>(&k_mood<5, 3, 3>);
}

}  // namespace gpm
