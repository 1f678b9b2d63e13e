// Dispatch from a runtime (Mord, ngp) to the compiled kernel instantiations.
#include "gpm_kernels.h"

namespace gpm {

// (1, 1) is the first-order Godunov method (FOG), the others GP-MOOD.
#define GPM_COMBOS(X) X(1, 1) X(3, 2) X(5, 3) X(7, 4)

hipError_t launch_stage(int mord, int ngp, int flux, bool gen, const StageArgs &a, int nblocks, hipStream_t s)
{
#define X(M, G) if (mord == M && ngp == G) return launch_stage_mg<M, G>(flux, gen, a, nblocks, s);
    GPM_COMBOS(X)
#undef X
    return hipErrorInvalidValue;
}

const void *mood_kernel(int mord, int ngp, int flux, bool gen, bool cooperative)
{
#define X(M, G) if (mord == M && ngp == G) return mood_kernel_mg<M, G>(flux, gen, cooperative);
    GPM_COMBOS(X)
#undef X
    return nullptr;
}

const void *mood_wide_kernel(int mord, int ngp, int flux, bool gen)
{
#define X(M, G) if (mord == M && ngp == G) return mood_wide_kernel_mg<M, G>(flux, gen);
    GPM_COMBOS(X)
#undef X
    return nullptr;
}

hipError_t launch_ring_face(int mord, int ngp, int flux, const RingArgs &a, hipStream_t s)
{
#define X(M, G) if (mord == M && ngp == G) return launch_ring_face_mg<M, G>(flux, a, s);
    GPM_COMBOS(X)
#undef X
    return hipErrorInvalidValue;
}

bool supported(int mord, int ngp) { return mord > 1 && mood_kernel(mord, ngp, kHLLC, false, true) != nullptr; }

}  // namespace gpm
