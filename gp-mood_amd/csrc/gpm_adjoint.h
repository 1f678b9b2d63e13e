// Adjoint kernels of the frozen-mask SSP-RK3 step (gpm_adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpm_device.h"

namespace gpm {

struct AdjFaceArgs {
    const double4 *__restrict__ V;   // stage input (interior, N*N)
    const double4 *__restrict__ Wb;  // adjoint of the stage's Euler output, before scaling by alpha
    const int8_t *__restrict__ mask; // frozen CellGPO of the stage (interior)
    const DevTables *__restrict__ T;
    double4 *__restrict__ ub;        // [face][g][(N+2)^2] adjoint of each reconstructed side state
    int8_t *__restrict__ ordx;       // [N][N+1] x-face orders
    int8_t *__restrict__ ordy;       // [N+1][N] y-face orders
    double alpha, dtdx, dtdy;
    int N, bc;
};

struct AdjGatherArgs {
    const double4 *__restrict__ ub;
    const int8_t *__restrict__ ordx;
    const int8_t *__restrict__ ordy;
    const DevTables *__restrict__ T;
    double4 *__restrict__ G;         // [(N+2R+2)^2] adjoint of the ghost-extended stage input
    int N;
};

struct AdjFoldArgs {
    const double4 *__restrict__ G;
    const double4 *__restrict__ Wn;  // adjoint of the stage output (RK-combined)
    const double4 *__restrict__ B;   // optional extra term
    double4 *__restrict__ out;
    double4 *A;                      // optional running sum for the step-start state
    double alpha, gamma, aA, aU;
    int N, bc;
};

hipError_t launch_adj_stage(int mord, int ngp, const AdjFaceArgs &fa, const AdjGatherArgs &ga, const AdjFoldArgs &fo,
                            hipStream_t s);
bool adj_supported(int mord, int ngp);
hipError_t launch_hllc_vjp(int dir, int n, const double *ul, const double *ur, const double *Fb, double *ulb,
                           double *urb, hipStream_t s);

}  // namespace gpm
