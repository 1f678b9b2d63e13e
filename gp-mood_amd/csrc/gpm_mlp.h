// Device view of the learned order predictor and its launch entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpm_device.h"

struct gpm_mlp;

namespace gpm {

// bf16 weights [out][in_pad] and fp32 biases of the four layers.
struct MlpDev {
    const uint16_t *W[4];
    const float *b[4];
};

MlpDev mlp_device_view(const gpm_mlp *m);
int mlp_stencil(const gpm_mlp *m);
int mlp_classes(const gpm_mlp *m);

hipError_t mlp_launch_gate(const double4 *U, int N, int bc, double dx, double dy, int gate, int mord, order_t *gpo,
                           int8_t *pred, int *list, int *count, hipStream_t s);
hipError_t mlp_launch_predict(const double4 *U, int N, int bc, const int *list, const int *count, const MlpDev &w,
                              int stencil, int C, int mord, order_t *gpo, int8_t *pred, long long *lowered,
                              hipStream_t s);
hipError_t mlp_launch_dump(const double4 *U, int N, int bc, double dx, double dy, const order_t *gpo, int mord,
                           int stencil, float keep_mord, unsigned long long seed, int gated_only, float *X, int *Y,
                           long long cap, unsigned long long *count, hipStream_t s);
hipError_t mlp_launch_features(const double4 *U, int N, int bc, int stencil, const int *cells, int n, float *X,
                               hipStream_t s);

}  // namespace gpm
