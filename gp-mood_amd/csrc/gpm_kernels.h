// Kernel argument blocks and launch entry points shared by the HIP kernels and
// the host side of libgpmood.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpm_device.h"

namespace gpm {

// First-pass tile: kTX x kTY cells per workgroup of kStageThreads threads.
#ifndef GPM_TX
#define GPM_TX 32
#endif
#ifndef GPM_TY
#define GPM_TY 11
#endif
#ifndef GPM_STAGE_WAVES_PER_EU
#define GPM_STAGE_WAVES_PER_EU 3
#endif
constexpr int kTX = GPM_TX;
constexpr int kTY = GPM_TY;
constexpr int kStageThreads = 256;
constexpr int kMoodThreads = 256;
constexpr int kSoloMax = 64;  // largest list a single workgroup takes over
// Halo width: the stencil radius R = (Mord-1)/2 plus one cell for the face
// neighbour, and at least 2 for the detection (U2 reads a cell's 5-point
// neighbours' 5-point stencils).
template <int MORD>
constexpr int kHalo = ((MORD - 1) / 2 + 1) > 2 ? ((MORD - 1) / 2 + 1) : 2;

constexpr int kMaxPasses = 4096;  // MOOD passes per stage before the run is declared failed
constexpr int kCountersPerStage = 2 * kMaxPasses + 2;

struct DevStats {
    long long steps;
    long long detected_step;
    long long detected_total;
    long long second_pass_cells;
    long long a_priori_lowered;
    int passes[3];
    int max_passes;
};

struct StageArgs {
    const double4 *__restrict__ Uin;  // stage input (interior, N*N)
    const double4 *__restrict__ Un;   // step-start state for the RK combination
    double4 *__restrict__ Uout;       // combined stage output
    order_t *__restrict__ gpo;        // packed CellGPO (interior)
    int *__restrict__ slist1;         // cells to recompute in the first correction pass
    int *cnt_s1;                      // their count
    int *cnt_f0;                      // troubled cells of the first pass
    int *mark;                        // per-cell pass stamp (dedup of neighbourhoods)
    const int *stamp;                 // running stamp base
    const DevTables *__restrict__ T;
    const double *__restrict__ tstate; // t, dt, dt/dx, dt/dy
    int N, bc, dmp, u2, u2tol, cascade;
    int apriori;                      // start from the a priori orders in pred (learned predictor)
    int frozen;                       // orders fixed to pred, no detection (frozen-mask step)
    const int8_t *__restrict__ pred;  // a priori orders (read only here; gpo is rewritten)
    double dx, dy;
    int tiles_x, ntiles;
};

struct MoodArgs {
    const double4 *__restrict__ Uin;
    const double4 *__restrict__ Un;
    double4 *__restrict__ Uout;
    order_t *gpo;
    int *slist[2];  // recompute lists, pass p reads slist[p & 1]
    int *cnt_f;     // [kMaxPasses+1] troubled-cell counts per pass (pass 0 = k_stage)
    int *cnt_s;     // [kMaxPasses+1] recomputed-cell counts per pass
    int *mark;
    int *stamp;
    int *err;
    int *solo_p;    // first pass after a run of workgroup-0-only passes
    DevStats *stats;
    const DevTables *__restrict__ T;
    const double *__restrict__ tstate;
    int N, bc, dmp, u2, u2tol, cascade, pmax, stage_idx;
    int solo_max;   // lists of at most this many cells run on workgroup 0 alone (0: never)
    double dx, dy;
};

enum { kDtCompute = 0, kDtGiven = 1 };

struct DtArgs {
    const double4 *__restrict__ U;
    double *block_min;
    unsigned *done;
    double *tstate;
    double *dt_log;
    int *counters;
    int ncounters;
    DevStats *stats;
    int N, mode, use_tmax, advance_t, step_idx, new_step;
    double dx, dy, cfl, tmax, dtfinal, dt_given;
};

template <int MORD, int NGP>
hipError_t launch_stage_mg(int flux, int stage, const StageArgs &a, int nblocks, hipStream_t s);
template <int MORD, int NGP>
const void *mood_kernel_mg(int flux, int stage);

// flux: numFlux id (kHLLC, kHLL, kUnsplit)
hipError_t launch_stage(int mord, int ngp, int flux, int stage, const StageArgs &a, int nblocks, hipStream_t s);
const void *mood_kernel(int mord, int ngp, int flux, int stage);
bool supported(int mord, int ngp);
hipError_t launch_dt(const DtArgs &a, int nblocks, hipStream_t s);
hipError_t launch_unpack_orders(const order_t *g, int8_t *out, size_t n, hipStream_t s);
hipError_t launch_orders_from_pred(const int8_t *pred, order_t *g, size_t n, hipStream_t s);

}  // namespace gpm
