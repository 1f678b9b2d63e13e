// Kernel argument blocks and launch entry points shared by the HIP kernels and
// the host side of libgpmood.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpm_device.h"

namespace gpm {

// First-pass tile: TX x TY cells per workgroup of kStageThreads threads, and
// the wavefronts per SIMD its launch bounds ask for.  Mord 3 (every problem
// file of the reference) runs 16 x 16 tiles at 4 wavefronts per SIMD, 5 %
// faster than the 32 x 11 / 3 of the higher orders, whose larger halo and
// register use favour the wider tile (results do not depend on the tiling).
#ifndef GPM_TX
#define GPM_TX 32
#endif
#ifndef GPM_TY
#define GPM_TY 11
#endif
#ifndef GPM_STAGE_WAVES_PER_EU
#define GPM_STAGE_WAVES_PER_EU 3
#endif
#ifndef GPM_TX3
#define GPM_TX3 16
#endif
#ifndef GPM_TY3
#define GPM_TY3 16
#endif
#ifndef GPM_STAGE_WAVES_PER_EU3
#define GPM_STAGE_WAVES_PER_EU3 4
#endif
__host__ __device__ constexpr int stage_tx(int mord) { return mord == 3 ? GPM_TX3 : GPM_TX; }
__host__ __device__ constexpr int stage_ty(int mord) { return mord == 3 ? GPM_TY3 : GPM_TY; }
constexpr int stage_waves(int mord) { return mord == 3 ? GPM_STAGE_WAVES_PER_EU3 : GPM_STAGE_WAVES_PER_EU; }
constexpr int kStageThreads = 256;
#ifndef GPM_MOOD_THREADS
#define GPM_MOOD_THREADS 512
#endif
constexpr int kMoodThreads = GPM_MOOD_THREADS;
// workgroup size of the wide launch of long-list MOOD passes (k_mood_wide)
constexpr int kWideThreads = 256;
constexpr int kSoloMax = 32;  // largest list a single workgroup takes over
// Halo width: the stencil radius R = (Mord-1)/2 plus one cell for the face
// neighbour, and at least 2 for the detection (U2 reads a cell's 5-point
// neighbours' 5-point stencils).
template <int MORD>
constexpr int kHalo = ((MORD - 1) / 2 + 1) > 2 ? ((MORD - 1) / 2 + 1) : 2;

// Per-stage int counters of the MOOD passes (zeroed by k_dt at every step).
// The recompute counts form a ring of three: pass p reads kCntS0 + p % 3, adds
// the next pass's cells to kCntS0 + (p + 1) % 3 and clears kCntS0 + (p + 2) %
// 3, whose last reader was pass p - 1; so a stage can run any number of
// passes (at most 3 x cells + 1: every pass but the last lowers a cell).
enum {
    kCntFail0 = 0,    // troubled cells of the first pass (k_stage)
    kCntDet = 1,      // troubled cells of the correction passes
    kCntS0 = 2,       // kCntS0 .. kCntS0 + 2: the ring of recompute counts
    kCntPass = 5,     // per-pass launches: the current pass minus one
    kCntDone = 6,     // per-pass launches: workgroups done with the launch
    kCntNext = 7,     // the pass that follows a launch or a run of workgroup-0 passes
    kCntSecond = 8,   // cells recomputed in the correction passes
    kCntPending = 9,  // per-pass launches: the current pass has cells to recompute
    kCntStop = 10,    // the first pass of workgroup 0's last run of passes, once that run has ended
    // ordinary persistent launch, each on a cache line of its own (the
    // workgroups poll them while the passes update the counters above):
    kCntArrive = 32,  // workgroups arrived (| kArriveAbort: one runs alone)
    kCntBar = 64,     // grid-barrier arrivals, gridDim.x per barrier
    kCntRel = 96,     // grid barriers released (the last arrival publishes each)
    kCountersPerStage = 128
};
constexpr int kArriveAbort = 1 << 30;
constexpr int kBarStalled = 1 << 29;  // kCntRel from here on: a grid-barrier wait stalled, the launch ends

// How k_mood's workgroups wait for each other between passes.
enum {
    kMoodCoop = 0,  // cooperative launch: cooperative-groups grid barrier
    kMoodOrd = 1,   // ordinary launch: rendezvous, then grid barriers on kCntBar / kCntRel
};

struct DevStats {
    long long steps;
    long long detected_step;
    long long detected_total;
    long long second_pass_cells;
    long long a_priori_lowered;
    int passes[5];  // MOOD passes of each stage of the last step (up to 5 stages: SSP_RK4)
    int max_passes;
    // the reference network (NN_GP_MOOD / NN_GP_MOOD_CC with its own weights)
    long long apriori_step;    // count_detected_cell_a_priori of the last step
    long long nn_passes_step;  // passes of the last step whose check lowered a cell
                               // (increments of count_steps_NN_produced_NAN, mod_FE.f90:154-157)
    long long nn_passes_total;
    // the longest list of a stage's first two correction passes since the
    // host last read it (the lists the wide launches take)
    int long_list;
    // ordinary persistent MOOD launches since the host last read them: stages
    // whose passes ran on one workgroup (the others were not resident
    // together), and grid-barrier waits that reached their bound (the stage's
    // results are then invalid and the call fails)
    int mood_alone;
    int mood_stall;
};

constexpr int kMaxStages = 5;

// How the Forward_Euler result `cand` of a stage (after the gravity source)
// becomes the stage output, with Un the step-start state and Uin the stage
// input (mod_SSP_RK2.f90:16-19, mod_SSP_RK3.f90:23-33, mod_SSP_RK4.f90:20-32;
// SSP_RK4's Forward_Euler returns Uout - Uin, mod_FE.f90:187).
enum {
    kCmbCopy = 0,  // FE, and stage 1 of SSP_RK2 / SSP_RK3: cand
    kCmbRK3b,      // 0.75*Un + 0.25*cand
    kCmbRK3c,      // (1./3)*Un + (2./3)*cand
    kCmbRK2b,      // 0.5*(Un + cand)
    kCmbRK4a,      // Un + c1*(cand - Un)
    kCmbRK4b,      // a20*Un + a21*Uin + c2*(cand - Uin)
    kCmbRK4c,      // a30*Un + a32*Uin + c3*(cand - Uin)
    kCmbRK4d,      // a40*Un + a43*Uin + c4*FU3, FU3 = cand - Uin (also written to W)
    kCmbRK4e,      // f2*X1 + f3*X2 + ff3*X3 + f4*Uin + ff4*(cand - Uin), X = (U2, U3, FU3)
};

// Buffers a stage uses only with SSP_RK4 or the ghost ring.  They live in
// device memory behind one pointer (they never change after set-up), which
// keeps the kernel argument, and the scalar registers it occupies, small.
struct StepExtra {
    const double4 *X1, *X2, *X3;  // kCmbRK4e operands: U2, U3, FU3
    double4 *W;                   // kCmbRK4d: FU3
    double4 *band;                // if set: pre-source FE values of the cells within 4 of the boundary
};

// Per-stage control of the step shared by k_stage and k_mood.
struct StepCtl {
    const double4 *__restrict__ Un;  // step-start state
    const StepExtra *ext;            // this stage's StepExtra (device memory)
    int kind;                        // kCmb*
    int gravity;                     // problem RT: source term after MOOD (mod_FE.f90:183-186)
    int trk;                         // t_rk of the stage: t (0), t+dt (1), t+0.5*dt (2)
    int band;                        // ext->band is set (the last stage of a ghost-ring step)
};

struct StageArgs {
    const double4 *__restrict__ Uin;  // stage input (interior, NX*NY)
    double4 *__restrict__ Uout;       // combined stage output
    order_t *__restrict__ gpo;        // packed CellGPO (interior)
    uint8_t *__restrict__ dflag;      // per cell: the stage-input parts of the decision (kFlag*)
    int *__restrict__ slist1;         // cells to recompute in the first correction pass
    int *cnt_s1;                      // their count
    int *cnt_f0;                      // troubled cells of the first pass
    int *mark;                        // per-cell pass stamp (dedup of neighbourhoods)
    const int *stamp;                 // running stamp base
    const DevTables *__restrict__ T;
    const double *__restrict__ tstate; // t, dt, dt/dx, dt/dy, t at the step start
    Bnd B;
    StepCtl C;
    int dmp, u2, u2tol, cascade;
    int apriori;                      // start from the a priori orders in pred (learned predictor)
    int frozen;                       // orders fixed to pred, no detection (frozen-mask step)
    const int8_t *__restrict__ pred;  // a priori orders (read only here; gpo is rewritten)
    const float *__restrict__ vnn;    // NN_GP_MOOD_CC: values_NN for the blended reconstruction (Mord 3), or null
    int recon;                        // kRecon*
    double dx, dy;
    int tiles_x, ntiles;
};

// Device-wide face-flux cache of a stage (open-addressing hash table): slot
// h holds key word kw = (stage tag << 32) | face-and-order key, the flux, and
// `ready` = the stage tag once the flux is written.  Entries of earlier
// stages are free slots; nothing is cleared between stages.
struct FluxTable {
    unsigned long long *kw;
    unsigned *ready;
    double4 *data;
    int bits;        // log2 of the slot count; 0: no table
    int radius;      // the helpers take the cells within this distance of workgroup 0's list
    int *beacon;     // [2 + kSoloMax]: workgroup 0's current pass, list length and list
    unsigned *done;  // per cell: the stage tag once a workgroup has taken the cell
};

struct MoodArgs {
    const double4 *__restrict__ Uin;
    double4 *__restrict__ Uout;
    order_t *gpo;
    const uint8_t *__restrict__ dflag;  // the first pass's record of the stage-input parts of the decision
    int *slist[2];  // recompute lists, pass p reads slist[p & 1]
    int *cnt;       // the stage's kCountersPerStage counters (kCnt*)
    int *mark;
    int *stamp;
    DevStats *stats;
    const DevTables *__restrict__ T;
    const double *__restrict__ tstate;
    Bnd B;
    StepCtl C;
    int dmp, u2, u2tol, cascade, stage_idx;
    int nn_passes;  // count the passes that lowered a cell (a step of the reference network)
    int solo_max;   // lists of at most this many cells run on workgroup 0 alone (0: never)
    int wide_min;   // a wide launch runs the current pass only when it lists more cells than this
    int fcache;     // workgroup 0's passes reuse face fluxes and compute likely next ones ahead
    int sync;       // kMoodCoop or kMoodOrd: how k_mood was launched
    int rv_spins;   // kMoodOrd: polls of the rendezvous before one workgroup takes every pass alone
    int rv_alone;   // kMoodOrd: the first workgroup to arrive takes every pass alone (tests)
    int bar_spins;  // kMoodOrd: polls of a grid-barrier wait before it gives up (a stall)
    FluxTable gfc;  // while workgroup 0 runs passes alone, the other workgroups fill it (persistent kernel)
    const float *__restrict__ vnn;  // as StageArgs::vnn
    int recon;                      // kRecon*
    double dx, dy;
};

enum { kDtCompute = 0, kDtGiven = 1 };

struct DtArgs {
    const double4 *__restrict__ U;
    const double4 *__restrict__ ring;  // ghost cells of U (ring layout), or null when they copy interior cells
    double *block_min;
    unsigned *done;
    double *tstate;
    double *dt_log;
    int *counters;
    int ncounters;
    DevStats *stats;
    int NX, NY, mode, use_tmax, advance_t, step_idx, new_step;
    double dx, dy, cfl, tmax, dtfinal, dt_given;
};

// Ghost cells of a state in a compact ring: rows -4..-1 and NY..NY+3 over
// columns -4..NX+3, then columns -4..-1 and NX..NX+3 of rows 0..NY-1.
__host__ __device__ __forceinline__ int ring_size(int NX, int NY) { return 8 * (NX + 8) + 8 * NY; }
__host__ __device__ __forceinline__ void ring_cell(int r, int NX, int NY, int &i, int &j)
{
    const int W = NX + 8;
    if (r < 8 * W) {
        const int row = r / W;
        i = r % W - 4;
        j = row < 4 ? row - 4 : NY + row - 4;
        return;
    }
    r -= 8 * W;
    j = r / 8;
    const int c = r % 8;
    i = c < 4 ? c - 4 : NX + c - 4;
}
__host__ __device__ __forceinline__ int ring_index(int i, int j, int NX, int NY)
{
    const int W = NX + 8;
    if (j < 0) return (j + 4) * W + i + 4;
    if (j >= NY) return (4 + j - NY) * W + i + 4;
    return 8 * W + j * 8 + (i < 0 ? i + 4 : 4 + i - NX);
}

// How the ring of a step's output is formed (the whole-array statements of the
// integrators act on the ghost cells too, so the reference's Setdt sees these
// values): G_t(A) is Boundary_C of A's interior at t_rk = t, S the gravity
// source, B the pre-source Forward_Euler values of the last stage (band).
enum {
    kRingFill = 0,  // G_t(A), t = tstate[0]
    kRingFE,        // S(G_t0(B))
    kRingRK2,       // 0.5*(G_t0(Un) + S(G_t0(B)))
    kRingRK3,       // (1./3)*G_t0(Un) + (2./3)*S(G_{t0+dt/2}(B))
    kRingFace,      // methods without DETECTION (FOG, GP_SE, Unlim_POL): the last stage's
                    // E = S(G_tb(B) + the face loop's updates of the first ghost layer), B the
                    // last stage's input; then E (FE), 0.5*(G_t0(Un) + E) (SSP_RK2) or
                    // (1./3)*G_t0(Un) + (2./3)*E (SSP_RK3, tb = t0 + dt/2)
    kRingRK4,       // SSP_RK4 (every stage at t_rk = t0; FE returns Uout - Uin, whose
                    // Uin ghosts Boundary_C has refilled): f2*G(U2) + f3*G(U3)
                    // + ff3*(S(G(B3)) - G(U3)) + f4*G(U4) + ff4*(S(G(B)) - G(U4)),
                    // B3 the fourth stage's band
};

struct RingArgs {
    const double4 *__restrict__ A;     // kRingFill: the state; else the step-start state Un
    const double4 *__restrict__ band;
    const double4 *__restrict__ Bin;  // kRingFace: the last stage's input
    double4 *__restrict__ ring;
    const double *__restrict__ tstate;
    Bnd B;
    const DevTables *__restrict__ T;  // kRingFace: the reconstruction tables and Gauss weights
    int mode, gravity, flux;
    int integ;                        // kRingFace: kRingFE, kRingRK2 or kRingRK3 (the combination)
    const double4 *__restrict__ X2, *__restrict__ X3, *__restrict__ X4;  // kRingRK4: U2, U3, U4
    const double4 *__restrict__ band3;                                   // kRingRK4: B3
};

template <int MORD, int NGP>
hipError_t launch_stage_mg(int flux, bool gen, const StageArgs &a, int nblocks, hipStream_t s);
template <int MORD, int NGP>
const void *mood_kernel_mg(int flux, bool gen, bool cooperative);
template <int MORD, int NGP>
const void *mood_wide_kernel_mg(int flux, bool gen);

// flux: numFlux id (kHLLC, kLLF, kHLL, kUnsplit); gen: the general boundary
// kernels (formula boundaries, or problem RT's source term) instead of the
// copy-boundary ones
hipError_t launch_stage(int mord, int ngp, int flux, bool gen, const StageArgs &a, int nblocks, hipStream_t s);
// The MOOD correction kernel: cooperative (every pass of a stage in one
// launch, grid barriers between passes) or the per-launch form (one pass, or
// one run of workgroup-0 passes, per launch; launches after the last pass
// return at once).  Both take one MoodArgs.
const void *mood_kernel(int mord, int ngp, int flux, bool gen, bool cooperative);
// the wide launch of long-list MOOD passes (k_mood_wide)
const void *mood_wide_kernel(int mord, int ngp, int flux, bool gen);
// The stage's MOOD statistics after per-launch passes (the cooperative kernel
// records them itself).
hipError_t launch_mood_finish(const MoodArgs &a, hipStream_t s);
bool supported(int mord, int ngp);
hipError_t launch_dt(const DtArgs &a, int nblocks, hipStream_t s);
hipError_t launch_ring(const RingArgs &a, hipStream_t s);
// kRingFace for (mord, ngp): (1, 1) FOG's cell values, else the face states of
// the order-mord reconstruction in T (GP weights, or the polynomial weights)
hipError_t launch_ring_face(int mord, int ngp, int flux, const RingArgs &a, hipStream_t s);
template <int MORD, int NGP>
hipError_t launch_ring_face_mg(int flux, const RingArgs &a, hipStream_t s);
hipError_t launch_unpack_orders(const order_t *g, int8_t *out, size_t n, hipStream_t s);
hipError_t launch_ddiv(const double *a, const double *b, double *q, long long n, bool shared, hipStream_t s);
hipError_t launch_run_reset(int *mark, order_t *gpo, size_t n, order_t g0, int *counters, int ncounters,
                            DevStats *stats, double *tstate, double t, unsigned *done, int *stamp, hipStream_t s);

// The reference's own network (mod_NN.f90): per-cell a priori orders of a
// stage input with a copy boundary.
struct RefNnArgs {
    const double4 *__restrict__ U;  // stage input (interior)
    const float *__restrict__ w;    // w0p [len][KP] | bias0 | w1p [len][kRefNnMaxLen] | bias1 | weight2(2,len) | bias2
                                    // (rows zero-padded; KP = refnn_input_pad(L))
    int8_t *pred;                   // CellGPO after the prediction (1 or 3)
    int8_t *priori;                 // CellGPO_priori (1, 3 or Mord)
    float *rowv;                    // [NY][2]: values_NN of each row's last cell (NN_GP_MOOD_CC), or null
    DevStats *stats;
    Bnd B;
    int L, len, mord, softmax;
};
constexpr int kRefNnMaxLen = 64;     // hidden units held in registers
constexpr int kRefNnThreads = 64;    // one wavefront per workgroup
hipError_t launch_refnn(const RefNnArgs &a, hipStream_t s);
int refnn_input_pad(int L);  // padded input width of the network kernel (64 or 128)

// Training samples in the reference network's format (append_to_NN_datatset).
struct RefDumpArgs {
    const double4 *__restrict__ U;  // the first stage's input
    const order_t *__restrict__ gpo;
    const double *__restrict__ tstate;
    Bnd B;
    float *inputs;                  // [capacity][L]
    float *labels;                  // [capacity][2]
    long long *cells;               // [capacity] cell ids, or null
    long long capacity;
    unsigned long long *count;
    long long *nr;                  // NR0, NR1 (running)
    long long *nr_start;            // NR0, NR1 at the start of the stage
    unsigned long long seed;
    long long step;
    int L;
};
hipError_t launch_refnn_dump(const RefDumpArgs &a, int mord, hipStream_t s);
// values_NN(-1:lf+2, -1:nf+2, 2) := (0, 1), then each row's values written at
// values_NN(L, n, :) in row order, as the reference indexes it (mod_NN.f90:41,44)
hipError_t launch_refnn_values(float *vnn, const float *rowv, int NX, int NY, int L, bool scatter, hipStream_t s);
// exp of the floats with bit patterns first .. first+n-1 by the network's expf
hipError_t launch_nn_expf_bits(uint32_t first, long long n, float *out, hipStream_t s);
void nn_expf_bits_host(uint32_t first, long long n, float *out);
hipError_t launch_orders_from_pred(const int8_t *pred, order_t *g, size_t n, hipStream_t s);
// host-array exchange of a ghosted state [NY+8][NX+8] through one contiguous copy
hipError_t launch_pack_ghosted(const double4 *U, const double4 *ring, const Bnd &B, double4 *out, hipStream_t s);
hipError_t launch_unpack_ghosted(const double4 *in, double4 *U, double4 *ring, int NX, int NY, hipStream_t s);
hipError_t launch_cellgpo_ghosted(const order_t *g, int32_t *out, int NX, int NY, int mord, int all_mord,
                                  hipStream_t s);

}  // namespace gpm
