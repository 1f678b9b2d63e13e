// Host side of libgpmood: the C ABI declared in include/gpmood.h.
//
// Owns the device buffers of one solver context and enqueues the step on the
// context's HIP stream: k_dt, then per RK stage k_stage (first MOOD pass) and
// the persistent k_mood (remaining passes), then, for boundaries whose ghost
// cells are not copies of interior cells, k_ring (the ghost values the next
// Setdt sees).  Host-array entry points convert between the reference's
// ghosted layout U(4, lb:le, nb:ne) and the device's interior-only layout.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpmood.h"
#include "gpm_internal.h"
#include "gpm_kernels.h"
#include "gpm_adjoint.h"
#include "gpm_mlp.h"

using namespace gpm;

struct gpm_ctx {
    gpm_config cfg{};
    int NX = 0, NY = 0, mord = 0, ngp = 0, bc = 0, device = 0;
    int N = 0;  // NX where only square grids are supported (learned predictor, adjoint)
    int integrator = 0, nstages = 0, gravity = 0;
    bool fog = false;        // method FOG: first-order Godunov, no detection
    bool nodetect = false;   // FOG, GP_SE, Unlim_POL: one pass, no DETECTION (mod_FE.f90:160-162)
    bool ring_mode = false;  // ghost cells are not all copies of interior cells: keep them in `ring`
    Bnd B{}, Bh{};           // the boundary for kernels (states in st_dev) and for host code (st_host)
    double4 st_host[kBndStates] = {};
    double4 *st_dev = nullptr;
    double dx = 0, dy = 0;
    hipStream_t stream = nullptr;
    DevTables hT{};
    DevTables *dT = nullptr;
    double4 *dU = nullptr, *dA = nullptr, *dB = nullptr;
    double4 *dC = nullptr, *dD = nullptr, *dE = nullptr, *dF = nullptr;  // SSP_RK4 stage states
    double4 *ring = nullptr, *band = nullptr, *band3 = nullptr;  // band3: SSP_RK4's fourth stage
    double4 *dt_ring = nullptr;  // copy boundaries: ghost cells the next Setdt reads once (gpm_set_dt_ghosts)
    bool dt_ring_once = false;
    StepExtra *ext = nullptr;  // [kMaxStages] in device memory (plan_stages' per-stage extras)
    order_t *gpo = nullptr;
    uint8_t *dflag = nullptr;  // the first pass's record of the stage-input parts of each cell's decision
    int *slistA = nullptr, *slistB = nullptr, *mark = nullptr;
    int *counters = nullptr, *stamp = nullptr;
    DevStats *stats = nullptr;
    double *tstate = nullptr, *block_min = nullptr, *dt_log = nullptr;
    unsigned *done = nullptr;
    int dt_log_cap = 0;
    int tiles_x = 0, ntiles = 0, dt_blocks = 0, mood_blocks = 0, mood_solo = 0;
    bool mood_coop = true;  // MOOD passes as one persistent launch per stage (else per-pass launches)
    bool mood_fc = true;    // workgroup 0's MOOD passes reuse face fluxes and compute likely next ones ahead
    // wide launches of long-list passes ahead of the persistent kernel: launches per stage, grid, shortest list
    int mood_wide = 0, mood_wide_blocks = 0, mood_wide_min = 0;
    bool mood_wide_on = true;  // the last call's first lists were long (or there was no call yet)
    // the persistent MOOD kernel as an ordinary launch with its own rendezvous
    // and grid barrier (else a cooperative launch); its rendezvous polls, and
    // the test switch that sends every pass to one workgroup
    bool mood_ord = false;
    int mood_rv_spins = 1 << 14;
    int mood_bar_spins = 1 << 21;  // about a second of polls; a pass takes milliseconds at most
    bool mood_rv_alone = false;
    bool mood_alone_told = false;
    int mood_chk[3] = {0, 0, 0};  // host copy of DevStats::long_list, mood_alone, mood_stall
    FluxTable gfc{};        // device-wide face-flux cache that the other workgroups fill meanwhile
    bool has_state = false;
    bool timing = false;
    int timing_every = 1;     // HIP events around the kernels of every timing_every-th step
    long long timing_step = 0;
    double stage_ms = 0, mood_ms = 0;
    long long stage_launches = 0, mood_launches = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_pairs;  // (start index, kind 0 stage / 1 mood)
    size_t ev_used = 0;
    std::string msg;
    // the reference's own network (mod_NN.f90) in device memory, Fortran layouts:
    // weight0 | bias0 | weight1 | bias1 | weight2 | bias2
    float *refnn_w = nullptr;
    int refnn_L = 0, refnn_len = 0;
    bool refnn_cc = false;           // method NN_GP_MOOD_CC: softmax and the blended reconstruction
    int8_t *priori = nullptr;        // CellGPO_priori of the last stage [NY][NX]
    bool priori_set = false;         // the last step ran the network (else CellGPO_priori = Mord)
    float *rowv = nullptr, *vnn = nullptr;  // NN_GP_MOOD_CC: values of each row's last cell; values_NN
    long long niter = 0;             // main.f90's niter: steps started so far
    // training samples in the reference network's format (not owned, but nr_start)
    RefDumpArgs rdump{};
    bool rdump_on = false;
    // the reference's own dataset arrays (gpm_set_reference_dataset and
    // gpm_append_reference_dataset): the device rows of the last step (owned)
    struct HostDataset {
        int32_t L = 0;
        float *d_in = nullptr, *d_lab = nullptr;
        long long *d_cells = nullptr, *d_nr = nullptr;
        unsigned long long *d_count = nullptr;
    } hds;
    bool hds_on = false;
    // host-array exchange (allocated on first use): a ghosted state and a ghosted CellGPO
    double4 *xfer = nullptr;
    int32_t *xfer_gpo = nullptr;
    // learned a priori predictor (not owned)
    gpm_mlp *nn = nullptr;
    int nn_gate = 1;
    bool fixed_orders = false;  // a priori orders given by the caller (pred holds them)
    int8_t *pred = nullptr;
    int *nnlist = nullptr, *nncount = nullptr;
    // training-data dump
    float *dump_x = nullptr;
    int32_t *dump_y = nullptr;
    long long dump_cap = 0;
    unsigned long long *dump_count = nullptr;
    double dump_keep = 0.0;
    unsigned long long dump_seed = 0;
    int dump_gated = 1, dump_stages = 0;
    long long step_counter = 0;
    // order masks of every RK stage, recorded for the frozen-mask adjoint
    int8_t *rec_masks = nullptr;
    int rec_cap = 0, rec_count = 0;
    // adjoint work buffers (allocated on first use)
    double4 *adj_ub = nullptr, *adj_G = nullptr, *adj_A = nullptr, *adj_tmp = nullptr;
    int8_t *adj_ordx = nullptr, *adj_ordy = nullptr;
};

namespace {

int fail(gpm_ctx *c, int code, const std::string &m)
{
    if (c) c->msg = m;
    return code;
}

int hipfail(gpm_ctx *c, hipError_t e, const char *what)
{
    return fail(c, GPM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                          \
    do {                                                           \
        hipError_t e_ = (expr);                                    \
        if (e_ != hipSuccess) return hipfail((ctx), e_, #expr);    \
    } while (0)

template <class T>
hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T) > 0 ? n * sizeof(T) : 1);
}

// Interior cells of a ghosted host state [NY+8][NX+8][4].
struct GhostedSrc {
    const double *U;
    int W;  // NX + 8
    double4 operator()(int i, int j) const
    {
        const double *p = U + ((size_t)(j + 4) * W + (i + 4)) * 4;
        return make_double4(p[0], p[1], p[2], p[3]);
    }
};

// The ghost cells of a ghosted host state by Boundary_C applied to its own
// interior, at the time whose DMR switch abscissa is xmin_top (the rule the
// kernels apply; it reads interior cells only).
void fill_ghosts(double *U, const Bnd &B, double xmin_top)
{
    const int NX = B.NX, NY = B.NY, W = NX + 8;
    const GhostedSrc src{U, W};
    for (int r = 0; r < ring_size(NX, NY); ++r) {
        int i, j;
        ring_cell(r, NX, NY, i, j);
        const double4 v = bnd_value(B, src, i, j, xmin_top);
        double *d = U + ((size_t)(j + 4) * W + (i + 4)) * 4;
        d[0] = v.x; d[1] = v.y; d[2] = v.z; d[3] = v.w;
    }
}



// The boundary descriptor of a configuration, with its fixed ghost states in
// st[kBndStates]: the formula rules' states by primitive_to_conservative
// (BC.f90:60-381, IC.f90:336-355).
Bnd make_bnd(int NX, int NY, int bc, double dx, double dy, double4 *st)
{
    Bnd B{};
    B.NX = NX; B.NY = NY; B.bc = bc; B.dx = dx; B.dy = dy;
    B.sqrt3 = std::sqrt(3.);
    for (int k = 0; k < kBndStates; ++k) st[k] = make_double4(0, 0, 0, 0);
    switch (bc) {
    case kMach800:
        st[0] = p2c(1.4, 0.0, 800.0, 1.0);
        break;
    case kDoubleMach800:
        st[0] = p2c(1.4, 0.0, 800.0, 1.0);
        st[1] = p2c(1.4, 0.0, -800.0, 1.0);
        break;
    case kDMR:
        st[0] = p2c(8., 7.1447096, -4.125, 116.5);
        st[1] = p2c(1.4, 0.0, 0.0, 1.0);
        break;
    case kRMI:
        st[0] = p2c(2.666666666666667, 2. * std::sqrt(1.4), 0., 4.5);
        break;
    case kDirichlet:
        st[0] = p2c(3.857143, 2.629369, 0., 10.33333);
        for (int g = 1; g <= 4; ++g) st[g] = p2c(1. + 0.2 * std::sin(5 * (mesh_c(NX - 1 + g, dx) - 4.5)), 0., 0., 1.);
        break;
    default:
        break;
    }
    B.st = st;
    return B;
}

int upload_tables(gpm_ctx *c, const double *zT, const double *d2x, const double *d2y, const double *gw,
                  const int32_t *ixiy)
{
    const int mord = c->mord, ngp = c->ngp, szm = stencil_size(mord);
    // The stencil gather order is compiled into the kernels; tables must use it.
    for (int ord = 1; ord <= mord; ord += 2)
        for (int s = 0; s < stencil_size(ord); ++s)
            if (ixiy[(ord - 1) + 7 * (s + 25 * 0)] != kStencilX[s] || ixiy[(ord - 1) + 7 * (s + 25 * 1)] != kStencilY[s])
                return fail(c, GPM_EINVAL, "ixiy_sp does not follow the sphere stencil order of GP_init.f90:97-176");
    DevTables &h = c->hT;
    std::memset(&h, 0, sizeof(h));
    for (int ord = 1; ord <= mord; ord += 2)
        for (int i = 0; i < 4; ++i)
            for (int s = 0; s < stencil_size(ord); ++s)
                for (int j = 0; j < ngp; ++j)
                    h.zT[(ord - 1) / 2][i][s][j] = zT[(ord - 1) + mord * (s + szm * (i + 4 * j))];
    if (mord >= 3)
        for (int k = 0; k < 5; ++k) {
            h.d2x[k] = d2x[2 + mord * k];
            h.d2y[k] = d2y[2 + mord * k];
        }
    for (int j = 0; j < ngp; ++j) h.gw[j] = gw[(ngp - 1) + 5 * j];
    if (c->cfg.method == GPM_POL_MOOD || c->cfg.method == GPM_UNLIM_POL) {
        // POL_MOOD and Unlim_POL reconstruct order 3 with Pol_zT_o3 (pol_presim, GP_init.f90:359-411)
        // over the same five stencil cells: o3_coef (constants.f90:59-61) placed
        // per face and Gauss point as (M, L, T, R, B)
        const double cf[5] = {15. / 18, 1. / 3, 1. / (4 * std::sqrt(3.)), -1. / (4 * std::sqrt(3.)), -1. / 6};
        enum { M = 0, CN = 1, CT = 2, FT = 3, FN = 4 };  // kM, kCloseNrm, kCloseTrans, kFarTrans, kFarNrm
        static const int tab[4][2][5] = {
            {{M, CN, FT, FN, CT}, {M, CN, CT, FN, FT}},  // iL
            {{M, CT, CN, FT, FN}, {M, FT, CN, CT, FN}},  // iT
            {{M, FN, FT, CN, CT}, {M, FN, CT, CN, FT}},  // iR
            {{M, CT, FN, FT, CN}, {M, FT, FN, CT, CN}},  // iB
        };
        for (int f = 0; f < 4; ++f)
            for (int g = 0; g < 2; ++g)
                for (int k = 0; k < 5; ++k) h.zT[1][f][k][g] = cf[tab[f][g][k]];
    }
    HIPCHK(c, hipMemcpyAsync(c->dT, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

hipEvent_t next_event(gpm_ctx *c)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

int collect_timing(gpm_ctx *c)
{
    if (!c->timing || c->ev_pairs.empty()) return GPM_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &pr : c->ev_pairs) {
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[pr.first], c->ev_pool[pr.first + 1]));
        if (pr.second == 0) { c->stage_ms += ms; ++c->stage_launches; }
        else { c->mood_ms += ms; ++c->mood_launches; }
    }
    c->ev_pairs.clear();
    c->ev_used = 0;
    return GPM_OK;
}

int launch_dt_kernel(gpm_ctx *c, int mode, int use_tmax, double tmax, double dtfinal, double dt_given,
                     int advance_t, int step_idx, bool log_dt, int new_step)
{
    DtArgs a{};
    a.U = c->dU;
    a.ring = c->ring_mode ? c->ring : (c->dt_ring_once && mode == kDtCompute ? c->dt_ring : nullptr);
    if (mode == kDtCompute) c->dt_ring_once = false;
    a.block_min = c->block_min;
    a.done = c->done;
    a.tstate = c->tstate;
    a.dt_log = log_dt ? c->dt_log : nullptr;
    a.counters = c->counters;
    a.ncounters = kMaxStages * kCountersPerStage;
    a.stats = c->stats;
    a.NX = c->NX;
    a.NY = c->NY;
    a.mode = mode;
    a.use_tmax = use_tmax;
    a.advance_t = advance_t;
    a.step_idx = step_idx;
    a.new_step = new_step;
    a.dx = c->dx;
    a.dy = c->dy;
    a.cfl = c->cfg.cfl;
    a.tmax = tmax;
    a.dtfinal = dtfinal;
    a.dt_given = dt_given;
    HIPCHK(c, launch_dt(a, mode == kDtCompute ? c->dt_blocks : 1, c->stream));
    return GPM_OK;
}

// One stage of the integrator: input, output and how the output is formed.
struct StagePlan {
    const double4 *in;
    double4 *out;
    StepCtl C;
};

// The stages of one step of the configured integrator (time_stepper.f90:17-21)
// on the context's buffers; the step's result lands in dA (swapped with dU
// afterwards).  Returns the stage count.  With `extra`, also the per-stage
// StepExtra (buffers that do not change from step to step).
int plan_stages(gpm_ctx *c, StagePlan *P, StepExtra *extra = nullptr)
{
    StepExtra ex[kMaxStages] = {};
    auto st = [&](int k, const double4 *in, double4 *out, int kind, int trk) {
        P[k] = StagePlan{in, out, StepCtl{}};
        P[k].C.Un = c->dU;
        P[k].C.ext = c->ext ? c->ext + k : nullptr;
        P[k].C.kind = kind;
        P[k].C.trk = trk;
        P[k].C.gravity = c->gravity;
    };
    int n = 0;
    switch (c->integrator) {
    case GPM_FE:  // mod_FE.f90:20-190
        st(n++, c->dU, c->dA, kCmbCopy, 0);
        break;
    case GPM_SSP_RK2:  // mod_SSP_RK2.f90:12-22
        st(n++, c->dU, c->dB, kCmbCopy, 0);
        st(n++, c->dB, c->dA, kCmbRK2b, 0);
        break;
    case GPM_SSP_RK4:  // mod_SSP_RK4.f90:12-34: U1 dB, U2 dC, U3 dD, U4 dE, FU3 dF
        st(n++, c->dU, c->dB, kCmbRK4a, 0);
        st(n++, c->dB, c->dC, kCmbRK4b, 0);
        st(n++, c->dC, c->dD, kCmbRK4c, 0);
        st(n++, c->dD, c->dE, kCmbRK4d, 0);
        ex[n - 1].W = c->dF;
        st(n++, c->dE, c->dA, kCmbRK4e, 0);
        ex[n - 1].X1 = c->dC;
        ex[n - 1].X2 = c->dD;
        ex[n - 1].X3 = c->dF;
        break;
    default:  // SSP_RK3, mod_SSP_RK3.f90:12-39
        st(n++, c->dU, c->dA, kCmbCopy, 0);
        st(n++, c->dA, c->dB, kCmbRK3b, 1);
        st(n++, c->dB, c->dA, kCmbRK3c, 2);
    }
    if (c->ring_mode && !c->nodetect) {
        ex[n - 1].band = c->band;
        P[n - 1].C.band = 1;
        if (c->integrator == GPM_SSP_RK4) {  // FU3's ghosts come from the fourth stage
            ex[n - 2].band = c->band3;
            P[n - 2].C.band = 1;
        }
    }
    if (extra)
        for (int k = 0; k < kMaxStages; ++k) extra[k] = ex[k];
    return n;
}

// Arguments of k_stage for stage `k` (0-based) of plan P.
StageArgs stage_args(gpm_ctx *c, int k, const StagePlan &P)
{
    int *cnt = c->counters + k * kCountersPerStage;
    StageArgs s{};
    s.Uin = P.in; s.Uout = P.out; s.C = P.C; s.B = c->B; s.gpo = c->gpo; s.dflag = c->dflag;
    s.slist1 = c->slistB; s.cnt_s1 = cnt + kCntS0 + 1; s.cnt_f0 = cnt + kCntFail0;
    s.mark = c->mark; s.stamp = c->stamp; s.T = c->dT; s.tstate = c->tstate;
    s.dmp = c->cfg.dmp; s.u2 = c->cfg.u2; s.u2tol = c->cfg.u2_tol;
    s.cascade = c->cfg.cascade;
    s.recon = c->cfg.method == GPM_POL_MOOD ? kReconPol : kReconGP;
    // GP_SE and Unlim_POL: every face at order Mord, no detection
    s.frozen = c->cfg.method == GPM_GP_SE || c->cfg.method == GPM_UNLIM_POL;
    s.dx = c->dx; s.dy = c->dy; s.tiles_x = c->tiles_x; s.ntiles = c->ntiles;
    return s;
}

// Ghost ring of the step's output (dA) from the step-start state (dU) and the
// last stage's band values, as the integrator's whole-array statements leave it.
int enqueue_ring(gpm_ctx *c)
{
    RingArgs r{};
    r.A = c->dU; r.band = c->band; r.ring = c->ring; r.tstate = c->tstate; r.B = c->B; r.gravity = c->gravity;
    r.T = c->dT; r.flux = c->cfg.num_flux;
    const int integ = c->integrator == GPM_FE        ? kRingFE
                      : c->integrator == GPM_SSP_RK2 ? kRingRK2
                      : c->integrator == GPM_SSP_RK4 ? kRingRK4
                                                     : kRingRK3;
    if (c->nodetect) {
        // the last stage's input: Un (FE), U1 (SSP_RK2) or U2 (SSP_RK3), still in its buffer
        StagePlan plan[kMaxStages];
        const int nst = plan_stages(c, plan);
        r.mode = kRingFace;
        r.integ = integ;
        r.Bin = plan[nst - 1].in;
        HIPCHK(c, launch_ring_face(c->fog ? 1 : c->mord, c->fog ? 1 : c->ngp, c->cfg.num_flux, r, c->stream));
        return GPM_OK;
    }
    r.mode = integ;
    if (integ == kRingRK4) {  // U2, U3, U4 still in their buffers (plan_stages)
        r.X2 = c->dC; r.X3 = c->dD; r.X4 = c->dE; r.band3 = c->band3;
    }
    HIPCHK(c, launch_ring(r, c->stream));
    return GPM_OK;
}

// Ghost ring of the current state dU by Boundary_C at t = tstate[0].
int enqueue_ring_fill(gpm_ctx *c)
{
    if (!c->ring_mode) return GPM_OK;
    RingArgs r{};
    r.A = c->dU; r.ring = c->ring; r.tstate = c->tstate; r.B = c->B; r.mode = kRingFill;
    HIPCHK(c, launch_ring(r, c->stream));
    return GPM_OK;
}

// Wide launches of the MOOD passes over long lists (k_mood_wide), ahead of the
// persistent kernel of each stage: as many workgroups as the GPU holds at
// once, for lists longer than one trip of the persistent grid.
// GPM_MOOD_WIDE sets the launches per stage (0 switches them off),
// GPM_MOOD_WIDE_MIN the shortest list they take.
void setup_mood_wide(gpm_ctx *c, int ncu, int flux)
{
    c->mood_wide = 0;
    // FOG runs no MOOD pass (its stage kernel is the first-order one)
    const void *k = c->fog || c->nodetect ? nullptr : mood_wide_kernel(c->mord, c->ngp, flux, c->ring_mode);
    if (!k) return;
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, kWideThreads, 0) != hipSuccess) {
        (void)hipGetLastError();
        per = 1;
    }
    if (per <= 0) per = 1;
    c->mood_wide_blocks = per * ncu;
    const int ngpp = c->ngp <= 1 ? 1 : c->ngp <= 2 ? 2 : c->ngp <= 4 ? 4 : 8;
    c->mood_wide_min = c->mood_blocks * (kMoodThreads / (4 * ngpp));
    c->mood_wide = 2;
    if (const char *env = std::getenv("GPM_MOOD_WIDE")) c->mood_wide = std::atoi(env) > 0 ? std::atoi(env) : 0;
    if (const char *env = std::getenv("GPM_MOOD_WIDE_MIN")) c->mood_wide_min = std::atoi(env) > 0 ? std::atoi(env) : 0;
}

// GPM_MOOD_ORD=1: the persistent MOOD kernel as an ordinary launch (0, the
// default: a cooperative launch); GPM_MOOD_RV_SPINS: its rendezvous polls;
// GPM_MOOD_RV_ALONE=1: one workgroup runs every stage's passes (tests);
// GPM_MOOD_BAR_SPINS: the polls of its grid-barrier waits (tests of a stall).
void mood_ord_env(gpm_ctx *c)
{
    if (const char *env = std::getenv("GPM_MOOD_BAR_SPINS")) c->mood_bar_spins = std::atoi(env) > 0 ? std::atoi(env) : 0;
    if (const char *env = std::getenv("GPM_MOOD_ORD")) c->mood_ord = std::atoi(env) != 0;
    if (const char *env = std::getenv("GPM_MOOD_RV_SPINS")) c->mood_rv_spins = std::atoi(env) > 0 ? std::atoi(env) : 0;
    if (const char *env = std::getenv("GPM_MOOD_RV_ALONE")) c->mood_rv_alone = std::atoi(env) != 0;
}

// An idle wide launch still costs a few microseconds per stage, so a context
// enqueues them only while its lists are long: the device records the longest
// first or second correction list (DevStats::long_list), a call that ends
// synchronised reads it (mood_wide_read before the synchronisation,
// mood_wide_decide after it), and the next call follows it.  Results do not
// depend on the launch plan.  The same read brings the ordinary persistent
// launches' reports: stages run by one workgroup (said once on stderr) and
// grid-barrier waits that reached their bound (the call fails).
static_assert(offsetof(DevStats, mood_alone) == offsetof(DevStats, long_list) + sizeof(int) &&
                  offsetof(DevStats, mood_stall) == offsetof(DevStats, long_list) + 2 * sizeof(int),
              "the MOOD launch reports are read as one block");
int mood_wide_read(gpm_ctx *c)
{
    if (!c->mood_coop) return GPM_OK;
    HIPCHK(c, hipMemcpyAsync(c->mood_chk, &c->stats->long_list, sizeof(c->mood_chk), hipMemcpyDeviceToHost, c->stream));
    return GPM_OK;
}
int mood_wide_decide(gpm_ctx *c)
{
    if (!c->mood_coop) return GPM_OK;
    if (c->mood_wide > 0) c->mood_wide_on = c->mood_chk[0] > c->mood_wide_min;
    HIPCHK(c, hipMemsetAsync(&c->stats->long_list, 0, sizeof(c->mood_chk), c->stream));
    if (c->mood_chk[2] > 0) return fail(c, GPM_EHIP, "a grid barrier of the MOOD passes did not complete");
    if (c->mood_chk[1] > 0 && !c->mood_alone_told) {
        c->mood_alone_told = true;
        std::fprintf(stderr,
                     "gpm: the MOOD passes of %d stage(s) ran on one workgroup (the workgroups of the persistent "
                     "launch were not resident together)\n",
                     c->mood_chk[1]);
    }
    return GPM_OK;
}

// The MOOD correction passes of one stage.  Persistent (the default): one
// launch of k_mood runs every pass, with grid barriers between passes; it is
// a cooperative launch, whose barrier is the hardware's (the default), or an
// ordinary launch whose workgroups meet at a rendezvous first (GPM_MOOD_ORD=1,
// and when the device refuses a cooperative launch, e.g. while another
// process holds that resource).  Otherwise (when asked for) per-pass
// launches in growing batches, with a look at the stage's pending flag
// between batches: a launch runs the current pass or one run of workgroup-0
// passes and completes at least one pass, and launches past the last pass
// return at once.
int enqueue_mood(gpm_ctx *c, MoodArgs &m)
{
    const int mord = c->mord, ngp = c->ngp, flux = c->cfg.num_flux;
    m.sync = kMoodCoop;
    m.rv_spins = c->mood_rv_spins;
    m.rv_alone = c->mood_rv_alone ? 1 : 0;
    m.bar_spins = c->mood_bar_spins;
    if (c->mood_coop) {
        void *args[] = {&m};
        if (c->mood_wide > 0 && c->mood_wide_on) {
            // the first passes over long lists in ordinary launches of the whole
            // GPU; each returns at once when its pass lists few cells
            m.wide_min = c->mood_wide_min;
            const void *wk = mood_wide_kernel(mord, ngp, flux, c->ring_mode);
            for (int w = 0; w < c->mood_wide; ++w)
                HIPCHK(c, hipLaunchKernel(wk, dim3(c->mood_wide_blocks), dim3(kWideThreads), args, 0, c->stream));
        }
        const void *k = mood_kernel(mord, ngp, flux, c->ring_mode, true);
        if (!c->mood_ord) {
            const hipError_t e =
                hipLaunchCooperativeKernel(k, dim3(c->mood_blocks), dim3(kMoodThreads), args, 0, c->stream);
            if (e == hipSuccess) return GPM_OK;
            (void)hipGetLastError();
            std::fprintf(stderr, "gpm: cooperative launch refused (%s); the MOOD kernel runs as an ordinary launch\n",
                         hipGetErrorString(e));
            c->mood_ord = true;
        }
        m.sync = kMoodOrd;
        HIPCHK(c, hipLaunchKernel(k, dim3(c->mood_blocks), dim3(kMoodThreads), args, 0, c->stream));
        return GPM_OK;
    }
    const void *k = mood_kernel(mord, ngp, flux, c->ring_mode, false);
    void *args[] = {&m};
    // a stage runs at most 3 x cells + 1 passes (every pass but the last lowers a cell)
    const long long max_launches = 3ll * c->NX * c->NY + 2;
    long long launched = 0;
    for (int batch = 2;; batch = batch < 64 ? 2 * batch : 64) {
        for (int b = 0; b < batch; ++b)
            HIPCHK(c, hipLaunchKernel(k, dim3(c->mood_blocks), dim3(kMoodThreads), args, 0, c->stream));
        launched += batch;
        int pending = 0;
        HIPCHK(c, hipMemcpyAsync(&pending, m.cnt + kCntPending, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (!pending) break;
        if (launched > max_launches) return fail(c, GPM_EHIP, "MOOD passes did not end");
    }
    HIPCHK(c, launch_mood_finish(m, c->stream));
    return GPM_OK;
}

// nsteps_with_no_NN (global_variables.f90:97): the reference network predicts
// from step 11 on; earlier steps run plain GP-MOOD with the configured DMP.
constexpr int kStepsWithNoNN = 10;

// One step of the configured integrator after k_dt has set dt.  A step first
// advances niter (main.f90:147).
int enqueue_stages(gpm_ctx *c)
{
    StagePlan plan[kMaxStages];
    const int nst = plan_stages(c, plan);
    ++c->niter;
    // the steps timed with HIP events (every step, or a sample)
    const bool timed = c->timing && c->timing_step++ % c->timing_every == 0;
    // the reference network's steps: a priori orders from the network and the
    // PAD-only check NN_DETECTION (mod_FE.f90:51-56,151-158)
    const bool refnn = c->refnn_w && c->niter > kStepsWithNoNN;
    c->priori_set = refnn;
    for (int k = 0; k < nst; ++k) {
        const int stage = k + 1;
        const double4 *in = plan[k].in;
        int *cnt = c->counters + k * kCountersPerStage;
        if (refnn) {
            RefNnArgs r{};
            r.U = in; r.w = c->refnn_w; r.pred = c->pred; r.priori = c->priori;
            r.rowv = c->refnn_cc ? c->rowv : nullptr; r.stats = c->stats; r.B = c->B;
            r.L = c->refnn_L; r.len = c->refnn_len; r.mord = c->mord; r.softmax = c->refnn_cc;
            HIPCHK(c, launch_refnn(r, c->stream));
            if (c->refnn_cc) HIPCHK(c, launch_refnn_values(c->vnn, c->rowv, c->NX, c->NY, c->refnn_L, true, c->stream));
        }
        if (c->nn) {
            // a priori orders of this stage's input (compute_CellGPO_with_NN, mod_NN.f90:14-58)
            HIPCHK(c, hipMemsetAsync(c->nncount, 0, sizeof(int), c->stream));
            HIPCHK(c, mlp_launch_gate(in, c->N, c->bc, c->dx, c->dy, c->nn_gate, c->mord, c->gpo, c->pred, c->nnlist,
                                      c->nncount, c->stream));
            HIPCHK(c, mlp_launch_predict(in, c->N, c->bc, c->nnlist, c->nncount, mlp_device_view(c->nn),
                                         mlp_stencil(c->nn), mlp_classes(c->nn), c->mord, c->gpo, c->pred,
                                         &c->stats->a_priori_lowered, c->stream));
        }
        if (c->fixed_orders && !c->nn)
            HIPCHK(c, launch_orders_from_pred(c->pred, c->gpo, (size_t)c->NX * c->NY, c->stream));
        StageArgs s = stage_args(c, k, plan[k]);
        s.apriori = c->nn != nullptr || c->fixed_orders || refnn;
        s.pred = c->pred;
        if (refnn) {
            s.dmp = 0;
            if (c->refnn_cc) {  // NN_GP_MOOD_CC_ replaces GP_MOOD_ (mod_Reconstruction.f90:26-32)
                s.vnn = c->vnn;
                s.recon = kReconCC;
            }
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) {
            size_t i0 = c->ev_used;
            e0 = next_event(c); e1 = next_event(c);
            if (!e0 || !e1) return fail(c, GPM_EHIP, "hipEventCreate failed");
            c->ev_pairs.push_back({(int)i0, 0});
            HIPCHK(c, hipEventRecord(e0, c->stream));
        }
        // a multiple of 8 workgroups, so the XCD-aware remap in k_stage is a bijection;
        // FOG runs the first-order kernel (Mord 1 template, one Gauss point)
        HIPCHK(c, launch_stage(c->fog ? 1 : c->mord, c->fog ? 1 : c->ngp, c->cfg.num_flux, c->ring_mode, s,
                               (c->ntiles + 7) / 8 * 8, c->stream));
        if (timed) HIPCHK(c, hipEventRecord(e1, c->stream));
        if (c->nodetect) continue;  // one pass, no detection (mod_FE.f90:160-162)

        MoodArgs m{};
        m.Uin = in; m.Uout = plan[k].out; m.C = plan[k].C; m.B = c->B; m.gpo = c->gpo; m.dflag = c->dflag;
        m.slist[0] = c->slistA; m.slist[1] = c->slistB;
        m.cnt = cnt; m.mark = c->mark; m.stamp = c->stamp; m.solo_max = c->mood_solo; m.fcache = c->mood_fc;
        m.gfc = c->gfc;
        m.stats = c->stats; m.T = c->dT; m.tstate = c->tstate;
        m.dmp = c->cfg.dmp; m.u2 = c->cfg.u2; m.u2tol = c->cfg.u2_tol;
        m.cascade = c->cfg.cascade; m.stage_idx = k;
        m.dx = c->dx; m.dy = c->dy;
        m.recon = s.recon;
        if (refnn) {
            m.dmp = 0;
            m.nn_passes = 1;
            if (c->refnn_cc) m.vnn = c->vnn;  // m.recon is kReconCC from s
        }
        if (timed) {
            size_t i0 = c->ev_used;
            e0 = next_event(c); e1 = next_event(c);
            if (!e0 || !e1) return fail(c, GPM_EHIP, "hipEventCreate failed");
            c->ev_pairs.push_back({(int)i0, 1});
            HIPCHK(c, hipEventRecord(e0, c->stream));
        }
        int rc = enqueue_mood(c, m);
        if (rc) return rc;
        if (timed) HIPCHK(c, hipEventRecord(e1, c->stream));
        if (c->rec_masks && c->rec_count < c->rec_cap) {
            const size_t n = (size_t)c->NX * c->NY;
            HIPCHK(c, launch_unpack_orders(c->gpo, c->rec_masks + ((size_t)c->rec_count * 3 + k) * n, n, c->stream));
        }
        if (c->rdump_on && k == 0) {
            // append_to_NN_datatset after the first stage (mod_FE.f90:176-181)
            RefDumpArgs d = c->rdump;
            d.U = in; d.gpo = c->gpo; d.tstate = c->tstate; d.B = c->B; d.step = c->step_counter;
            HIPCHK(c, launch_refnn_dump(d, c->mord, c->stream));
        }
        if (c->dump_x && (c->dump_stages & (1 << k))) {
            // (stencil of the stage input, final order of the stage) pairs
            const int S = 2 * ((c->mord - 1) / 2) + 1;
            HIPCHK(c, mlp_launch_dump(in, c->N, c->bc, c->dx, c->dy, c->gpo, c->mord, S, (float)c->dump_keep,
                                      c->dump_seed ^ (0x632BE59BD9B4E019ull * (unsigned long long)(c->step_counter * 3 + stage)),
                                      c->dump_gated, c->dump_x, c->dump_y, c->dump_cap, c->dump_count, c->stream));
        }
    }
    if (c->ring_mode) {
        int rc = enqueue_ring(c);
        if (rc) return rc;
    }
    std::swap(c->dU, c->dA);
    ++c->step_counter;
    if (c->rec_masks && c->rec_count < c->rec_cap) ++c->rec_count;
    return GPM_OK;
}

// The learned predictor, the training-data dump and the differentiable form
// are built for square grids with a copy boundary, SSP_RK3 and no source term.
bool square_copy_rk3(const gpm_ctx *c)
{
    return c->NX == c->NY && index_map_bc(c->bc) && !c->gravity && c->integrator == GPM_SSP_RK3;
}

// ---- allocation-free runs (gpm_run_noalloc) ---------------------------------
// Every device buffer of such a run lives in `scratch`, one static array per
// GPU, laid out as below; its runs are limited to kNaMaxCells cells.
constexpr size_t kNaMaxCells = (size_t)1 << 20;
constexpr int kNaMaxDtBlocks = 1024;
constexpr size_t na_align(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t kNaTables = 0;
constexpr size_t kNaCounters = na_align(kNaTables + sizeof(DevTables));
constexpr size_t kNaStats = na_align(kNaCounters + sizeof(int) * kMaxStages * kCountersPerStage);
constexpr size_t kNaTstate = na_align(kNaStats + sizeof(DevStats));
constexpr size_t kNaBlockMin = na_align(kNaTstate + 8 * sizeof(double));
constexpr size_t kNaDone = na_align(kNaBlockMin + kNaMaxDtBlocks * sizeof(double));
constexpr size_t kNaStamp = na_align(kNaDone + sizeof(unsigned));
constexpr size_t kNaA = na_align(kNaStamp + sizeof(int));
constexpr size_t kNaB = kNaA + kNaMaxCells * sizeof(double4);
constexpr size_t kNaGpo = kNaB + kNaMaxCells * sizeof(double4);
constexpr size_t kNaListA = kNaGpo + kNaMaxCells * sizeof(order_t);
constexpr size_t kNaListB = kNaListA + kNaMaxCells * sizeof(int);
constexpr size_t kNaMark = kNaListB + kNaMaxCells * sizeof(int);
constexpr size_t kNaFlag = kNaMark + kNaMaxCells * sizeof(int);
constexpr size_t kNaBytes = kNaFlag + kNaMaxCells;

}  // namespace

// The scratch memory of gpm_run_noalloc (kNaBytes, about 80 MiB).  In host code
// this name is not a device address: host code uses the address that
// hipGetSymbolAddress returns for the current device (na_scratch).
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
unsigned char *g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on device `dev` (the current device), looked up
// once per device.
int na_scratch(int dev, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "gpm_run_noalloc: hipGetSymbolAddress: %s\n", hipGetErrorString(e));
            return GPM_EHIP;
        }
        g_na_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_na_ptr[dev];
    return GPM_OK;
}

}  // namespace

extern "C" {

int gpm_create(const gpm_config *cfg, gpm_ctx **out)
{
    if (!cfg || !out) return GPM_EINVAL;
    *out = nullptr;
    gpm_ctx *c = new gpm_ctx();
    c->cfg = *cfg;
    c->NX = cfg->lf;
    c->NY = cfg->nf;
    c->N = cfg->lf;
    c->mord = cfg->mord;
    c->ngp = cfg->ngp;
    c->bc = cfg->bc_type;
    c->integrator = cfg->integrator;
    c->gravity = cfg->problem == GPM_PROBLEM_RT;
    c->fog = cfg->method == GPM_FOG;
    c->nodetect = c->fog || cfg->method == GPM_GP_SE || cfg->method == GPM_UNLIM_POL;
    auto bad = [&](int code, const char *m) {
        std::fprintf(stderr, "gpm_create: %s\n", m);
        delete c;
        return code;
    };
    const int bcs[] = {GPM_NEUMANN, GPM_PERIODIC, GPM_DIRICHLET, GPM_REFLECTIVE, GPM_KH_BC, GPM_MACH800_BC,
                       GPM_DOUBLEMACH800_BC, GPM_RMI_BC, GPM_DMR_BC, GPM_RT_BC};
    bool bc_ok = false;
    for (int b : bcs) bc_ok = bc_ok || cfg->bc_type == b;
    if (cfg->lf < 8 || cfg->nf < 8) return bad(GPM_EINVAL, "grids need lf, nf >= 8");
    if (cfg->integrator != GPM_FE && cfg->integrator != GPM_SSP_RK2 && cfg->integrator != GPM_SSP_RK3 &&
        cfg->integrator != GPM_SSP_RK4)
        return bad(GPM_EUNSUP, "integrator must be FE, SSP_RK2, SSP_RK3 or SSP_RK4");
    if (cfg->num_flux != GPM_HLLC && cfg->num_flux != GPM_LLF && cfg->num_flux != GPM_HLL &&
        cfg->num_flux != GPM_UNSPLIT)
        return bad(GPM_EUNSUP, "numFlux must be HLLC, LLF, HLL or unsplit");
    const bool nn_method = cfg->method == GPM_NN_GP_MOOD || cfg->method == GPM_NN_GP_MOOD_CC;
    const bool pol = cfg->method == GPM_POL_MOOD || cfg->method == GPM_UNLIM_POL;
    if (cfg->method != GPM_GP_MOOD && !nn_method && cfg->method != GPM_FOG && !pol && cfg->method != GPM_GP_SE)
        return bad(GPM_EUNSUP, "method must be GP_MOOD, NN_GP_MOOD, NN_GP_MOOD_CC, POL_MOOD, Unlim_POL, GP_SE or FOG");
    // Pol_zT_o3 holds 2 Gauss points and the polynomial methods have no order
    // above 3 (POL_MOOD_ takes the cell value at any other order,
    // mod_pol_MOOD.f90:62-76; Unlim_POL_ stops, mod_Unlim_Pol.f90:33-36)
    if (pol && (cfg->mord != 3 || cfg->ngp != 2))
        return bad(GPM_EUNSUP, "POL_MOOD and Unlim_POL need Mord = 3 and ngp = 2");
    if (c->fog && cfg->ngp != 1)
        return bad(GPM_EUNSUP, "FOG needs ngp = 1 (the problems/*_FOG.f90 set-up)");
    if (!bc_ok) return bad(GPM_EUNSUP, "unknown BC_type");
    // FOG leaves the face loop's updates in the first ghost layer (no Boundary_C
    // after the step), so its ghost cells are never copies of interior cells
    c->ring_mode = !index_map_bc(cfg->bc_type) || c->gravity || c->nodetect;
    if (c->nodetect && cfg->integrator == GPM_SSP_RK4)
        return bad(GPM_EUNSUP, "SSP_RK4 needs a method with DETECTION (not FOG, GP_SE or Unlim_POL)");
    // the network reads the stage input's ghost cells as copies of interior cells
    if (nn_method && (!index_map_bc(cfg->bc_type) || c->gravity))
        return bad(GPM_EUNSUP, "NN_GP_MOOD needs a copy boundary (Neumann, Periodic, Reflective, KH_BC) and no source term");
    if (cfg->method == GPM_NN_GP_MOOD_CC && cfg->mord != 3)
        return bad(GPM_EUNSUP, "NN_GP_MOOD_CC needs Mord = 3 (the reference stops at other orders, "
                               "mod_NN_GP_MOOD_CC.f90:57-59)");
    if (cfg->cascade != GPM_CASCADE_REFERENCE && cfg->cascade != GPM_CASCADE_FULL)
        return bad(GPM_EINVAL, "unknown cascade");
    if (!c->fog && !supported(cfg->mord, cfg->ngp)) return bad(GPM_EUNSUP, "(Mord, ngp) combination not compiled");
    if ((long long)cfg->lf * cfg->nf > (1ll << 31) - 1) return bad(GPM_EUNSUP, "grid too large for int32 cell ids");
    // dx = real(Lx_16/lf, 8) (global_variables.f90:19-21)
    c->dx = quad_ratio(cfg->lx, cfg->lf);
    c->dy = quad_ratio(cfg->ly, cfg->nf);
    c->Bh = make_bnd(c->NX, c->NY, c->bc, c->dx, c->dy, c->st_host);
    c->B = c->Bh;
    c->nstages = cfg->integrator == GPM_FE ? 1 : cfg->integrator == GPM_SSP_RK2 ? 2 : cfg->integrator == GPM_SSP_RK4 ? 5 : 3;
    const size_t n = (size_t)c->NX * c->NY;
    hipError_t e;
    if ((e = hipGetDevice(&c->device)) != hipSuccess) return bad(GPM_EHIP, hipGetErrorString(e));
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bad(GPM_EHIP, hipGetErrorString(e));
    bool ok = dalloc(&c->dT, 1) == hipSuccess && dalloc(&c->dU, n) == hipSuccess && dalloc(&c->dA, n) == hipSuccess &&
              dalloc(&c->dB, n) == hipSuccess && dalloc(&c->gpo, n) == hipSuccess && dalloc(&c->dflag, n) == hipSuccess &&
              dalloc(&c->slistA, n) == hipSuccess && dalloc(&c->slistB, n) == hipSuccess &&
              dalloc(&c->mark, n) == hipSuccess &&
              dalloc(&c->counters, kMaxStages * kCountersPerStage) == hipSuccess && dalloc(&c->stamp, 1) == hipSuccess &&
              dalloc(&c->stats, 1) == hipSuccess &&
              dalloc(&c->tstate, 8) == hipSuccess && dalloc(&c->done, 1) == hipSuccess &&
              dalloc(&c->pred, n) == hipSuccess && dalloc(&c->nnlist, n) == hipSuccess &&
              dalloc(&c->priori, n) == hipSuccess &&
              dalloc(&c->nncount, 1) == hipSuccess;
    if (ok && c->integrator == GPM_SSP_RK4)
        ok = dalloc(&c->dC, n) == hipSuccess && dalloc(&c->dD, n) == hipSuccess && dalloc(&c->dE, n) == hipSuccess &&
             dalloc(&c->dF, n) == hipSuccess;
    if (ok && c->ring_mode)
        ok = dalloc(&c->ring, (size_t)ring_size(c->NX, c->NY)) == hipSuccess && dalloc(&c->band, n) == hipSuccess;
    if (ok && c->ring_mode && c->integrator == GPM_SSP_RK4) ok = dalloc(&c->band3, n) == hipSuccess;
    if (ok) ok = dalloc(&c->st_dev, kBndStates) == hipSuccess && dalloc(&c->ext, kMaxStages) == hipSuccess;
    // The device-wide face-flux cache of the MOOD passes: a quarter slot per
    // cell (powers of two from 2^12 to 2^24 slots, 44 B each; 46 MB at 2048^2),
    // and a 4-byte claim word per cell.
    // GPM_MOOD_HELP=0 leaves it out (same results).
    bool help = !c->nodetect && c->mord > 1;
    if (const char *env = std::getenv("GPM_MOOD_HELP")) help = help && std::atoi(env) != 0;
    if (ok && help) {
        int bits = 12;
        while (bits < 24 && ((size_t)1 << bits) < n / 4) ++bits;
        const size_t slots = (size_t)1 << bits;
        ok = dalloc(&c->gfc.kw, slots) == hipSuccess && dalloc(&c->gfc.ready, slots) == hipSuccess &&
             dalloc(&c->gfc.data, slots) == hipSuccess && dalloc(&c->gfc.beacon, 2 + kSoloMax) == hipSuccess &&
             dalloc(&c->gfc.done, n) == hipSuccess;
        if (ok) {
            c->gfc.bits = bits;
            c->gfc.radius = 3;
            if (const char *env = std::getenv("GPM_MOOD_HELP_R")) c->gfc.radius = std::atoi(env);
            ok = hipMemsetAsync(c->gfc.kw, 0, slots * sizeof(*c->gfc.kw), c->stream) == hipSuccess &&
                 hipMemsetAsync(c->gfc.ready, 0, slots * sizeof(*c->gfc.ready), c->stream) == hipSuccess &&
                 hipMemsetAsync(c->gfc.beacon, 0, (2 + kSoloMax) * sizeof(int), c->stream) == hipSuccess &&
                 hipMemsetAsync(c->gfc.done, 0, n * sizeof(unsigned), c->stream) == hipSuccess;
        }
    }
    if (!ok) {
        gpm_destroy(c);
        std::fprintf(stderr, "gpm_create: device allocation failed\n");
        return GPM_ENOMEM;
    }
    int ncu = 0;
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device);
    if (ncu <= 0) ncu = 256;
    c->dt_blocks = ncu;
    if (const char *env = std::getenv("GPM_DT_BLOCKS")) c->dt_blocks = std::atoi(env) > 0 ? std::atoi(env) : c->dt_blocks;
    if (dalloc(&c->block_min, c->dt_blocks) != hipSuccess) {
        gpm_destroy(c);
        return GPM_ENOMEM;
    }
    // the tiles of the stage kernel that runs (FOG: the Mord-1 kernel)
    const int tx = stage_tx(c->fog ? 1 : c->mord), ty = stage_ty(c->fog ? 1 : c->mord);
    c->tiles_x = (c->NX + tx - 1) / tx;
    c->ntiles = c->tiles_x * ((c->NY + ty - 1) / ty);
    // Cooperative grid of the MOOD kernel: all its workgroups must be resident.
    // The per-pass grid barrier gets cheaper with fewer workgroups, while the
    // first correction passes (thousands of cells) want more lanes, and the
    // helpers of workgroup 0's passes need a few dozen: one workgroup of 512
    // threads per four CUs (64 on MI355X) measured best.  GPM_MOOD_BLOCKS
    // overrides the count.
    int want = ncu / 4 > 0 ? ncu / 4 : 1;
    if (const char *env = std::getenv("GPM_MOOD_BLOCKS")) want = std::atoi(env) > 0 ? std::atoi(env) : ncu;
    {
        int per = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, mood_kernel(c->fog ? 1 : c->mord, c->fog ? 1 : c->ngp, c->cfg.num_flux, c->ring_mode, true),
            kMoodThreads, 0);
        if (e != hipSuccess || per <= 0) per = 1;
        c->mood_blocks = want < per * ncu ? want : per * ncu;
    }
    setup_mood_wide(c, ncu, c->cfg.num_flux);
    // GPM_MOOD_COOP=0 selects the per-pass launches of the MOOD passes.
    if (const char *env = std::getenv("GPM_MOOD_COOP")) c->mood_coop = std::atoi(env) != 0;
    mood_ord_env(c);
    // Passes over at most this many cells run on one workgroup (GPM_MOOD_SOLO overrides; 0 disables).
    c->mood_solo = kSoloMax;
    if (const char *env = std::getenv("GPM_MOOD_SOLO")) c->mood_solo = std::atoi(env) > 0 ? std::atoi(env) : 0;
    if (c->mood_solo > kSoloMax) c->mood_solo = kSoloMax;
    // GPM_MOOD_FC=0 switches the face-flux cache of workgroup 0's passes off (same results).
    if (const char *env = std::getenv("GPM_MOOD_FC")) c->mood_fc = std::atoi(env) != 0;
    hipMemsetAsync(c->mark, 0, n * sizeof(int), c->stream);
    hipMemsetD32Async(c->gpo, pack_order(c->mord, c->mord, 0), n, c->stream);
    hipMemsetAsync(c->counters, 0, kMaxStages * kCountersPerStage * sizeof(int), c->stream);
    hipMemsetAsync(c->stats, 0, sizeof(DevStats), c->stream);
    hipMemsetAsync(c->tstate, 0, 8 * sizeof(double), c->stream);
    hipMemsetAsync(c->done, 0, sizeof(unsigned), c->stream);
    int one = 1;
    hipMemcpyAsync(c->stamp, &one, sizeof(int), hipMemcpyHostToDevice, c->stream);
    hipMemcpyAsync(c->st_dev, c->st_host, sizeof(c->st_host), hipMemcpyHostToDevice, c->stream);
    c->B.st = c->st_dev;
    StepExtra ext_host[kMaxStages];
    {
        StagePlan plan[kMaxStages];
        plan_stages(c, plan, ext_host);
    }
    hipMemcpyAsync(c->ext, ext_host, sizeof(ext_host), hipMemcpyHostToDevice, c->stream);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) {
        gpm_destroy(c);
        std::fprintf(stderr, "gpm_create: %s\n", hipGetErrorString(e));
        return GPM_EHIP;
    }
    // GP tables from the library's own precompute (GP_presim_sphere).
    std::vector<double> zT((size_t)c->mord * stencil_size(c->mord) * 4 * c->ngp), d2x(c->mord * stencil_size(c->mord)),
        d2y(d2x.size()), gw(25);
    std::vector<int32_t> ixiy(7 * 25 * 2);
    if (gp_tables(c->mord, c->ngp, cfg->ell_o_dx, cfg->lx, cfg->ly, cfg->lf, cfg->nf, zT.data(), d2x.data(),
                  d2y.data(), gw.data(), ixiy.data()) != 0) {
        gpm_destroy(c);
        return GPM_EINVAL;
    }
    int rc = upload_tables(c, zT.data(), d2x.data(), d2y.data(), gw.data(), ixiy.data());
    if (rc) {
        std::fprintf(stderr, "gpm_create: %s\n", c->msg.c_str());
        gpm_destroy(c);
        return rc;
    }
    *out = c;
    return GPM_OK;
}

int gpm_destroy(gpm_ctx *c)
{
    if (!c) return GPM_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    void *bufs[] = {c->dT, c->dU, c->dA, c->dB, c->dC, c->dD, c->dE, c->dF, c->ring, c->dt_ring, c->band, c->band3,
                    c->st_dev, c->ext,
                    c->gpo, c->dflag, c->slistA, c->slistB, c->mark, c->counters,
                    c->stamp, c->stats, c->tstate, c->block_min, c->dt_log, c->done, c->pred, c->nnlist,
                    c->nncount, c->adj_ub, c->adj_G, c->adj_A, c->adj_tmp, c->adj_ordx, c->adj_ordy,
                    c->priori, c->refnn_w, c->rowv, c->vnn, c->xfer, c->xfer_gpo, c->rdump.nr_start,
                    c->gfc.kw, c->gfc.ready, c->gfc.data, c->gfc.beacon, c->gfc.done};
    for (void *b : bufs)
        if (b) hipFree(b);
    void *hbufs[] = {c->hds.d_in, c->hds.d_lab, c->hds.d_cells, c->hds.d_nr, c->hds.d_count};
    for (void *b : hbufs)
        if (b) hipFree(b);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return GPM_OK;
}

const char *gpm_last_error(const gpm_ctx *c) { return c ? c->msg.c_str() : "null context"; }

int gpm_set_gp_tables(gpm_ctx *c, const double *zT, const double *d2x, const double *d2y, const double *gw,
                      const int32_t *ixiy)
{
    if (!c || !zT || !d2x || !d2y || !gw || !ixiy) return fail(c, GPM_EINVAL, "null argument");
    return upload_tables(c, zT, d2x, d2y, gw, ixiy);
}

int gpm_gp_tables(int32_t mord, int32_t ngp, double ell_o_dx, double lx, double ly, int32_t lf, int32_t nf,
                  double *zT, double *d2x, double *d2y, double *gw, int32_t *ixiy)
{
    if (!zT || !d2x || !d2y || !gw || !ixiy) return GPM_EINVAL;
    return gp_tables(mord, ngp, ell_o_dx, lx, ly, lf, nf, zT, d2x, d2y, gw, ixiy) == 0 ? GPM_OK : GPM_EINVAL;
}

int gpm_boundary_c(const gpm_config *cfg, double t_rk, double *U)
{
    if (!cfg || !U || cfg->lf < 8 || cfg->nf < 8) return GPM_EINVAL;
    const double dx = quad_ratio(cfg->lx, cfg->lf), dy = quad_ratio(cfg->ly, cfg->nf);
    double4 st[kBndStates];
    const Bnd B = make_bnd(cfg->lf, cfg->nf, cfg->bc_type, dx, dy, st);
    fill_ghosts(U, B, dmr_xmin_top(t_rk, B.sqrt3));
    return GPM_OK;
}

int gpm_set_state(gpm_ctx *c, const double *U)
{
    if (!c || !U) return fail(c, GPM_EINVAL, "null argument");
    const size_t n = (size_t)c->NX * c->NY;
    // the whole ghosted array in one contiguous copy, split on the device; in
    // ring mode the caller's ghost cells are what Setdt sees before the first step
    const size_t full = (size_t)(c->NX + 8) * (c->NY + 8);
    if (!c->xfer) HIPCHK(c, dalloc(&c->xfer, full));
    HIPCHK(c, hipMemcpyAsync(c->xfer, U, full * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_unpack_ghosted(c->xfer, c->dU, c->ring_mode ? c->ring : nullptr, c->NX, c->NY, c->stream));
    HIPCHK(c, hipMemsetD32Async(c->gpo, pack_order(c->mord, c->mord, 0), n, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->has_state = true;
    c->dt_ring_once = false;
    return GPM_OK;
}

int gpm_set_dt_ghosts(gpm_ctx *c, const double *U)
{
    if (!c || !U) return fail(c, GPM_EINVAL, "null argument");
    if (!c->has_state) return fail(c, GPM_ESTATE, "no state set");
    const size_t full = (size_t)(c->NX + 8) * (c->NY + 8);
    if (!c->xfer) HIPCHK(c, dalloc(&c->xfer, full));
    double4 *ring = c->ring;
    if (!c->ring_mode) {
        if (!c->dt_ring) HIPCHK(c, dalloc(&c->dt_ring, (size_t)ring_size(c->NX, c->NY)));
        ring = c->dt_ring;
    }
    HIPCHK(c, hipMemcpyAsync(c->xfer, U, full * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_unpack_ghosted(c->xfer, nullptr, ring, c->NX, c->NY, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!c->ring_mode) c->dt_ring_once = true;
    return GPM_OK;
}

int gpm_get_state(gpm_ctx *c, double *U, int32_t *cellgpo)
{
    if (!c || !U) return fail(c, GPM_EINVAL, "null argument");
    if (!c->has_state) return fail(c, GPM_ESTATE, "no state set");
    // ghosted on the device (ring, or the boundary rule for copy boundaries),
    // then one contiguous copy each for the state and CellGPO
    const size_t full = (size_t)(c->NX + 8) * (c->NY + 8);
    if (!c->xfer) HIPCHK(c, dalloc(&c->xfer, full));
    HIPCHK(c, launch_pack_ghosted(c->dU, c->ring_mode ? c->ring : nullptr, c->B, c->xfer, c->stream));
    HIPCHK(c, hipMemcpyAsync(U, c->xfer, full * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
    if (cellgpo) {
        if (!c->xfer_gpo) HIPCHK(c, dalloc(&c->xfer_gpo, full));
        // FOG leaves CellGPO = Mord (mod_FE.f90:48)
        HIPCHK(c, launch_cellgpo_ghosted(c->gpo, c->xfer_gpo, c->NX, c->NY, c->mord, c->fog, c->stream));
        HIPCHK(c, hipMemcpyAsync(cellgpo, c->xfer_gpo, full * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

int gpm_set_state_device(gpm_ctx *c, const double *u_dev)
{
    if (!c || !u_dev) return fail(c, GPM_EINVAL, "null argument");
    const size_t n = (size_t)c->NX * c->NY;
    HIPCHK(c, hipMemcpyAsync(c->dU, u_dev, n * 4 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemsetD32Async(c->gpo, pack_order(c->mord, c->mord, 0), n, c->stream));
    int rc = enqueue_ring_fill(c);  // ghost cells by Boundary_C at the current time
    if (rc) return rc;
    c->has_state = true;
    c->dt_ring_once = false;
    return GPM_OK;
}

int gpm_get_state_device(gpm_ctx *c, double *u_dev, int8_t *order_dev)
{
    if (!c) return GPM_EINVAL;
    if (!c->has_state) return fail(c, GPM_ESTATE, "no state set");
    const size_t n = (size_t)c->NX * c->NY;
    if (u_dev) HIPCHK(c, hipMemcpyAsync(u_dev, c->dU, n * 4 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (order_dev && c->fog) HIPCHK(c, hipMemsetAsync(order_dev, c->mord, n, c->stream));
    else if (order_dev) HIPCHK(c, launch_unpack_orders(c->gpo, order_dev, n, c->stream));
    return GPM_OK;
}

int gpm_compute_dt(gpm_ctx *c, double dtfinal, double *dt)
{
    if (!c || !dt) return fail(c, GPM_EINVAL, "null argument");
    if (!c->has_state) return fail(c, GPM_ESTATE, "no state set");
    int rc = launch_dt_kernel(c, kDtCompute, 0, 0.0, dtfinal, 0.0, 0, 0, false, 0);
    if (rc) return rc;
    double ts[4];
    HIPCHK(c, hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *dt = ts[1];
    return GPM_OK;
}

int gpm_set_time(gpm_ctx *c, double t)
{
    if (!c) return GPM_EINVAL;
    double ts[8] = {t, 0, 0, 0, t, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(c->tstate, ts, sizeof(ts), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

int gpm_time_stepping(gpm_ctx *c, double *Uin, double *Uout, double t, double dt, int32_t *cellgpo,
                      gpm_stats *stats)
{
    if (!c || !Uin || !Uout) return fail(c, GPM_EINVAL, "null argument");
    if (c->hds_on)  // this step's rows start at 0
        HIPCHK(c, hipMemsetAsync(c->hds.d_count, 0, sizeof(unsigned long long), c->stream));
    int rc = gpm_set_state(c, Uin);
    if (rc) return rc;
    // Boundary_C(Uin) of the first Forward_Euler stage (mod_FE.f90:64, t_rk = t)
    // refills Uin's ghosts.
    fill_ghosts(Uin, c->Bh, dmr_xmin_top(t, c->Bh.sqrt3));
    rc = gpm_set_time(c, t);
    if (rc) return rc;
    rc = launch_dt_kernel(c, kDtGiven, 0, 0.0, 0.0, dt, 0, 0, false, 1);
    if (rc) return rc;
    rc = enqueue_stages(c);
    if (rc) return rc;
    rc = collect_timing(c);
    if (rc) return rc;
    rc = mood_wide_read(c);
    if (rc) return rc;
    rc = gpm_get_state(c, Uout, cellgpo);
    if (rc) return rc;
    rc = mood_wide_decide(c);
    if (rc) return rc;
    if (stats) {
        rc = gpm_get_stats(c, stats);
        if (rc) return rc;
        stats->t = t;
        stats->dt = dt;
    }
    return GPM_OK;
}

int gpm_run(gpm_ctx *c, int32_t nsteps, double tmax, double *dts)
{
    if (!c || nsteps < 0) return fail(c, GPM_EINVAL, "invalid argument");
    if (!c->has_state) return fail(c, GPM_ESTATE, "no state set");
    if (c->hds_on) return fail(c, GPM_ESTATE, "the reference dataset arrays are filled step by step (gpm_time_stepping)");
    if (dts && nsteps > c->dt_log_cap) {
        if (c->dt_log) hipFree(c->dt_log);
        c->dt_log = nullptr;
        HIPCHK(c, dalloc(&c->dt_log, nsteps));
        c->dt_log_cap = nsteps;
    }
    for (int s = 0; s < nsteps; ++s) {
        int rc = launch_dt_kernel(c, kDtCompute, 1, tmax, 0.0, 0.0, 1, s, dts != nullptr, 1);
        if (rc) return rc;
        rc = enqueue_stages(c);
        if (rc) return rc;
        if (c->timing && c->ev_pairs.size() > 8192) {
            rc = collect_timing(c);
            if (rc) return rc;
        }
    }
    int rc = collect_timing(c);
    if (rc) return rc;
    if (dts && nsteps > 0)
        HIPCHK(c, hipMemcpyAsync(dts, c->dt_log, nsteps * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    rc = mood_wide_read(c);
    if (rc) return rc;
    // synchronous: the steps are done when gpm_run returns
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return mood_wide_decide(c);
}

int gpm_get_stats(gpm_ctx *c, gpm_stats *st)
{
    if (!c || !st) return fail(c, GPM_EINVAL, "null argument");
    DevStats d{};
    double ts[4];
    HIPCHK(c, hipMemcpyAsync(&d, c->stats, sizeof(d), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ts, c->tstate, sizeof(ts), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    st->steps = d.steps;
    st->count_detected_cell = d.detected_step;
    st->detected_total = d.detected_total;
    for (int k = 0; k < 3; ++k) st->passes[k] = d.passes[k];
    st->max_passes = d.max_passes;
    st->second_pass_cells = d.second_pass_cells;
    st->a_priori_lowered = d.a_priori_lowered;
    st->t = ts[0];
    st->dt = ts[1];
    st->count_detected_cell_a_priori = d.apriori_step;
    st->nn_nan_passes = d.nn_passes_step;
    st->nn_nan_passes_total = d.nn_passes_total;
    st->niter = c->niter;
    return GPM_OK;
}

int gpm_reset_stats(gpm_ctx *c)
{
    if (!c) return GPM_EINVAL;
    HIPCHK(c, hipMemsetAsync(c->stats, 0, sizeof(DevStats), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->stage_ms = c->mood_ms = 0;
    c->stage_launches = c->mood_launches = 0;
    c->ev_pairs.clear();
    c->ev_used = 0;
    return GPM_OK;
}

void *gpm_stream(gpm_ctx *c) { return c ? (void *)c->stream : nullptr; }

int gpm_set_mood_launch(gpm_ctx *c, int32_t cooperative)
{
    if (!c) return GPM_EINVAL;
    c->mood_coop = cooperative != 0;
    return GPM_OK;
}

int gpm_get_mood_launch(gpm_ctx *c, int32_t *cooperative)
{
    if (!c || !cooperative) return GPM_EINVAL;
    *cooperative = c->mood_coop ? 1 : 0;
    return GPM_OK;
}

int gpm_set_timing(gpm_ctx *c, int32_t enable)
{
    if (!c || enable < 0) return GPM_EINVAL;
    int rc = collect_timing(c);
    c->timing = enable != 0;
    c->timing_every = enable > 1 ? enable : 1;
    c->timing_step = 0;
    return rc;
}

int gpm_kernel_time(gpm_ctx *c, double *avg_stage_ms, int64_t *stage_launches, double *avg_mood_ms,
                    int64_t *mood_launches)
{
    if (!c) return GPM_EINVAL;
    int rc = collect_timing(c);
    if (rc) return rc;
    if (avg_stage_ms) *avg_stage_ms = c->stage_launches ? c->stage_ms / c->stage_launches : 0.0;
    if (stage_launches) *stage_launches = c->stage_launches;
    if (avg_mood_ms) *avg_mood_ms = c->mood_launches ? c->mood_ms / c->mood_launches : 0.0;
    if (mood_launches) *mood_launches = c->mood_launches;
    return GPM_OK;
}

int gpm_set_predictor(gpm_ctx *c, gpm_mlp *m, int32_t gate)
{
    if (!c) return GPM_EINVAL;
    if (m && !square_copy_rk3(c))
        return fail(c, GPM_EUNSUP, "the predictor needs lf = nf, a copy boundary, SSP_RK3 and no source term");
    if (m) {
        const int R = (c->mord - 1) / 2;
        if (mlp_stencil(m) != 2 * R + 1 || mlp_classes(m) != (c->mord + 1) / 2)
            return fail(c, GPM_EINVAL, "predictor stencil/classes do not match Mord");
    }
    if (m && c->refnn_w) return fail(c, GPM_ESTATE, "the reference network is set; a context runs one predictor");
    c->nn = m;
    c->nn_gate = gate;
    return GPM_OK;
}

int gpm_set_reference_nn(gpm_ctx *c, int32_t L, int32_t length, const float *weight0, const float *bias0,
                         const float *weight1, const float *bias1, const float *weight2, const float *bias2)
{
    if (!c) return GPM_EINVAL;
    if (c->cfg.method != GPM_NN_GP_MOOD && c->cfg.method != GPM_NN_GP_MOOD_CC)
        return fail(c, GPM_ESTATE, "the reference network needs method NN_GP_MOOD or NN_GP_MOOD_CC");
    if (c->nn) return fail(c, GPM_ESTATE, "a learned predictor is set; a context runs one predictor");
    if (c->mord != 3 && c->mord != 5)
        return fail(c, GPM_EUNSUP, "the reference network's stencil (szs_sphere(Mord+2)) is built for Mord 3 and 5");
    const int S = c->mord == 3 ? 13 : 25;
    if (!weight0 || !bias0 || !weight1 || !bias1 || !weight2 || !bias2) return fail(c, GPM_EINVAL, "null argument");
    if (L < 4 * S || L > 128)
        return fail(c, GPM_EINVAL, "L must hold the 4*szs_sphere(Mord+2) stencil values and be at most 128");
    if (length < 1 || length > kRefNnMaxLen) return fail(c, GPM_EUNSUP, "length must be 1..64");
    const bool cc = c->cfg.method == GPM_NN_GP_MOOD_CC;
    // values_NN(L, n, :) must lie inside values_NN(-1:lf+2, -1:nf+2, 2) for every row
    if (cc && (long long)L + 1 >= 3ll * (c->NX + 4))
        return fail(c, GPM_EUNSUP, "NN_GP_MOOD_CC with L >= 3*(lf+4)-1: the reference writes past values_NN");
    // the kernel's layout: rows of weight0 and weight1 zero-padded to the
    // kernel's widths, then the biases and weight2 as given
    const int KP = refnn_input_pad(L), HP = kRefNnMaxLen;
    const size_t total = (size_t)length * KP + length + (size_t)length * HP + length + 2 * (size_t)length + 2;
    std::vector<float> h(total, 0.0f);
    float *p0 = h.data(), *pb0 = p0 + (size_t)length * KP, *p1 = pb0 + length, *pb1 = p1 + (size_t)length * HP;
    float *p2 = pb1 + length, *pb2 = p2 + 2 * (size_t)length;
    for (int i = 0; i < length; ++i) {
        for (int k = 0; k < L; ++k) p0[(size_t)i * KP + k] = weight0[i + (size_t)k * length];
        for (int k = 0; k < length; ++k) p1[(size_t)i * HP + k] = weight1[i + (size_t)k * length];
        pb0[i] = bias0[i];
        pb1[i] = bias1[i];
    }
    std::memcpy(p2, weight2, 2 * (size_t)length * sizeof(float));
    std::memcpy(pb2, bias2, 2 * sizeof(float));
    if (c->refnn_w) { HIPCHK(c, hipStreamSynchronize(c->stream)); hipFree(c->refnn_w); c->refnn_w = nullptr; }
    HIPCHK(c, dalloc(&c->refnn_w, total));
    HIPCHK(c, hipMemcpyAsync(c->refnn_w, h.data(), total * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (cc && !c->vnn) {
        HIPCHK(c, dalloc(&c->rowv, 2 * (size_t)c->NY));
        HIPCHK(c, dalloc(&c->vnn, 2 * (size_t)(c->NX + 4) * (c->NY + 4)));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->refnn_L = L;
    c->refnn_len = length;
    c->refnn_cc = cc;
    return GPM_OK;
}

int gpm_read_reference_nn(const char *nn_filename, int32_t L, int32_t length, float *weight0, float *bias0,
                          float *weight1, float *bias1, float *weight2, float *bias2)
{
    if (!nn_filename || !weight0 || !bias0 || !weight1 || !bias1 || !weight2 || !bias2) return GPM_EINVAL;
    if (L < 1 || length < 1 || L > 4096 || length > 4096) return GPM_EINVAL;
    // load_NN (mod_NN.f90:60-117): "<NN_filename>.txt", one number per record,
    // weight0 row by row (j fastest: update_ij), bias0, weight1 row by row,
    // bias1, weight2 row by row, bias2; records past up5 are not read
    std::string path = std::string(nn_filename);
    while (!path.empty() && path.back() == ' ') path.pop_back();
    size_t b = 0;
    while (b < path.size() && path[b] == ' ') ++b;
    path = path.substr(b) + ".txt";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return GPM_EINVAL;
    // (Fortran array, rows, columns) in reading order; element (i, j) at i + j*rows
    struct Part { float *a; int rows, cols; } parts[6] = {{weight0, length, L}, {bias0, length, 1},
                                                         {weight1, length, length}, {bias1, length, 1},
                                                         {weight2, 2, length}, {bias2, 2, 1}};
    char line[512];
    for (const Part &p : parts)
        for (int i = 0; i < p.rows; ++i)
            for (int j = 0; j < p.cols; ++j) {
                char *end = nullptr;
                float v = 0.0f;
                bool got = false;
                while (!got && std::fgets(line, sizeof line, f)) {
                    char *q = line;
                    while (*q == ' ' || *q == '\t') ++q;
                    if (*q == '\n' || *q == '\r' || *q == 0) continue;  // list-directed reads skip blank records
                    v = std::strtof(q, &end);
                    if (end == q) { std::fclose(f); return GPM_EINVAL; }
                    got = true;
                }
                if (!got) { std::fclose(f); return GPM_EINVAL; }  // the file ends before all weights are read
                p.a[i + (size_t)j * p.rows] = v;
            }
    std::fclose(f);
    return GPM_OK;
}

int gpm_load_reference_nn(gpm_ctx *c, const char *nn_filename, int32_t L, int32_t length)
{
    if (!c || !nn_filename) return GPM_EINVAL;
    if (L < 1 || length < 1 || L > 4096 || length > 4096) return fail(c, GPM_EINVAL, "invalid L or length");
    std::vector<float> w0((size_t)length * L), b0(length), w1((size_t)length * length), b1(length),
        w2(2 * (size_t)length), b2(2);
    if (gpm_read_reference_nn(nn_filename, L, length, w0.data(), b0.data(), w1.data(), b1.data(), w2.data(),
                              b2.data()))
        return fail(c, GPM_EINVAL, std::string("cannot read ") + nn_filename + ".txt: missing file, a record "
                                   "that is not a number, or fewer records than the network holds");
    return gpm_set_reference_nn(c, L, length, w0.data(), b0.data(), w1.data(), b1.data(), w2.data(), b2.data());
}

int gpm_set_reference_dump(gpm_ctx *c, float *inputs_dev, float *labels_dev, int64_t *cells_dev, int64_t capacity,
                           uint64_t *count_dev, int64_t *nr_dev, uint64_t seed, int32_t L)
{
    if (!c) return GPM_EINVAL;
    if (!inputs_dev) {
        c->rdump_on = false;
        return GPM_OK;
    }
    if (!labels_dev || !count_dev || !nr_dev || capacity < 0) return fail(c, GPM_EINVAL, "null argument");
    if (c->mord != 3 && c->mord != 5)
        return fail(c, GPM_EUNSUP, "the reference network's stencil (szs_sphere(Mord+2)) is built for Mord 3 and 5");
    if (L < 4 * (c->mord == 3 ? 13 : 25)) return fail(c, GPM_EINVAL, "L must hold the 4*szs_sphere(Mord+2) stencil values");
    if (!c->rdump.nr_start) HIPCHK(c, dalloc(&c->rdump.nr_start, 2));
    c->rdump.inputs = inputs_dev; c->rdump.labels = labels_dev; c->rdump.cells = reinterpret_cast<long long *>(cells_dev);
    c->rdump.capacity = capacity; c->rdump.count = reinterpret_cast<unsigned long long *>(count_dev);
    c->rdump.nr = reinterpret_cast<long long *>(nr_dev); c->rdump.seed = seed; c->rdump.L = L;
    c->rdump_on = true;
    return GPM_OK;
}

int gpm_set_reference_dataset(gpm_ctx *c, int32_t L, int32_t nr0, int32_t nr1)
{
    if (!c) return GPM_EINVAL;
    auto &h = c->hds;
    if (L <= 0) {
        c->hds_on = false;
        c->rdump_on = false;
        return GPM_OK;
    }
    if (nr0 < 0 || nr1 < 0) return fail(c, GPM_EINVAL, "NR0 and NR1 must not be negative");
    const size_t n = (size_t)c->NX * c->NY;
    if (!h.d_in || h.L != L) {
        if (h.d_in) { HIPCHK(c, hipStreamSynchronize(c->stream)); hipFree(h.d_in); h.d_in = nullptr; }
        HIPCHK(c, dalloc(&h.d_in, n * (size_t)L));
    }
    if (!h.d_lab) HIPCHK(c, dalloc(&h.d_lab, 2 * n));
    if (!h.d_cells) HIPCHK(c, dalloc(&h.d_cells, n));
    if (!h.d_nr) HIPCHK(c, dalloc(&h.d_nr, 2));
    if (!h.d_count) HIPCHK(c, dalloc(&h.d_count, 1));
    h.L = L;
    int rc = gpm_set_reference_dump(c, h.d_in, h.d_lab, reinterpret_cast<int64_t *>(h.d_cells), (int64_t)n,
                                    reinterpret_cast<uint64_t *>(h.d_count), reinterpret_cast<int64_t *>(h.d_nr),
                                    0x5EED0D5Eull, L);
    if (rc) return rc;
    const long long nr[2] = {nr0, nr1};
    HIPCHK(c, hipMemsetAsync(h.d_count, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipMemcpyAsync(h.d_nr, nr, sizeof(nr), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->hds_on = true;
    return GPM_OK;
}

int gpm_append_reference_dataset(gpm_ctx *c, float *inputs, float *labels, int32_t dataset_size, int32_t *index,
                                 int32_t *n_overwrite, int32_t *nr0, int32_t *nr1)
{
    if (!c) return GPM_EINVAL;
    if (!c->hds_on) return fail(c, GPM_ESTATE, "gpm_set_reference_dataset has not switched the samples on");
    if (!inputs || !labels || !index || !n_overwrite || !nr0 || !nr1 || dataset_size < 1)
        return fail(c, GPM_EINVAL, "null argument or dataset_size < 1");
    if (*index < 1 || *index > dataset_size) return fail(c, GPM_EINVAL, "index must be in 1..dataset_size");
    auto &h = c->hds;
    unsigned long long count = 0;
    HIPCHK(c, hipMemcpyAsync(&count, h.d_count, sizeof(count), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->NX * c->NY;
    if (count > n) count = n;
    std::vector<float> in(count * (size_t)h.L), lab(count * 2);
    std::vector<long long> cells(count);
    if (count) {
        HIPCHK(c, hipMemcpyAsync(in.data(), h.d_in, in.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(lab.data(), h.d_lab, lab.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(cells.data(), h.d_cells, cells.size() * sizeof(long long), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    // append_to_NN_datatset's loop order (n outer, l inner): rows by cell id
    std::vector<long long> row_of(n, -1);
    for (size_t r = 0; r < count; ++r) row_of[(size_t)cells[r]] = (long long)r;
    for (size_t cell = 0; cell < n; ++cell) {
        const long long r = row_of[cell];
        if (r < 0) continue;
        const int64_t i = *index - 1;  // inputs(index, 1:L), labels(index, 1:2), column-major
        for (int k = 0; k < h.L; ++k) inputs[i + (int64_t)k * dataset_size] = in[(size_t)r * h.L + k];
        labels[i] = lab[2 * (size_t)r];
        labels[i + dataset_size] = lab[2 * (size_t)r + 1];
        if (lab[2 * (size_t)r + 1] == 1.0f) ++*nr1;
        else ++*nr0;
        if (*index == dataset_size) {
            *index = 1;
            ++*n_overwrite;
        } else {
            ++*index;
        }
    }
    // the next step's order-3 keep probability starts from these counts
    const long long nr[2] = {*nr0, *nr1};
    HIPCHK(c, hipMemsetAsync(h.d_count, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipMemcpyAsync(h.d_nr, nr, sizeof(nr), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

int gpm_set_niter(gpm_ctx *c, int64_t niter)
{
    if (!c || niter < 0) return fail(c, GPM_EINVAL, "invalid argument");
    c->niter = niter;
    return GPM_OK;
}

int gpm_get_cellgpo_priori(gpm_ctx *c, int32_t *priori)
{
    if (!c || !priori) return fail(c, GPM_EINVAL, "null argument");
    const size_t n = (size_t)c->NX * c->NY;
    std::vector<int8_t> h(n, (int8_t)c->mord);
    if (c->priori_set) {
        HIPCHK(c, hipMemcpyAsync(h.data(), c->priori, n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    const int W = c->NX + 8, H = c->NY + 8;
    for (int jj = 0; jj < H; ++jj)
        for (int ii = 0; ii < W; ++ii) {
            const int i = ii - 4, j = jj - 4;
            priori[(size_t)jj * W + ii] =
                (i < 0 || i >= c->NX || j < 0 || j >= c->NY) ? c->mord : (int32_t)h[(size_t)j * c->NX + i];
        }
    return GPM_OK;
}

int gpm_nn_expf_bits_host(uint32_t first, int64_t n, float *y)
{
    if (!y || n < 0) return GPM_EINVAL;
    nn_expf_bits_host(first, n, y);
    return GPM_OK;
}

int gpm_nn_expf_bits(gpm_ctx *c, uint32_t first, int64_t n, float *y_dev)
{
    if (!c || !y_dev || n < 0) return fail(c, GPM_EINVAL, "invalid argument");
    HIPCHK(c, launch_nn_expf_bits(first, n, y_dev, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

int gpm_set_dump(gpm_ctx *c, float *x_dev, int32_t *y_dev, int64_t capacity, uint64_t *count_dev, double keep_mord,
                 uint64_t seed, int32_t gated_only, int32_t stage_mask)
{
    if (!c) return GPM_EINVAL;
    if (x_dev && (!y_dev || !count_dev || capacity < 0)) return fail(c, GPM_EINVAL, "dump buffers incomplete");
    if (x_dev && !square_copy_rk3(c))
        return fail(c, GPM_EUNSUP, "the dump needs lf = nf, a copy boundary, SSP_RK3 and no source term");
    const int R = (c->mord - 1) / 2;
    if (x_dev && R != 2 && R != 3) return fail(c, GPM_EUNSUP, "dump needs Mord 5 or 7");
    c->dump_x = x_dev;
    c->dump_y = y_dev;
    c->dump_cap = capacity;
    c->dump_count = reinterpret_cast<unsigned long long *>(count_dev);
    c->dump_keep = keep_mord;
    c->dump_seed = seed;
    c->dump_gated = gated_only;
    c->dump_stages = stage_mask;
    return GPM_OK;
}

int gpm_features(gpm_ctx *c, const int32_t *cells_dev, int32_t n, float *x_dev)
{
    if (!c || !cells_dev || !x_dev || n < 0) return GPM_EINVAL;
    if (!c->has_state) return fail(c, GPM_ESTATE, "no state set");
    if (!square_copy_rk3(c)) return fail(c, GPM_EUNSUP, "features need lf = nf and a copy boundary");
    const int R = (c->mord - 1) / 2;
    HIPCHK(c, mlp_launch_features(c->dU, c->N, c->bc, 2 * R + 1, cells_dev, n, x_dev, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

int gpm_get_apriori_orders(gpm_ctx *c, int8_t *pred_dev)
{
    if (!c || !pred_dev) return GPM_EINVAL;
    HIPCHK(c, hipMemcpyAsync(pred_dev, c->pred, (size_t)c->NX * c->NY, hipMemcpyDeviceToDevice, c->stream));
    return GPM_OK;
}

int gpm_set_apriori_orders(gpm_ctx *c, const int8_t *orders)
{
    if (!c) return GPM_EINVAL;
    if (!orders) {
        c->fixed_orders = false;
        return GPM_OK;
    }
    const size_t n = (size_t)c->NX * c->NY;
    for (size_t i = 0; i < n; ++i)
        if (orders[i] < 1 || orders[i] > c->mord || !(orders[i] & 1))
            return fail(c, GPM_EINVAL, "a priori orders must be odd and in [1, Mord]");
    HIPCHK(c, hipMemcpyAsync(c->pred, orders, n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->fixed_orders = true;
    return GPM_OK;
}

// ---- differentiable form: frozen-mask step and its adjoint ------------------

int gpm_record_masks(gpm_ctx *c, int8_t *masks_dev, int32_t max_steps)
{
    if (!c || max_steps < 0) return GPM_EINVAL;
    if (!masks_dev || max_steps == 0) {  // stop recording; the count stays readable
        c->rec_masks = nullptr;
        c->rec_cap = 0;
        return GPM_OK;
    }
    if (!square_copy_rk3(c))
        return fail(c, GPM_EUNSUP, "mask recording needs lf = nf, a copy boundary, SSP_RK3 and no source term");
    c->rec_masks = masks_dev;
    c->rec_cap = max_steps;
    c->rec_count = 0;
    return GPM_OK;
}

int gpm_recorded_steps(gpm_ctx *c, int32_t *nsteps)
{
    if (!c || !nsteps) return GPM_EINVAL;
    *nsteps = c->rec_count;
    return GPM_OK;
}

int gpm_frozen_run(gpm_ctx *c, const int8_t *masks_dev, const double *dts, int32_t nsteps, double *stage_states_dev)
{
    if (!c || !masks_dev || !dts || nsteps < 0) return fail(c, GPM_EINVAL, "invalid argument");
    if (!c->has_state) return fail(c, GPM_ESTATE, "no state set");
    if (!square_copy_rk3(c))
        return fail(c, GPM_EUNSUP, "the frozen run needs lf = nf, a copy boundary, SSP_RK3 and no source term");
    const size_t n = (size_t)c->NX * c->NY;
    for (int st = 0; st < nsteps; ++st) {
        int rc = launch_dt_kernel(c, kDtGiven, 0, 0.0, 0.0, dts[st], 1, 0, false, 1);
        if (rc) return rc;
        StagePlan plan[kMaxStages];
        const int nst = plan_stages(c, plan);
        for (int k = 0; k < nst; ++k) {
            if (stage_states_dev)
                HIPCHK(c, hipMemcpyAsync(stage_states_dev + ((size_t)st * 3 + k) * n * 4, plan[k].in,
                                         n * sizeof(double4), hipMemcpyDeviceToDevice, c->stream));
            StageArgs s = stage_args(c, k, plan[k]);
            s.apriori = 1;
            s.frozen = 1;
            s.pred = masks_dev + ((size_t)st * 3 + k) * n;
            HIPCHK(c, launch_stage(c->mord, c->ngp, c->cfg.num_flux, false, s, (c->ntiles + 7) / 8 * 8, c->stream));
        }
        std::swap(c->dU, c->dA);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

int gpm_adjoint(gpm_ctx *c, const double *stage_states_dev, const int8_t *masks_dev, const double *dts, int32_t nsteps,
                double *ubar_dev)
{
    if (!c || !stage_states_dev || !masks_dev || !dts || !ubar_dev || nsteps < 0)
        return fail(c, GPM_EINVAL, "invalid argument");
    if (!adj_supported(c->mord, c->ngp)) return fail(c, GPM_EUNSUP, "adjoint needs (Mord, ngp) in {(3,2), (5,3), (7,4)}");
    if (!square_copy_rk3(c))
        return fail(c, GPM_EUNSUP, "the adjoint needs lf = nf, a copy boundary, SSP_RK3 and no source term");
    if (c->cfg.num_flux != GPM_HLLC) return fail(c, GPM_EUNSUP, "adjoint needs numFlux = HLLC");
    const int N = c->N, R = (c->mord - 1) / 2;
    const size_t n = (size_t)N * N, W = (size_t)(N + 2), E = (size_t)(N + 2 * R + 2);
    if (!c->adj_ub) {
        const bool ok = dalloc(&c->adj_ub, 4 * (size_t)c->ngp * W * W) == hipSuccess &&
                        dalloc(&c->adj_G, E * E) == hipSuccess && dalloc(&c->adj_A, n) == hipSuccess &&
                        dalloc(&c->adj_tmp, n) == hipSuccess && dalloc(&c->adj_ordx, (size_t)N * (N + 1)) == hipSuccess &&
                        dalloc(&c->adj_ordy, (size_t)N * (N + 1)) == hipSuccess;
        if (!ok) return fail(c, GPM_ENOMEM, "adjoint work buffers");
    }
    const double4 *S = reinterpret_cast<const double4 *>(stage_states_dev);
    double4 *cur = reinterpret_cast<double4 *>(ubar_dev), *nxt = c->adj_tmp;
    AdjFaceArgs fa{};
    fa.T = c->dT; fa.ub = c->adj_ub; fa.ordx = c->adj_ordx; fa.ordy = c->adj_ordy; fa.N = N; fa.bc = c->bc;
    AdjGatherArgs ga{};
    ga.ub = c->adj_ub; ga.ordx = c->adj_ordx; ga.ordy = c->adj_ordy; ga.T = c->dT; ga.G = c->adj_G; ga.N = N;
    AdjFoldArgs fo{};
    fo.G = c->adj_G; fo.N = N; fo.bc = c->bc;
    // SSP-RK3 (mod_SSP_RK3.f90:12-39): U1 = E1(U0); U2 = 3/4 U0 + 1/4 E2(U1); U3 = 1/3 U0 + 2/3 E3(U2),
    // E_s(V) = V + dt L_s(V).  Reverse, from the adjoint of U3 in `cur`:
    //   U2b = 2/3 E3'^T U3b,   A  = 1/3 U3b
    //   U1b = 1/4 E2'^T U2b,   A += 3/4 U2b
    //   U0b = E1'^T U1b + A
    const double third = 1.0 / 3, two_thirds = 2.0 / 3;
    for (int st = nsteps - 1; st >= 0; --st) {
        fa.dtdx = dts[st] / c->dx;
        fa.dtdy = dts[st] / c->dy;
        const int8_t *m = masks_dev + (size_t)st * 3 * n;
        // stage 3: input U2
        fa.V = S + ((size_t)st * 3 + 2) * n; fa.mask = m + 2 * n; fa.Wb = cur; fa.alpha = two_thirds;
        fo.Wn = cur; fo.alpha = two_thirds; fo.B = nullptr; fo.gamma = 0.0; fo.out = nxt;
        fo.A = c->adj_A; fo.aA = 0.0; fo.aU = third;
        HIPCHK(c, launch_adj_stage(c->mord, c->ngp, fa, ga, fo, c->stream));
        std::swap(cur, nxt);
        // stage 2: input U1
        fa.V = S + ((size_t)st * 3 + 1) * n; fa.mask = m + n; fa.Wb = cur; fa.alpha = 0.25;
        fo.Wn = cur; fo.alpha = 0.25; fo.out = nxt; fo.aA = 1.0; fo.aU = 0.75;
        HIPCHK(c, launch_adj_stage(c->mord, c->ngp, fa, ga, fo, c->stream));
        std::swap(cur, nxt);
        // stage 1: input U0
        fa.V = S + (size_t)st * 3 * n; fa.mask = m; fa.Wb = cur; fa.alpha = 1.0;
        fo.Wn = cur; fo.alpha = 1.0; fo.out = nxt; fo.A = nullptr; fo.B = c->adj_A; fo.gamma = 1.0;
        HIPCHK(c, launch_adj_stage(c->mord, c->ngp, fa, ga, fo, c->stream));
        std::swap(cur, nxt);
    }
    if (cur != reinterpret_cast<double4 *>(ubar_dev))
        HIPCHK(c, hipMemcpyAsync(ubar_dev, cur, n * sizeof(double4), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return GPM_OK;
}

int gpm_run_noalloc(const gpm_config *cfg, const double *zT_sp, const double *gp_d2x_sp, const double *gp_d2y_sp,
                    const double *gauss_weight, const int32_t *ixiy_sp, double *u_dev, int32_t nsteps, double t,
                    double tmax, double *t_end, void *stream)
{
    auto bad = [](int code, const char *m) {
        std::fprintf(stderr, "gpm_run_noalloc: %s\n", m);
        return code;
    };
    if (!cfg || !zT_sp || !gp_d2x_sp || !gp_d2y_sp || !gauss_weight || !ixiy_sp || !u_dev || nsteps < 0)
        return bad(GPM_EINVAL, "null argument or negative nsteps");
    if (cfg->lf < 8 || cfg->nf < 8) return bad(GPM_EINVAL, "grids need lf, nf >= 8");
    if ((long long)cfg->lf * cfg->nf > (long long)kNaMaxCells)
        return bad(GPM_EINVAL, "lf * nf exceeds the 1048576 cells of the scratch buffer");
    if (cfg->method != GPM_GP_MOOD) return bad(GPM_EUNSUP, "method must be GP_MOOD");
    if (!index_map_bc(cfg->bc_type) || cfg->problem == GPM_PROBLEM_RT)
        return bad(GPM_EUNSUP, "needs a copy boundary (Neumann, Periodic, Reflective, KH_BC) and no source term");
    if (cfg->integrator != GPM_FE && cfg->integrator != GPM_SSP_RK2 && cfg->integrator != GPM_SSP_RK3)
        return bad(GPM_EUNSUP, "integrator must be FE, SSP_RK2 or SSP_RK3");
    if (cfg->num_flux != GPM_HLLC && cfg->num_flux != GPM_LLF && cfg->num_flux != GPM_HLL &&
        cfg->num_flux != GPM_UNSPLIT)
        return bad(GPM_EUNSUP, "numFlux must be HLLC, LLF, HLL or unsplit");
    if (cfg->cascade != GPM_CASCADE_REFERENCE && cfg->cascade != GPM_CASCADE_FULL)
        return bad(GPM_EINVAL, "unknown cascade");
    if (!supported(cfg->mord, cfg->ngp)) return bad(GPM_EUNSUP, "(Mord, ngp) combination not compiled");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) return bad(GPM_EHIP, "no usable device");
    unsigned char *base = nullptr;
    if (int rc = na_scratch(dev, &base)) return rc;

    // A context on the stack whose buffers are carved from the scratch memory.
    gpm_ctx c{};
    c.cfg = *cfg;
    c.NX = cfg->lf; c.NY = cfg->nf; c.N = cfg->lf;
    c.mord = cfg->mord; c.ngp = cfg->ngp; c.bc = cfg->bc_type; c.integrator = cfg->integrator;
    c.device = dev;
    c.nstages = cfg->integrator == GPM_FE ? 1 : cfg->integrator == GPM_SSP_RK2 ? 2 : 3;
    c.dx = quad_ratio(cfg->lx, cfg->lf);
    c.dy = quad_ratio(cfg->ly, cfg->nf);
    c.Bh = make_bnd(c.NX, c.NY, c.bc, c.dx, c.dy, c.st_host);
    c.B = c.Bh;
    c.B.st = nullptr;  // copy boundaries read no fixed ghost state
    c.stream = static_cast<hipStream_t>(stream);
    c.dT = reinterpret_cast<DevTables *>(base + kNaTables);
    c.counters = reinterpret_cast<int *>(base + kNaCounters);
    c.stats = reinterpret_cast<DevStats *>(base + kNaStats);
    c.tstate = reinterpret_cast<double *>(base + kNaTstate);
    c.block_min = reinterpret_cast<double *>(base + kNaBlockMin);
    c.done = reinterpret_cast<unsigned *>(base + kNaDone);
    c.stamp = reinterpret_cast<int *>(base + kNaStamp);
    c.dA = reinterpret_cast<double4 *>(base + kNaA);
    c.dB = reinterpret_cast<double4 *>(base + kNaB);
    c.gpo = reinterpret_cast<order_t *>(base + kNaGpo);
    c.slistA = reinterpret_cast<int *>(base + kNaListA);
    c.slistB = reinterpret_cast<int *>(base + kNaListB);
    c.mark = reinterpret_cast<int *>(base + kNaMark);
    c.dflag = base + kNaFlag;
    c.dU = reinterpret_cast<double4 *>(u_dev);
    int ncu = 0;
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (ncu <= 0) ncu = 256;
    c.dt_blocks = ncu < kNaMaxDtBlocks ? ncu : kNaMaxDtBlocks;
    c.tiles_x = (c.NX + stage_tx(c.mord) - 1) / stage_tx(c.mord);
    c.ntiles = c.tiles_x * ((c.NY + stage_ty(c.mord) - 1) / stage_ty(c.mord));
    {
        int per = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, mood_kernel(c.mord, c.ngp, cfg->num_flux, false, true), kMoodThreads, 0);
        if (e != hipSuccess || per <= 0) per = 1;
        const int want = ncu / 4 > 0 ? ncu / 4 : 1;
        c.mood_blocks = want < per * ncu ? want : per * ncu;
    }
    c.mood_solo = kSoloMax;
    // no earlier call to choose the wide launches' plan from: the
    // persistent kernel alone, as for short lists
    c.mood_wide = 0;
    if (const char *env = std::getenv("GPM_MOOD_COOP")) c.mood_coop = std::atoi(env) != 0;
    mood_ord_env(&c);
    c.has_state = true;

    int rc = upload_tables(&c, zT_sp, gp_d2x_sp, gp_d2y_sp, gauss_weight, ixiy_sp);
    if (rc) return bad(rc, c.msg.c_str());
    const size_t n = (size_t)c.NX * c.NY;
    if (hipError_t e = launch_run_reset(c.mark, c.gpo, n, pack_order(c.mord, c.mord, 0), c.counters,
                                        kMaxStages * kCountersPerStage, c.stats, c.tstate, t, c.done, c.stamp, c.stream))
        return bad(GPM_EHIP, hipGetErrorString(e));
    for (int st = 0; st < nsteps; ++st) {
        rc = launch_dt_kernel(&c, kDtCompute, 1, tmax, 0.0, 0.0, 1, st, false, 1);
        if (!rc) rc = enqueue_stages(&c);
        if (rc) return bad(rc, c.msg.c_str());
    }
    // the last step's output lies in u_dev or in a scratch state buffer
    if (c.dU != reinterpret_cast<double4 *>(u_dev) &&
        hipMemcpyAsync(u_dev, c.dU, n * sizeof(double4), hipMemcpyDeviceToDevice, c.stream) != hipSuccess)
        return bad(GPM_EHIP, "copy of the final state");
    if (t_end && hipMemcpyAsync(t_end, c.tstate, sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        return bad(GPM_EHIP, "copy of the final time");
    // the MOOD launch reports (an ordinary persistent launch's stalls)
    rc = mood_wide_read(&c);
    if (rc) return bad(rc, c.msg.c_str());
    if (hipStreamSynchronize(c.stream) != hipSuccess) return bad(GPM_EHIP, "stream synchronisation");
    if (c.mood_chk[2] > 0) return bad(GPM_EHIP, "a grid barrier of the MOOD passes did not complete");
    return GPM_OK;
}

int gpm_ddiv(int64_t n, const double *a_dev, const double *b_dev, double *q_dev)
{
    if (n < 0 || (n > 0 && (!a_dev || !b_dev || !q_dev))) return GPM_EINVAL;
    if (launch_ddiv(a_dev, b_dev, q_dev, n, false, nullptr) != hipSuccess) return GPM_EHIP;
    return hipDeviceSynchronize() == hipSuccess ? GPM_OK : GPM_EHIP;
}

int gpm_qdiv(int64_t n, const double *a_dev, const double *b_dev, double *q_dev)
{
    if (n < 0 || (n > 0 && (!a_dev || !b_dev || !q_dev))) return GPM_EINVAL;
    if (launch_ddiv(a_dev, b_dev, q_dev, n, true, nullptr) != hipSuccess) return GPM_EHIP;
    return hipDeviceSynchronize() == hipSuccess ? GPM_OK : GPM_EHIP;
}

int gpm_hllc_vjp(int32_t dir, int32_t n, const double *ul, const double *ur, const double *fbar, double *ulbar,
                 double *urbar)
{
    if (n < 0 || (dir != 0 && dir != 1) || (n > 0 && (!ul || !ur || !fbar || !ulbar || !urbar))) return GPM_EINVAL;
    if (launch_hllc_vjp(dir, n, ul, ur, fbar, ulbar, urbar, nullptr) != hipSuccess) return GPM_EHIP;
    return hipDeviceSynchronize() == hipSuccess ? GPM_OK : GPM_EHIP;
}

}  // extern "C"
