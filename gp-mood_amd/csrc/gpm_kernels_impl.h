// HIP kernels of the GP-MOOD step for gfx950 (MI355X).
//
// Per RK stage (Forward_Euler, mod_FE.f90:20-190):
//   k_stage  first MOOD pass over the whole grid.  A workgroup stages a tile of
//            the stage input with its halo in LDS, computes every face flux of
//            the tile once (GP reconstruction at the Gauss points + the
//            Riemann solver + quadrature), applies the four face updates per
//            cell in the reference's order, runs the detection, writes the
//            output combined by the integrator and appends troubled cells to
//            a work list (wavefront ballot + one atomic per wavefront).  With
//            Mord 1 it is the first-order Godunov method (FOG), one pass.
//   k_mood   the remaining MOOD passes, as one persistent kernel:
//            each pass lowers the order of the listed cells, marks their
//            5-point neighbourhoods (deduplicated with a pass stamp), and
//            recomputes exactly those cells from scratch with the current
//            orders; grid barriers separate the phases; the loop ends on the
//            device when a pass finds no troubled cell.
//   k_dt     Setdt (physics.f90:78-129) as a grid min-reduction, plus per-step
//            bookkeeping (t, dt/dx, dt/dy, counter reset).
//
// Recomputing a cell from scratch reproduces the reference's stored-flux
// reuse bit for bit: a stored face flux is reused only while neither
// neighbour's order changed (a failing cell re-flags all four of its faces,
// mod_subroutine_MOOD.f90:171-174), so recomputation with the current orders
// yields the same value.
#pragma once
#include <hip/hip_cooperative_groups.h>
#include <type_traits>
#include <hip/hip_runtime.h>

#include "gpm_device.h"
#include "gpm_kernels.h"

namespace cg = cooperative_groups;

namespace gpm {

// Accessor of a stage input staged in LDS (tile coordinates incl. halo).
struct LdsAcc {
    const double4 *s;
    int pitch;
    __device__ __forceinline__ double4 operator()(int i, int j) const { return s[j * pitch + i]; }
};
// the tile anchored at (i, j): the same accessor on a shifted pointer
__device__ __forceinline__ LdsAcc anchor(const LdsAcc &a, int i, int j) { return LdsAcc{a.s + (j * a.pitch + i), a.pitch}; }

// Time t_rk of a stage (time_stepper.f90:17, mod_SSP_RK3.f90:22-31) from the
// step-start time tstate[4] and dt tstate[1].
__device__ __forceinline__ double stage_time(const double *__restrict__ ts, int trk)
{
    const double t0 = ts[4];
    if (trk == 1) return t0 + ts[1];
    if (trk == 2) return t0 + 0.5 * ts[1];
    return t0;
}

__device__ __forceinline__ double stage_xmin_top(const Bnd &B, const double *__restrict__ ts, int trk)
{
    return B.bc == kDMR ? dmr_xmin_top(stage_time(ts, trk), B.sqrt3) : 0.0;
}

// Cells whose values the ghost rules read: within 4 of the boundary.
__device__ __forceinline__ bool in_band(int i, int j, int NX, int NY)
{
    return i < 4 || j < 4 || i >= NX - 4 || j >= NY - 4;
}

// Gravity source of problem RT after the MOOD loop (mod_FE.f90:183-186), with
// dtg = dt*g: E first (with the old momentum), then the y momentum.
__device__ __forceinline__ void gravity_source(double c[4], double dtg)
{
    c[3] = c[3] - dtg * c[2];
    c[2] = c[2] - dtg * c[0];
}

// SSP_RK4 coefficients (constants.f90:49-55).
constexpr double kRK4c1 = 0.391752226571890, kRK4a20 = 0.444370493651235, kRK4a21 = 0.555629506348765,
                 kRK4c2 = 0.368410593050371, kRK4a30 = 0.620101851488403, kRK4a32 = 0.379898148511597,
                 kRK4c3 = 0.251891774271694, kRK4a40 = 0.178079954393132, kRK4a43 = 0.821920045606868,
                 kRK4c4 = 0.544974750228521, kRK4f2 = 0.517231671970585, kRK4f3 = 0.096059710526147,
                 kRK4f4 = 0.386708617503269, kRK4ff3 = 0.063692468666290, kRK4ff4 = 0.226007483236906;

// SSP_RK4 stage outputs (kCmbRK4*) of one component from the step-start value
// un, the stage input uin and the Forward_Euler value c; x1..x3 are the
// kCmbRK4e operands; *w receives FU3 for kCmbRK4d.  Kept out of the common
// path: SSP_RK4 is the only integrator that needs it.
__device__ __forceinline__ double rk4_out1(int kind, double un, double uin, double c, double x1, double x2, double x3,
                                           double *w)
{
    switch (kind) {
    case kCmbRK4a: return un + kRK4c1 * (c - un);
    case kCmbRK4b: return kRK4a20 * un + kRK4a21 * uin + kRK4c2 * (c - uin);
    case kCmbRK4c: return kRK4a30 * un + kRK4a32 * uin + kRK4c3 * (c - uin);
    case kCmbRK4d: {
        const double f = c - uin;
        *w = f;
        return kRK4a40 * un + kRK4a43 * uin + kRK4c4 * f;
    }
    default: return kRK4f2 * x1 + kRK4f3 * x2 + kRK4ff3 * x3 + kRK4f4 * uin + kRK4ff4 * (c - uin);
    }
}

// The combination kind of a stage: CMB when fixed at compile time (the
// SSP-RK3 stages of the copy-boundary kernels), else the run-time C.kind.
constexpr int kCmbRuntime = -1;
template <int CMB>
__device__ __forceinline__ int comb_kind(const StepCtl &C)
{
    return CMB == kCmbRuntime ? C.kind : CMB;
}

// Writes the output of interior cell gc: the band copy of the Forward_Euler
// value (before the source term), the gravity source, the RK combination and,
// for kCmbRK4d, FU3.  un: the step-start state of the cell (loaded by the
// caller when the kind needs it), uin: the stage input.  The band and the
// source term exist only with GEN (formula boundaries or problem RT).
template <bool GEN, int CMB>
__device__ __forceinline__ void write_stage_output(const StepCtl &C, double4 *__restrict__ Uout, size_t gc,
                                                   bool band, const double4 &un, const double4 &uin,
                                                   double cand[4], double dtg)
{
    if (GEN && band) C.ext->band[gc] = make_double4(cand[0], cand[1], cand[2], cand[3]);
    if (GEN && C.gravity) gravity_source(cand, dtg);
    const double u[4] = {un.x, un.y, un.z, un.w};
    double o[4];
    const int kind = comb_kind<CMB>(C);
    switch (kind) {
    case kCmbCopy:
        Uout[gc] = make_double4(cand[0], cand[1], cand[2], cand[3]);
        return;
    case kCmbRK3b:
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = 0.75 * u[k] + 0.25 * cand[k];
        break;
    case kCmbRK3c:
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = (1.0 / 3) * u[k] + (2.0 / 3) * cand[k];
        break;
    case kCmbRK2b:
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = 0.5 * (u[k] + cand[k]);
        break;
    default: {
        const double v[4] = {uin.x, uin.y, uin.z, uin.w};
        double4 x1 = make_double4(0, 0, 0, 0), x2 = x1, x3 = x1;
        if (kind == kCmbRK4e) { x1 = C.ext->X1[gc]; x2 = C.ext->X2[gc]; x3 = C.ext->X3[gc]; }
        const double a1[4] = {x1.x, x1.y, x1.z, x1.w}, a2[4] = {x2.x, x2.y, x2.z, x2.w}, a3[4] = {x3.x, x3.y, x3.z, x3.w};
        double w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[k] = rk4_out1(kind, u[k], v[k], cand[k], a1[k], a2[k], a3[k], &w[k]);
        if (kind == kCmbRK4d) C.ext->W[gc] = make_double4(w[0], w[1], w[2], w[3]);
    }
    }
    Uout[gc] = make_double4(o[0], o[1], o[2], o[3]);
}

// A stage input's cell (i, j) with the boundary rule: GEN selects the
// general rule (formula boundaries) over the copy boundaries' index map.
template <bool GEN>
__device__ __forceinline__ double4 load_stage(const double4 *__restrict__ U, const Bnd &B, int i, int j,
                                              double xmin_top)
{
    if (GEN) return load_bnd(U, B, i, j, xmin_top);
    return load_copy(U, B.NX, B.NY, B.bc, i, j);
}

// Append the cells of the lanes with `flag` set to list/cnt: one atomic per
// wavefront, slots by prefix count of the ballot.
__device__ __forceinline__ void wave_append(bool flag, int value, int *__restrict__ list, int *cnt)
{
    const unsigned long long m = __ballot(flag);
    if (m == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt, __popcll(m));
    base = __shfl(base, leader);
    if (flag) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// Append up to K values per lane (value k where add[k]) to list/cnt: one
// atomic per wavefront for all of them.  Many cells of a pass can fail at
// once, and one atomic per entry on the single count word serialises in L2.
template <int K>
__device__ __forceinline__ void wave_append_n(const bool (&add)[K], const int (&val)[K], int *__restrict__ list,
                                              int *cnt)
{
    unsigned long long m[K];
    int tot = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        m[k] = __ballot(add[k]);
        tot += __popcll(m[k]);
    }
    if (tot == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)__ballot(1)) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(cnt, tot);
    base = __shfl(base, leader);
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (add[k]) list[base + __popcll(m[k] & lt)] = val[k];
        base += __popcll(m[k]);
    }
}

// Cell update of the face loop (mod_FE.f90:100-101, 134-135) in the order the
// reference's n-outer, l-inner loop applies it to an interior cell.
__device__ __forceinline__ void apply_fluxes(const double4 &u, const double fyb[4], const double fxl[4],
                                             const double fxr[4], const double fyt[4], double dtdx, double dtdy,
                                             double cand[4])
{
    const double uu[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double v = uu[k];
        v = v + dtdy * fyb[k];
        v = v + dtdx * fxl[k];
        v = v - dtdx * fxr[k];
        v = v - dtdy * fyt[k];
        cand[k] = v;
    }
}

// Face flux of the first-order Godunov method (FOG_, mod_FOG.f90:11-25): the
// face states are the two cells' values themselves (no reconstruction sum), one
// Gauss point, L = 0 + F * gauss_weight(1, 1) (mod_FE.f90:80-98).
template <int DIR, int FLUX>
__device__ __forceinline__ void fog_flux(const DevTables *__restrict__ T, const double4 &a, const double4 &b,
                                         double L[4])
{
    const double ua[4] = {a.x, a.y, a.z, a.w}, ub[4] = {b.x, b.y, b.z, b.w};
    double F[4];
    riemann<FLUX, DIR>(ua, ub, F);
    const double w = T->gw[0];
#pragma unroll
    for (int k = 0; k < 4; ++k) L[k] = 0.0 + F[k] * w;
}

// XCD-aware tile order: consecutive workgroups are dispatched round-robin over
// the 8 XCDs; give each XCD a contiguous band of tiles so halo reads of
// neighbouring tiles hit the same L2.  nb (the grid size) is a multiple of 8,
// which makes the map a bijection on [0, nb); ids past the last tile return -1.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int per = nb / 8;
    const int x = b % 8, k = b / 8;
    const int t = x * per + k;
    return t < nb ? t : -1;
}

// Flag the 5-point neighbourhood of a failing cell for recomputation in the
// next pass (DetCell, mod_subroutine_MOOD.f90:165-169): each cell enters the
// next pass's list once, deduplicated by the pass stamp.  Called by every
// lane of the wavefront (dec: this lane's cell fails).
__device__ __forceinline__ void mark_neighbourhood(bool dec, int i, int j, int NX, int NY, int stamp, int *mark,
                                                   int *list, int *cnt)
{
    const int ni[5] = {i, i - 1, i, i + 1, i}, nj[5] = {j - 1, j, j, j, j + 1};
    bool add[5];
    int d[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        d[k] = nj[k] * NX + ni[k];
        add[k] = dec && !(ni[k] < 0 || ni[k] >= NX || nj[k] < 0 || nj[k] >= NY) && atomicExch(&mark[d[k]], stamp) != stamp;
    }
    wave_append_n<5>(add, d, list, cnt);
}

template <int MORD, int NGP, int FLUX, bool GEN, int CMB>
__global__ __launch_bounds__(kStageThreads, stage_waves(MORD)) void k_stage(StageArgs a)
{
    constexpr int kTX = stage_tx(MORD), kTY = stage_ty(MORD);
    constexpr int H = kHalo<MORD>;
    constexpr int PW = kTX + 2 * H, PH = kTY + 2 * H;
    __shared__ double4 sU[PW * PH];
    __shared__ double4 sFx[kTY * (kTX + 1)];
    __shared__ double4 sFy[(kTY + 1) * kTX];

    const int NX = a.B.NX, NY = a.B.NY;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile < 0 || tile >= a.ntiles) return;
    const int x0 = (tile % a.tiles_x) * kTX, y0 = (tile / a.tiles_x) * kTY;
    const int tid = threadIdx.x;
    const double xmin_top = GEN ? stage_xmin_top(a.B, a.tstate, a.C.trk) : 0.0;

    for (int idx = tid; idx < PW * PH; idx += kStageThreads) {
        const int ti = idx % PW, tj = idx / PW;
        sU[idx] = load_stage<GEN>(a.Uin, a.B, x0 - H + ti, y0 - H + tj, xmin_top);
    }
    // a priori orders of the tile and its one-cell ring (ghost cells: Mord)
    constexpr int OW = kTX + 2;
    __shared__ int8_t sO[(kTY + 2) * OW];
    if (a.apriori) {
        for (int idx = tid; idx < (kTY + 2) * OW; idx += kStageThreads) {
            const int gi = x0 - 1 + idx % OW, gj = y0 - 1 + idx / OW;
            sO[idx] = (gi < 0 || gi >= NX || gj < 0 || gj >= NY) ? (int8_t)MORD : a.pred[(size_t)gj * NX + gi];
        }
    }
    __syncthreads();

    const DevTables *__restrict__ T = a.T;
    const LdsAcc acc{sU, PW};
    constexpr int NFX = (kTX + 1) * kTY, NFY = kTX * (kTY + 1);
    for (int f = tid; f < NFX + NFY; f += kStageThreads) {
        double L[4];
        if (f < NFX) {
            const int fy = f / (kTX + 1), fx = f % (kTX + 1);
            if (MORD == 1) {  // FOG
                fog_flux<0, FLUX>(T, acc(fx - 1 + H, fy + H), acc(fx + H, fy + H), L);
            } else if (!a.apriori) {
                face_flux<MORD, NGP, 0, FLUX>(T, acc, fx - 1 + H, fy + H, fx + H, fy + H, L);
            } else if (MORD == 3 && a.recon == kReconCC) {
                const int o = min(sO[(fy + 1) * OW + fx], sO[(fy + 1) * OW + fx + 1]);
                double va1, va2, vb1, vb2;
                vnn_at(a.vnn, NX, NY, x0 + fx - 1, y0 + fy, va1, va2);
                vnn_at(a.vnn, NX, NY, x0 + fx, y0 + fy, vb1, vb2);
                face_flux_cc<NGP, 0, FLUX>(T, acc, o, fx - 1 + H, fy + H, fx + H, fy + H, va1, va2, vb1, vb2, L);
            } else {
                const int o = min(sO[(fy + 1) * OW + fx], sO[(fy + 1) * OW + fx + 1]);
                face_flux_dyn<MORD, NGP, 0, FLUX>(T, acc, o, fx - 1 + H, fy + H, fx + H, fy + H, L);
            }
            sFx[f] = make_double4(L[0], L[1], L[2], L[3]);
        } else {
            const int g = f - NFX;
            const int gy = g / kTX, gx = g % kTX;
            if (MORD == 1) {
                fog_flux<1, FLUX>(T, acc(gx + H, gy - 1 + H), acc(gx + H, gy + H), L);
            } else if (!a.apriori) {
                face_flux<MORD, NGP, 1, FLUX>(T, acc, gx + H, gy - 1 + H, gx + H, gy + H, L);
            } else if (MORD == 3 && a.recon == kReconCC) {
                const int o = min(sO[gy * OW + gx + 1], sO[(gy + 1) * OW + gx + 1]);
                double va1, va2, vb1, vb2;
                vnn_at(a.vnn, NX, NY, x0 + gx, y0 + gy - 1, va1, va2);
                vnn_at(a.vnn, NX, NY, x0 + gx, y0 + gy, vb1, vb2);
                face_flux_cc<NGP, 1, FLUX>(T, acc, o, gx + H, gy - 1 + H, gx + H, gy + H, va1, va2, vb1, vb2, L);
            } else {
                const int o = min(sO[gy * OW + gx + 1], sO[(gy + 1) * OW + gx + 1]);
                face_flux_dyn<MORD, NGP, 1, FLUX>(T, acc, o, gx + H, gy - 1 + H, gx + H, gy + H, L);
            }
            sFy[g] = make_double4(L[0], L[1], L[2], L[3]);
        }
    }
    __syncthreads();

    const double dtdx = a.tstate[2], dtdy = a.tstate[3];
    const double dtg = GEN ? a.tstate[1] * kGravity : 0.0;
    const MoodParams P{a.dmp, a.u2, a.u2tol, a.dx, a.dy, a.recon};
    const int stamp1 = *a.stamp + 1;
    for (int c = tid; c < kTX * kTY; c += kStageThreads) {
        const int cx = c % kTX, cy = c / kTX;
        const int gi = x0 + cx, gj = y0 + cy;
        const bool inside = gi < NX && gj < NY;
        bool dec = false;
        if (inside) {
            const double4 u = sU[(cy + H) * PW + cx + H];
            const double4 b = sFy[cy * kTX + cx], l = sFx[cy * (kTX + 1) + cx];
            const double4 r = sFx[cy * (kTX + 1) + cx + 1], t = sFy[(cy + 1) * kTX + cx];
            const double fyb[4] = {b.x, b.y, b.z, b.w}, fxl[4] = {l.x, l.y, l.z, l.w};
            const double fxr[4] = {r.x, r.y, r.z, r.w}, fyt[4] = {t.x, t.y, t.z, t.w};
            double cand[4];
            apply_fluxes(u, fyb, fxl, fxr, fyt, dtdx, dtdy, cand);
            // CellGPO := Mord (mod_FE.f90:48), or the a priori order
            const int me = a.apriori ? (int)sO[(cy + 1) * OW + cx + 1] : MORD;
            const size_t gc = (size_t)gj * NX + gi;
            if (me > 1 && !a.frozen) {
                uint8_t fl;
                dec = decide_record(P, T, acc, cx + H, cy + H, cand, fl);
                a.dflag[gc] = fl;
            }
            const double4 un = comb_kind<CMB>(a.C) != kCmbCopy ? a.C.Un[gc] : u;
            write_stage_output<GEN, CMB>(a.C, a.Uout, gc, GEN && a.C.band && in_band(gi, gj, NX, NY), un, u, cand,
                                         dtg);
            // a failing cell takes the cascade at once
            a.gpo[gc] = pack_order(dec ? lower_order(me, a.cascade) : me, me, 0);
        }
        mark_neighbourhood(dec, gi, gj, NX, NY, stamp1, a.mark, a.slist1, a.cnt_s1);
        const unsigned long long m = __ballot(dec);
        if (__lane_id() == 0 && m) atomicAdd(a.cnt_f0, (int)__popcll(m));
    }
}

// A workgroup barrier that orders LDS accesses only: global loads still in
// flight stay in flight (a full __syncthreads waits for them), so a barrier
// that guards LDS buffers does not serialise unrelated global latencies.
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int load_relaxed(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Memory operations of the MOOD passes at a chosen coherence scope: device
// scope while all workgroups take part in a pass, workgroup scope while
// workgroup 0 runs passes alone (its own writes then stay in its XCD's L2).
template <int SC>
struct Scoped {
    static __device__ __forceinline__ int load(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, SC); }
    static __device__ __forceinline__ order_t load(const order_t *p)
    {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, SC);
    }
    static __device__ __forceinline__ void store(order_t *p, order_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, SC); }
    static __device__ __forceinline__ int add(int *p, int v) { return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, SC); }
    static __device__ __forceinline__ int exch(int *p, int v) { return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, SC); }
};

// The next pass's list of a grid-wide pass: the 5-point neighbourhood of a
// failing cell, each cell listed once (deduplicated by the pass stamp in the
// global mark array; DetCell, mod_subroutine_MOOD.f90:165-169).  Called by
// every lane of the wavefront: `on` asks for neighbour k (0 bottom, 1 left,
// 2 the cell, 3 right, 4 top) of cell (i, j); the new entries go to the list
// with one atomic per wavefront.
struct GridMark {
    int *mark, *list, *cnt;
    int stamp;
    __device__ __forceinline__ void operator()(bool on, int i, int j, int k, int NX, int NY) const
    {
        const int ni = i + (k == 1 ? -1 : k == 3 ? 1 : 0), nj = j + (k == 0 ? -1 : k == 4 ? 1 : 0);
        bool add[1] = {false};
        int d[1] = {nj * NX + ni};
        if (on && !(ni < 0 || ni >= NX || nj < 0 || nj >= NY))
            add[0] = Scoped<__HIP_MEMORY_SCOPE_AGENT>::exch(&mark[d[0]], stamp) != stamp;
        wave_append_n<1>(add, d, list, cnt);
    }
};

// The same for the passes workgroup 0 runs alone: the list lives in LDS and
// an LDS hash set (open addressing, -1 = empty) deduplicates it, so these
// passes make no global atomic round trips.  The set is at most half full
// (static_assert below), so a probe meets an empty slot within the table; the
// probe loops still stop after one sweep of the table, so that damaged LDS
// contents give wrong bits, which the parity tests catch, rather than a hang.
constexpr int kSoloHashBits = 9;
constexpr int kSoloHash = 1 << kSoloHashBits;  // slots; a solo pass lists at most 5 * kSoloMax cells
__device__ __forceinline__ bool lds_set_insert(int *hash, int bits, int key)
{
    unsigned h = ((unsigned)key * 0x9E3779B1u) >> (32 - bits);
    for (int probe = 0; probe < (1 << bits); ++probe) {
        const int old = atomicCAS(&hash[h], -1, key);
        if (old == -1) return true;
        if (old == key) return false;
        h = (h + 1) & ((1u << bits) - 1u);
    }
    return false;
}
__device__ __forceinline__ bool lds_set_has(const int *hash, int bits, int key)
{
    unsigned h = ((unsigned)key * 0x9E3779B1u) >> (32 - bits);
    for (int probe = 0; probe < (1 << bits); ++probe) {
        const int v = hash[h];
        if (v == key) return true;
        if (v == -1) return false;
        h = (h + 1) & ((1u << bits) - 1u);
    }
    return false;
}
static_assert(5 * kSoloMax <= kSoloHash / 2, "the solo hash set must stay at most half full");

// Face-flux cache of the passes workgroup 0 runs alone.  A face's flux at a
// given order (the quadrature sum of its Gauss-point Riemann fluxes) depends
// only on the stage input and that order, so within a stage it can be reused
// bit for bit.  Entries live in a direct-mapped window of kFcW x kFcW cells
// (a cell owns its left x-face and bottom y-face; two slots per face, by the
// parity of the order index, so a face's current order and the next lower
// one coexist).  A slot's 64-bit tag holds the face-and-order key and a
// version: writers claim a slot by compare-and-swap (version + 1, busy), write
// the flux, and publish (version + 2, key); readers compare the tag before and
// after reading the flux, so a torn read is a miss, never a wrong value.
constexpr int kFcW = 16;
constexpr int kFcSlots = kFcW * kFcW * 4;
constexpr unsigned kFcEmpty = 0xFFFFFFFFu, kFcBusy = 0xFFFFFFFEu;
__device__ __forceinline__ int fc_slot(int dir, int fi, int fj, int oi)
{
    return ((((fj & (kFcW - 1)) * kFcW + (fi & (kFcW - 1))) * 2 + dir) * 2) + (oi & 1);
}
// x-face (fi, fj) lies between cells (fi - 1, fj) and (fi, fj), fi in [0, NX];
// y-face (fi, fj) between (fi, fj - 1) and (fi, fj), fj in [0, NY]
__device__ __forceinline__ unsigned fc_key(int dir, int fi, int fj, int oi, int NX, int NY)
{
    const unsigned id = dir == 0 ? (unsigned)fj * (unsigned)(NX + 1) + (unsigned)fi
                                 : (unsigned)(NX + 1) * (unsigned)NY + (unsigned)fj * (unsigned)NX + (unsigned)fi;
    return id * 4u + (unsigned)oi;
}
// whether every face key of an NX x NY grid stays below kFcBusy
__host__ __device__ __forceinline__ bool fc_fits(int NX, int NY)
{
    const unsigned long long faces = (unsigned long long)(NX + 1) * NY + (unsigned long long)NX * (NY + 1);
    return 4ull * faces < (unsigned long long)kFcBusy;
}

// A stage input's 5-point neighbourhood held in registers, as an accessor in
// the staged-box coordinates of the detection (centre at (h, h)).
struct FiveAcc {
    double4 c, l, r, b, t;
    int h;
    __device__ __forceinline__ double4 operator()(int i, int j) const
    {
        const int di = i - h, dj = j - h;
        return di < 0 ? l : di > 0 ? r : dj < 0 ? b : dj > 0 ? t : c;
    }
};

// LDS writes of a wavefront's lanes visible to its other lanes (no workgroup
// barrier: the lanes that stage a cell's neighbourhood are the lanes that read it).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One MOOD correction pass (mod_FE.f90:69-174) over a list of cells, shared
// by the persistent kernel and the per-pass launches.  Pass p recomputes the
// cells of its list from scratch with the orders they had when the pass
// began; a cell that fails takes the cascade at once (version p) and flags its
// neighbourhood for pass p+1.  Each cell is handled by G lanes, one per (face,
// Gauss point): the lanes reconstruct and solve their Riemann problem, then
// cross-lane shuffles form the quadrature sum per face and the four-face
// update in the reference's order on the cell's first lane.
//
// A pass is a chain of dependent memory round trips, so the loads that do not
// depend on the staged neighbourhood (the face's two order words, the cell's
// own order word and step-start state, the next trip's list entry) are
// issued together with the neighbourhood loads.
//
// WIDE selects the form of the wide launch (k_mood_wide): grid-wide passes
// only, in workgroups of kWideThreads threads whose LDS holds the tables and
// the staged neighbourhoods alone, so that several workgroups share a CU.
template <int MORD, int NGP, int FLUX, bool GEN, bool WIDE = false>
struct MoodPass {
    static constexpr int NGPP = NGP <= 1 ? 1 : NGP <= 2 ? 2 : NGP <= 4 ? 4 : 8;
    static constexpr int G = 4 * NGPP;
    static_assert(64 % G == 0, "lanes per cell must divide the wavefront");
    // Each cell's (2H+1)^2 neighbourhood of the stage input is staged in LDS by
    // its G lanes (all loads in flight at once); reconstruction and detection
    // then read LDS.  H covers the stencils of the cell and its four
    // neighbours and the detection's 2-cell reach.
    static constexpr int HB = kHalo<MORD>;
    static constexpr int BOX = 2 * HB + 1;
    static constexpr int NTH = WIDE ? kWideThreads : kMoodThreads;
    static constexpr int CPB = NTH / G;  // cells per workgroup per trip
    // speculative jobs of one workgroup-0 pass: for each of the first 16 listed
    // cells, the cell at its next lower order and its four neighbours
    static constexpr int FCN = MORD == 1 ? 1 : kFcSlots;
    static constexpr int PW = 64 / G;  // jobs per wavefront
    static constexpr int KK = (5 + G - 1) / G;  // neighbours of a failing cell per lane
    struct LdsWide {
        DevTables T;
        double4 nb[CPB][BOX * BOX];
        // a trip's new list entries and failures, gathered over the workgroup
        // (double-buffered by trip parity) before one device atomic each
        int wbuf[2][5 * CPB];
        int wcnt[2], wfail[2], wbase;
    };
    struct LdsFull {
        DevTables T;  // the GP tables (the reconstruction's weights are read per lane)
        double4 nb[CPB][BOX * BOX];
        // face-flux cache of the workgroup-0 passes (none at first order, where
        // a pass never runs: FOG has no detection)
        unsigned long long ctag[FCN];
        double4 cflux[FCN];
        // lists of the passes workgroup 0 runs alone (a failing cell adds at
        // most five cells, so a list of at most kSoloMax cells yields at most
        // 5x).  List q + 1 is built during pass q in hash[(q + 1) % 3]; pass q
        // tells list q's cells by hash[q % 3] and clears hash[(q + 2) % 3] and
        // cnt[(q + 2) % 3], so a pass needs no barrier but the one before it.
        int list[2][5 * kSoloMax];
        int hash[3][kSoloHash];
        int cnt[3];   // list q has cnt[q % 3] cells
        int fail[3];  // failures of pass q in fail[q % 3]
        // the other workgroups' precompute (persistent kernel): workgroup
        // 0's pass and list as last read, cells taken this round
        int hseq, hm, hn, hstop;
        int hlist[16];
        // workgroup 0's passes: the listed cells' wavefronts' barrier, and
        // what they tell the speculative wavefronts
        int ebar_cnt, ebar_gen, spass, sdone, qend;
    };
    using Lds = typename std::conditional<WIDE, LdsWide, LdsFull>::type;

    const MoodArgs &a;
    Lds &L;
    int NX, NY, lane, gl, gbase, f, jg, fdir, fa, fb, oai, oaj, obi, obj, cs, stamp_base;
    bool fc_on, g_on;
    unsigned gst;  // this stage's tag in the device-wide cache
    // trips of grid-wide passes run by this workgroup since the launch began:
    // their parity picks wide_append's LDS buffers, alternating across passes
    // as well as within one
    int trip = 0;
    MoodParams P;
    double dtdx, dtdy, dtg, xmin_top;

    __device__ __forceinline__ MoodPass(const MoodArgs &args, Lds &lds) : a(args), L(lds)
    {
        NX = a.B.NX;
        NY = a.B.NY;
        lane = __lane_id();
        gl = lane % G;
        gbase = lane - gl;
        f = gl / NGPP;
        jg = gl % NGPP;
        // face f of a cell (i, j): 0 bottom y-face, 1 left x-face, 2 right x-face, 3 top y-face
        fdir = (f == 1 || f == 2) ? 0 : 1;
        fa = fdir == 0 ? kR : kT;
        fb = fdir == 0 ? kL : kB;
        oai = f == 1 ? -1 : 0;
        oaj = f == 0 ? -1 : 0;
        obi = f == 2 ? 1 : 0;
        obj = f == 3 ? 1 : 0;
        cs = threadIdx.x / G;
        stamp_base = *a.stamp;
        fc_on = MORD > 1 && a.fcache && fc_fits(NX, NY);
        g_on = fc_on && a.gfc.bits > 0;
        gst = (unsigned)stamp_base + 1u;
        P = MoodParams{a.dmp, a.u2, a.u2tol, a.dx, a.dy, a.recon};
        dtdx = a.tstate[2];
        dtdy = a.tstate[3];
        dtg = GEN ? a.tstate[1] * kGravity : 0.0;
        xmin_top = GEN ? stage_xmin_top(a.B, a.tstate, a.C.trk) : 0.0;
    }

    // Copy the GP tables to LDS (all threads of the workgroup, at the start
    // of a launch).
    __device__ __forceinline__ void init_lds()
    {
        const double *src = reinterpret_cast<const double *>(a.T);
        double *dst = reinterpret_cast<double *>(&L.T);
        for (int k = threadIdx.x; k < (int)(sizeof(DevTables) / sizeof(double)); k += blockDim.x) dst[k] = src[k];
        if constexpr (!WIDE) {
            for (int k = threadIdx.x; k < FCN; k += blockDim.x) L.ctag[k] = kFcEmpty;
        } else if (threadIdx.x < 2) {
            L.wcnt[threadIdx.x] = 0;
            L.wfail[threadIdx.x] = 0;
        }
        __syncthreads();
    }

    // One pass over the nS cells of list S, spread over the lanes [t0, t0 +
    // nthr) of the grid (all workgroups) or of workgroup 0 alone; d0 is this
    // lane's entry of the first trip (loaded with the count).  cnt_fail counts
    // the cells that fail; mark lists their neighbourhoods for pass p+1.
    template <int SC, class Mark>
    __device__ __forceinline__ void run(int p, int nS, int d0, int t0, int nthr, const int *S, int *cnt_fail,
                                        const Mark &mark)
    {
        using M = Scoped<SC>;
        const DevTables *__restrict__ T = &L.T;
        const LdsAcc lacc{&L.nb[cs][0], BOX};
        auto order_of = [&](int ii, int jj) -> int {
            if (ii < 0 || ii >= NX || jj < 0 || jj >= NY) return MORD;  // ghost cells keep Mord
            return order_at_pass(M::load(&a.gpo[(size_t)jj * NX + ii]), p);
        };
        const int per_iter = nthr / G;
        const int niter = (nS + per_iter - 1) / per_iter;
        int dnext = d0;
        for (int it = 0; it < niter; ++it) {
            const int t = it * per_iter + t0 / G;
            const bool valid = t < nS;
            const int d = valid ? dnext : 0;
            const int i = d % NX, j = d / NX;
            // independent of the neighbourhood: issue before staging it
            int ord = 1;
            if (valid && jg < NGP) ord = min(order_of(i + oai, j + oaj), order_of(i + obi, j + obj));
            int me = 1;
            double4 un = make_double4(0.0, 0.0, 0.0, 0.0);
            if (valid && gl == 0) {
                me = order_at_pass(M::load(&a.gpo[d]), p);
                if (a.C.kind != kCmbCopy) un = a.C.Un[d];
            }
            const int tn = t + per_iter;
            if (it + 1 < niter && tn < nS) dnext = S[tn];
            uint8_t fl = 0;
            if (valid && gl == 0 && me > 1) fl = a.dflag[d];
            // a group's staged neighbourhood is written and read by its own
            // lanes (one wavefront): no workgroup barrier around it
            wave_lds_sync();  // the previous trip's readers are done with nb
            for (int q = gl; q < BOX * BOX; q += G)
                L.nb[cs][q] = load_stage<GEN>(a.Uin, a.B, i - HB + q % BOX, j - HB + q / BOX, xmin_top);
            wave_lds_sync();
            double F[4] = {0.0, 0.0, 0.0, 0.0};
            if (valid && jg < NGP) {
                double va1 = 0.0, va2 = 1.0, vb1 = 0.0, vb2 = 1.0;
                const int recon = MORD == 3 ? a.recon : kReconGP;
                if (recon == kReconCC) {
                    vnn_at(a.vnn, NX, NY, i + oai, j + oaj, va1, va2);
                    vnn_at(a.vnn, NX, NY, i + obi, j + obj, vb1, vb2);
                }
                face_gp_flux_dyn<MORD, FLUX>(T, lacc, ord, fa, fb, fdir, jg, HB + oai, HB + oaj, HB + obi, HB + obj, F,
                                             recon, va1, va2, vb1, vb2);
            }
            // quadrature over the face's Gauss points, j = 1..ngp from 0
            double Lf[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int jj = 0; jj < NGP; ++jj) {
                const int src = gbase + f * NGPP + jj;
                const double w = T->gw[jj];
#pragma unroll
                for (int k = 0; k < 4; ++k) Lf[k] = Lf[k] + __shfl(F[k], src) * w;
            }
            double fyb[4], fxl[4], fxr[4], fyt[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                fyb[k] = __shfl(Lf[k], gbase);
                fxl[k] = __shfl(Lf[k], gbase + NGPP);
                fxr[k] = __shfl(Lf[k], gbase + 2 * NGPP);
                fyt[k] = __shfl(Lf[k], gbase + 3 * NGPP);
            }
            bool dec = false;
            if (valid && gl == 0) {
                double cand[4];
                const double4 uin = lacc(HB, HB);
                apply_fluxes(uin, fyb, fxl, fxr, fyt, dtdx, dtdy, cand);
                if (me > 1) dec = decide_with_flags(P, lacc, HB, HB, cand, fl);
                write_stage_output<GEN, kCmbRuntime>(a.C, a.Uout, d, GEN && a.C.band && in_band(i, j, NX, NY), un, uin,
                                                     cand, dtg);
                if (dec) M::store(&a.gpo[d], pack_order(lower_order(me, a.cascade), me, p));
            }
            // the failures, one count per wavefront; a failing cell's 5-point
            // neighbourhood enters list p + 1, one neighbour per lane of its group
            const unsigned long long fm = __ballot(dec);
            const bool decg = __shfl(dec ? 1 : 0, gbase) != 0;
            if constexpr (WIDE) {
                wide_append(trip++ & 1, decg, fm, i, j, cnt_fail, mark);
            } else {
                if (fm && lane == __ffsll((long long)fm) - 1) M::add(cnt_fail, (int)__popcll(fm));
#pragma unroll
                for (int k0 = 0; k0 < 5; k0 += G) mark(decg && k0 + gl < 5, i, j, k0 + gl, NX, NY);
            }
        }
    }

    // The wide launch's list appends and failure count for one trip: every
    // wavefront of a wide launch appends at once, and one device atomic per
    // wavefront on the single count word still serialises (thousands of
    // wavefronts), so the workgroup gathers its trip's entries and failures
    // in LDS first, and one thread reserves the workgroup's range.  Called by
    // every thread of the workgroup once per trip; b is the parity of the
    // workgroup's trip count since the launch (init_lds clears both buffers'
    // counts; trip t clears those of trip t + 1 between its two barriers, and
    // trip t + 1 does not write the buffer trip t copies out).  The persistent
    // kernel's short, latency-bound passes keep one atomic per wavefront: the
    // two barriers cost them about what the atomics would.
    __device__ __forceinline__ void wide_append(int b, bool decg, unsigned long long fm, int i, int j,
                                                int *cnt_fail, const GridMark &mk)
    {
        constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
        const unsigned long long lt = (1ull << lane) - 1ull;
        const int leader = __ffsll((long long)__ballot(1)) - 1;
#pragma unroll
        for (int kk = 0; kk < KK; ++kk) {
            const int k = kk * G + gl;  // neighbour k of the group's cell: 0 bottom, 1 left, 2 the cell, 3 right, 4 top
            const int ni = i + (k == 1 ? -1 : k == 3 ? 1 : 0), nj = j + (k == 0 ? -1 : k == 4 ? 1 : 0);
            const int d = nj * NX + ni;
            bool add = false;
            if (decg && k < 5 && !(ni < 0 || ni >= NX || nj < 0 || nj >= NY))
                add = Scoped<__HIP_MEMORY_SCOPE_AGENT>::exch(&mk.mark[d], mk.stamp) != mk.stamp;
            const unsigned long long m = __ballot(add);
            int off = 0;
            if (lane == leader && m) off = __hip_atomic_fetch_add(&L.wcnt[b], (int)__popcll(m), __ATOMIC_RELAXED, WG);
            off = __shfl(off, leader);
            if (add) L.wbuf[b][off + __popcll(m & lt)] = d;
        }
        if (lane == leader && fm) __hip_atomic_fetch_add(&L.wfail[b], (int)__popcll(fm), __ATOMIC_RELAXED, WG);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int n = L.wcnt[b], f = L.wfail[b];
            L.wbase = n ? atomicAdd(mk.cnt, n) : 0;
            if (f) Scoped<__HIP_MEMORY_SCOPE_AGENT>::add(cnt_fail, f);
            L.wcnt[b ^ 1] = 0;
            L.wfail[b ^ 1] = 0;
        }
        __syncthreads();
        const int n = L.wcnt[b], base = L.wbase;
        for (int q = threadIdx.x; q < n; q += NTH) mk.list[base + q] = L.wbuf[b][q];
    }

    // Face-flux cache access (workgroup 0's passes).
    __device__ __forceinline__ bool fc_get(unsigned key, int slot, double4 &F)
    {
        constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
        unsigned long long *tp = &L.ctag[slot];
        const unsigned long long t1 = __hip_atomic_load(tp, __ATOMIC_RELAXED, WG);
        if ((unsigned)t1 != key) return false;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        F = L.cflux[slot];
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        return __hip_atomic_load(tp, __ATOMIC_RELAXED, WG) == t1;
    }
    __device__ __forceinline__ void fc_put(unsigned key, int slot, const double4 &F)
    {
        constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
        unsigned long long *tp = &L.ctag[slot];
        unsigned long long t = __hip_atomic_load(tp, __ATOMIC_RELAXED, WG);
        if ((unsigned)t == key || (unsigned)t == kFcBusy) return;
        const unsigned long long ver = t >> 32;
        if (!__hip_atomic_compare_exchange_strong(tp, &t, ((ver + 1) << 32) | kFcBusy, __ATOMIC_RELAXED,
                                                  __ATOMIC_RELAXED, WG))
            return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        L.cflux[slot] = F;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __hip_atomic_store(tp, ((ver + 2) << 32) | key, __ATOMIC_RELAXED, WG);
    }

    // Device-wide cache access.  Probe sequences are short and bounded; a
    // full or crowded table only costs misses.  A writer claims a free slot
    // (an empty one or one of an earlier stage) by compare-and-swap of its
    // key word, writes the flux and then `ready`, with a device-scope release
    // between; a reader that finds the key and `ready` reads the flux after
    // them.  The flux of a face at an order is the same bits whoever computes it.
    static constexpr int kGProbe = 8;
    __device__ __forceinline__ unsigned g_hash(unsigned key) const
    {
        return (key * 0x9E3779B1u) >> (32 - a.gfc.bits);
    }
    __device__ __forceinline__ bool gfc_get(unsigned key, double4 &F) const
    {
        constexpr int AG = __HIP_MEMORY_SCOPE_AGENT;
        const unsigned long long want = ((unsigned long long)gst << 32) | key;
        const unsigned mask = (1u << a.gfc.bits) - 1u;
        unsigned h = g_hash(key);
        for (int pr = 0; pr < kGProbe; ++pr) {
            const unsigned long long kw = __hip_atomic_load(&a.gfc.kw[h], __ATOMIC_RELAXED, AG);
            const unsigned r = __hip_atomic_load(&a.gfc.ready[h], __ATOMIC_RELAXED, AG);
            if (kw == want) {
                if (r != gst) return false;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                const double *src = reinterpret_cast<const double *>(&a.gfc.data[h]);
                F.x = __hip_atomic_load(src + 0, __ATOMIC_RELAXED, AG);
                F.y = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, AG);
                F.z = __hip_atomic_load(src + 2, __ATOMIC_RELAXED, AG);
                F.w = __hip_atomic_load(src + 3, __ATOMIC_RELAXED, AG);
                return true;
            }
            if ((unsigned)(kw >> 32) != gst) return false;  // the first free slot ends the sequence
            h = (h + 1) & mask;
        }
        return false;
    }
    __device__ __forceinline__ void gfc_put(unsigned key, const double4 &F) const
    {
        constexpr int AG = __HIP_MEMORY_SCOPE_AGENT;
        const unsigned long long want = ((unsigned long long)gst << 32) | key;
        const unsigned mask = (1u << a.gfc.bits) - 1u;
        unsigned h = g_hash(key);
        for (int pr = 0; pr < kGProbe; ++pr) {
            unsigned long long kw = __hip_atomic_load(&a.gfc.kw[h], __ATOMIC_RELAXED, AG);
            if (kw == want) return;
            if ((unsigned)(kw >> 32) != gst &&
                __hip_atomic_compare_exchange_strong(&a.gfc.kw[h], &kw, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED, AG)) {
                double *dst = reinterpret_cast<double *>(&a.gfc.data[h]);
                __hip_atomic_store(dst + 0, F.x, __ATOMIC_RELAXED, AG);
                __hip_atomic_store(dst + 1, F.y, __ATOMIC_RELAXED, AG);
                __hip_atomic_store(dst + 2, F.z, __ATOMIC_RELAXED, AG);
                __hip_atomic_store(dst + 3, F.w, __ATOMIC_RELAXED, AG);
                __hip_atomic_store(&a.gfc.ready[h], gst, __ATOMIC_RELEASE, AG);
                return;
            }
            if (kw == want) return;  // claimed for this key meanwhile
            h = (h + 1) & mask;
        }
    }

    // Speculative job number t (this lane's group's) of workgroup-0 pass q,
    // listed by the whole wavefront from list q: for each of its first 16
    // cells the cell at its next lower order (its faces if it fails now),
    // then each neighbour not in the list at the current orders (the faces of
    // the cells that enter list q + 1 if it fails).  Every wavefront that
    // runs speculative jobs lists them itself, in the same order.  Which jobs
    // run changes which fluxes are cached, never a result.  Returns false
    // when there is no job t.
    __device__ __forceinline__ bool spec_job(int q, int m, int t, int &d, int &var)
    {
        const int k = lane >> 2, dd = lane & 3;
        bool el = false, ec = false;
        int vl = 0, vc = 0;
        const int c = k < m ? L.list[q & 1][k] : -1;
        if (c >= 0 && c < NX * NY) {
            const int ci = c % NX, cj = c / NX;
            el = dd == 0;
            vl = 2 * c + 1;
            const int ni = ci + (dd == 1 ? -1 : dd == 2 ? 1 : 0), nj = cj + (dd == 0 ? -1 : dd == 3 ? 1 : 0);
            if (ni >= 0 && ni < NX && nj >= 0 && nj < NY) {
                const int n = nj * NX + ni;
                ec = !lds_set_has(L.hash[q % 3], kSoloHashBits, n);
                vc = 2 * n;
            }
        }
        const unsigned long long bl = __ballot(el), bc = __ballot(ec);
        const unsigned long long lt = (1ull << lane) - 1ull;
        const int nl = __popcll(bl), il = __popcll(bl & lt), ic = nl + __popcll(bc & lt);
        int v = -1;
#pragma unroll
        for (int g = 0; g < PW; ++g) {
            const int tg = __shfl(t, g * G);
            const unsigned long long sl = __ballot(el && il == tg), sc = __ballot(ec && ic == tg);
            const int src = sl ? __ffsll((long long)sl) - 1 : sc ? __ffsll((long long)sc) - 1 : 0;
            const int val = __shfl(sl ? vl : vc, src);
            if (lane / G == g && (sl | sc)) v = val;
        }
        if (v < 0) return false;
        d = v >> 1;
        var = v & 1;
        return true;
    }

    // One job of a workgroup-0 pass q on this lane's group of G lanes:
    // kind 1 recomputes listed cell d (the pass's work, as in run()), kind 2
    // computes the face fluxes of cell d at the current orders (var 0) or with
    // the cell at its next lower order (var 1) into the cache, kind 3 (the
    // other workgroups) computes all four face fluxes of cell d at order
    // 2 var + 1 into the device-wide cache, kind 0 idles.  Each face flux
    // comes from a cache (workgroup 0: its LDS cache, then the device-wide
    // one) or is computed and cached; the staged neighbourhood is written and
    // read by the group's own lanes, so jobs of different wavefronts never
    // wait for each other.
    __device__ __forceinline__ void job(int q, int kind, int d, int var)
    {
        using M = Scoped<__HIP_MEMORY_SCOPE_WORKGROUP>;
        const DevTables *__restrict__ T = &L.T;
        const bool valid = kind != 0;
        const int i = d % NX, j = d / NX;
        auto order_of = [&](int ii, int jj) -> int {
            if (ii < 0 || ii >= NX || jj < 0 || jj >= NY) return MORD;  // ghost cells keep Mord
            return order_at_pass(M::load(&a.gpo[(size_t)jj * NX + ii]), q);
        };
        // this lane's face f: the neighbour across it, the face's coordinates
        const int nbi = i + (f == 1 ? -1 : f == 2 ? 1 : 0), nbj = j + (f == 0 ? -1 : f == 3 ? 1 : 0);
        const int fi = i + (f == 2 ? 1 : 0), fj = j + (f == 3 ? 1 : 0);
        int me = 1, ord = 1;
        if (kind == 1 || kind == 2) {
            me = order_of(i, j);
            ord = min(var ? lower_order(me, a.cascade) : me, order_of(nbi, nbj));
        } else if (kind == 3) {
            ord = 2 * var + 1;
        }
        uint8_t fl = 0;
        const double4 zero = make_double4(0.0, 0.0, 0.0, 0.0);
        double4 un = zero, c0 = zero, cl = zero, cr = zero, cb = zero, ct = zero;
        if (kind == 1 && gl == 0) {
            if (a.C.kind != kCmbCopy) un = a.C.Un[d];
            fl = a.dflag[d];
            c0 = load_stage<GEN>(a.Uin, a.B, i, j, xmin_top);
            cl = load_stage<GEN>(a.Uin, a.B, i - 1, j, xmin_top);
            cr = load_stage<GEN>(a.Uin, a.B, i + 1, j, xmin_top);
            cb = load_stage<GEN>(a.Uin, a.B, i, j - 1, xmin_top);
            ct = load_stage<GEN>(a.Uin, a.B, i, j + 1, xmin_top);
        }
        const int oi = (ord - 1) >> 1;
        const unsigned key = fc_on ? fc_key(fdir, fi, fj, oi, NX, NY) : 0u;
        const int slot = fc_slot(fdir, fi, fj, oi);
        double4 Lc = zero;
        bool hit = false;
        if (valid && fc_on && jg == 0) {
            if (kind != 3) hit = fc_get(key, slot, Lc);
            if (!hit && g_on) {
                hit = gfc_get(key, Lc);
                if (hit && kind != 3) fc_put(key, slot, Lc);
            }
        }
        // the group's four face lanes are gbase + f * NGPP
        const unsigned long long hm = __ballot(hit);
        const unsigned long long lead = (1ull | (1ull << NGPP) | (1ull << (2 * NGPP)) | (1ull << (3 * NGPP))) << gbase;
        const bool allhit = (hm & lead) == lead;
        const bool fhit = (hm >> (gbase + f * NGPP)) & 1ull;
        double Lf[4] = {Lc.x, Lc.y, Lc.z, Lc.w};
        // a speculative job only copies what the device-wide cache holds into
        // the LDS one; with no device-wide cache it computes the faces itself
        if (valid && !allhit && !(kind == 2 && g_on)) {
            for (int qq = gl; qq < BOX * BOX; qq += G)
                L.nb[cs][qq] = load_stage<GEN>(a.Uin, a.B, i - HB + qq % BOX, j - HB + qq / BOX, xmin_top);
            wave_lds_sync();
            const LdsAcc lacc{&L.nb[cs][0], BOX};
            double F[4] = {0.0, 0.0, 0.0, 0.0};
            if (!fhit && jg < NGP) {
                double va1 = 0.0, va2 = 1.0, vb1 = 0.0, vb2 = 1.0;
                const int recon = MORD == 3 ? a.recon : kReconGP;
                if (recon == kReconCC) {
                    vnn_at(a.vnn, NX, NY, i + oai, j + oaj, va1, va2);
                    vnn_at(a.vnn, NX, NY, i + obi, j + obj, vb1, vb2);
                }
                face_gp_flux_dyn<MORD, FLUX>(T, lacc, ord, fa, fb, fdir, jg, HB + oai, HB + oaj, HB + obi, HB + obj, F,
                                             recon, va1, va2, vb1, vb2);
            }
            // quadrature over the face's Gauss points, j = 1..ngp from 0
            double S[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int jj = 0; jj < NGP; ++jj) {
                const int src = gbase + f * NGPP + jj;
                const double w = T->gw[jj];
#pragma unroll
                for (int k = 0; k < 4; ++k) S[k] = S[k] + __shfl(F[k], src) * w;
            }
            if (!fhit) {
#pragma unroll
                for (int k = 0; k < 4; ++k) Lf[k] = S[k];
                if (jg == 0 && fc_on) {
                    if (kind == 3) gfc_put(key, make_double4(S[0], S[1], S[2], S[3]));
                    else fc_put(key, slot, make_double4(S[0], S[1], S[2], S[3]));
                }
            }
        }
        double fyb[4], fxl[4], fxr[4], fyt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            fyb[k] = __shfl(Lf[k], gbase);
            fxl[k] = __shfl(Lf[k], gbase + NGPP);
            fxr[k] = __shfl(Lf[k], gbase + 2 * NGPP);
            fyt[k] = __shfl(Lf[k], gbase + 3 * NGPP);
        }
        bool dec = false;
        if (kind == 1 && gl == 0) {
            double cand[4];
            apply_fluxes(c0, fyb, fxl, fxr, fyt, dtdx, dtdy, cand);
            if (me > 1) dec = decide_with_flags(P, FiveAcc{c0, cl, cr, cb, ct, HB}, HB, HB, cand, fl);
            write_stage_output<GEN, kCmbRuntime>(a.C, a.Uout, d, GEN && a.C.band && in_band(i, j, NX, NY), un, c0,
                                                 cand, dtg);
            if (dec) {
                M::store(&a.gpo[d], pack_order(lower_order(me, a.cascade), me, q));
                __hip_atomic_fetch_add(&L.fail[q % 3], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        // a failing cell's 5-point neighbourhood enters list q + 1, one
        // neighbour per lane (lanes 0..4 of the group)
        if (__shfl(dec ? 1 : 0, gbase) && gl < 5) {
            const int ni = i + (gl == 1 ? -1 : gl == 3 ? 1 : 0), nj = j + (gl == 0 ? -1 : gl == 4 ? 1 : 0);
            if (ni >= 0 && ni < NX && nj >= 0 && nj < NY) {
                const int dn = nj * NX + ni;
                if (lds_set_insert(L.hash[(q + 1) % 3], kSoloHashBits, dn))
                    L.list[(q + 1) & 1][atomicAdd(&L.cnt[(q + 1) % 3], 1)] = dn;
            }
        }
    }

    // The other workgroups of the persistent kernel while workgroup 0 runs
    // the passes from p alone: they fill the device-wide cache ahead of it.
    // Workgroup 0 publishes each pass's list (the first 16 cells); a failure
    // chain moves by one cell per pass, so the helpers take the cells within
    // gfc.radius (Chebyshev distance) of the listed ones (each cell once per
    // stage, claimed in `done`) and compute their four face fluxes at every
    // order.  They stop when workgroup 0 has ended this run of passes.
    __device__ __forceinline__ void helper(int p)
    {
        if (!g_on || gridDim.x < 2) return;
        constexpr int AG = __HIP_MEMORY_SCOPE_AGENT;
        constexpr int K = (MORD + 1) / 2;  // orders 1, 3, ..., Mord
        const int hr = a.gfc.radius < 0 ? 0 : a.gfc.radius > 8 ? 8 : a.gfc.radius;
        const int WIN = 2 * hr + 1, NWIN = WIN * WIN;
        const int mcap = 3 * kSoloHash / NWIN;  // listed cells whose windows fit the solo tables
        const int tid = threadIdx.x;
        const int nh = (int)gridDim.x - 1, hw = (int)blockIdx.x - 1;
        const int slots = (int)blockDim.x / G;
        int *cells = &L.hash[0][0];  // the solo tables are unused outside workgroup 0
        int last = -1;
        for (;;) {
            if (tid == 0) {
                // workgroup 0 ended this run of passes, or (ordinary launch) a
                // grid barrier stalled and workgroup 0 may have left
                L.hstop = load_relaxed(&a.cnt[kCntStop]) == p ||
                          (a.sync == kMoodOrd && load_relaxed(&a.cnt[kCntRel]) >= kBarStalled);
                L.hseq = __hip_atomic_load(&a.gfc.beacon[0], __ATOMIC_RELAXED, AG);
                const int m = __hip_atomic_load(&a.gfc.beacon[1], __ATOMIC_RELAXED, AG);
                L.hm = m < 0 ? 0 : m > 16 ? 16 : m;
                L.hn = 0;
            }
            if (tid < 16) L.hlist[tid] = __hip_atomic_load(&a.gfc.beacon[2 + tid], __ATOMIC_RELAXED, AG);
            __syncthreads();
            const bool stop = L.hstop, fresh = L.hseq != last;
            const int m = L.hm < mcap ? L.hm : mcap;
            last = L.hseq;
            if (stop) break;
            if (fresh) {
                // this workgroup's share of the window cells, each taken once per stage
                for (int k = tid; k < m * NWIN; k += blockDim.x) {
                    if (k % nh != hw) continue;
                    const int c0 = L.hlist[k / NWIN], o = k % NWIN;
                    if (c0 < 0 || c0 >= NX * NY) continue;
                    const int ci = c0 % NX + o % WIN - hr, cj = c0 / NX + o / WIN - hr;
                    if (ci < 0 || ci >= NX || cj < 0 || cj >= NY) continue;
                    const int c = cj * NX + ci;
                    if (__hip_atomic_exchange(&a.gfc.done[c], gst, __ATOMIC_RELAXED, AG) != gst)
                        cells[atomicAdd(&L.hn, 1)] = c;
                }
            }
            __syncthreads();
            const int nj = fresh ? L.hn * K : 0;
            for (int b = 0; b < nj; b += slots) {
                const int e = b + cs;
                int kind = 0, c = 0, var = 0;
                if (e < nj) {
                    kind = 3;
                    c = cells[e / K];
                    var = e % K;
                }
                job(p, kind, c, var);
            }
            if (!fresh) __builtin_amdgcn_s_sleep(8);
            __syncthreads();  // every lane is done with the round's LDS fields
        }
    }

    // Short lists (the long tail of a failure chain): workgroup 0 runs the
    // passes from p on alone while the list has 1..solo_max cells, with the
    // lists, counts and deduplication in LDS and workgroup-scope memory
    // operations, which cost far less than grid barriers and device-scope
    // round trips.  Same passes, same order of effects: the results are
    // those of the grid-wide path.  Its first kEWaves wavefronts run the
    // listed cells, pass after pass, separated by their own barrier; the
    // other wavefronts run the speculative jobs of the latest pass those
    // publish, filling the LDS face-flux cache for the passes that follow,
    // and never hold a pass up.  The current list goes back to global
    // memory with its count; the stage's counters record the passes run.
    // Returns the first pass not run.
    // Barrier of the kEWaves wavefronts that run the listed cells of
    // workgroup 0's passes (the speculative wavefronts do not wait for the
    // passes): an arrival count and a generation word in LDS, with
    // workgroup-scope release and acquire around it, as a workgroup barrier
    // has.  The wait is bounded; false means it timed out.
    static constexpr int kEWaves = 4;
    static_assert(kMoodThreads >= 64 * kEWaves, "the listed cells' wavefronts are part of the workgroup");
    __device__ __forceinline__ bool e_barrier(int &gen)
    {
        constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        int ok = 1;
        if (lane == 0) {
            const int g = gen;
            if (__hip_atomic_fetch_add(&L.ebar_cnt, 1, __ATOMIC_RELAXED, WG) == kEWaves - 1) {
                __hip_atomic_store(&L.ebar_cnt, 0, __ATOMIC_RELAXED, WG);
                __hip_atomic_store(&L.ebar_gen, g + 1, __ATOMIC_RELEASE, WG);
            } else {
                for (int spin = 0; __hip_atomic_load(&L.ebar_gen, __ATOMIC_ACQUIRE, WG) == g; ++spin) {
                    if (spin > (1 << 22)) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        ++gen;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        return __shfl(ok, 0) != 0;
    }

    __device__ __forceinline__ int solo(int p, int nS)
    {
        constexpr int WG = __HIP_MEMORY_SCOPE_WORKGROUP;
        const int tid = threadIdx.x, wave = tid >> 6;
        constexpr int ES = kEWaves * PW;  // job slots of the listed cells' wavefronts
        for (int k = tid; k < nS; k += blockDim.x) L.list[p & 1][k] = a.slist[p & 1][k];
        for (int k = tid; k < 3 * kSoloHash; k += blockDim.x) L.hash[k / kSoloHash][k % kSoloHash] = -1;
        if (tid < 3) {
            L.cnt[tid] = tid == p % 3 ? nS : 0;
            L.fail[tid] = 0;
        }
        if (tid == 0) {
            L.ebar_cnt = 0;
            L.ebar_gen = 0;
            L.spass = -1;
            L.sdone = 0;
        }
        __syncthreads();
        int second = 0, det = 0;
        if (wave < kEWaves) {
            int q = p, gen = 0;
            for (;;) {
                // the previous pass's order words and outputs are visible to
                // these wavefronts, its list and counts are final
                if (q > p && !e_barrier(gen)) break;
                const int m = L.cnt[q % 3];
                if (tid == 0 && q > p) {
                    det += L.fail[(q - 1) % 3];
                    L.fail[(q - 1) % 3] = 0;  // next written in pass q + 2
                }
                if (m == 0 || m > a.solo_max) break;
                second += m;
                // the tables of pass q + 1 (no reader or writer in this pass)
                if (tid == 0) L.cnt[(q + 2) % 3] = 0;
                for (int k = tid; k < kSoloHash; k += 64 * kEWaves) L.hash[(q + 2) % 3][k] = -1;
                if (g_on && tid < 16) {
                    // the pass and its list for the other workgroups (read as hints only)
                    constexpr int AG = __HIP_MEMORY_SCOPE_AGENT;
                    if (tid < m) __hip_atomic_store(&a.gfc.beacon[2 + tid], L.list[q & 1][tid], __ATOMIC_RELAXED, AG);
                    if (tid == 0) {
                        __hip_atomic_store(&a.gfc.beacon[1], m, __ATOMIC_RELAXED, AG);
                        __hip_atomic_store(&a.gfc.beacon[0], q, __ATOMIC_RELAXED, AG);
                    }
                }
                if (tid == 0) __hip_atomic_store(&L.spass, q, __ATOMIC_RELEASE, WG);
                const int niter = (m + ES - 1) / ES;
                for (int it = 0; it < niter; ++it) {
                    const int e = it * ES + cs;
                    job(q, e < m ? 1 : 0, e < m ? L.list[q & 1][e] : 0, 0);
                }
                ++q;
            }
            if (tid == 0) {
                L.qend = q;
                __hip_atomic_store(&L.sdone, 1, __ATOMIC_RELEASE, WG);
            }
        } else if (fc_on) {
            // speculative wavefronts: the jobs of the latest pass published,
            // until the passes end (they never hold a pass up)
            int last = -1;
            for (;;) {
                if (__hip_atomic_load(&L.sdone, __ATOMIC_ACQUIRE, WG)) break;
                const int cur = __hip_atomic_load(&L.spass, __ATOMIC_ACQUIRE, WG);
                if (cur < 0 || cur == last) {
                    __builtin_amdgcn_s_sleep(2);
                    continue;
                }
                last = cur;
                int mm = L.cnt[cur % 3];
                mm = mm < 0 ? 0 : mm > 16 ? 16 : mm;
                int d = 0, var = 0;
                const bool has = spec_job(cur, mm, cs - ES, d, var);
                job(cur, has ? 2 : 0, d, var);
            }
        }
        __syncthreads();
        // hand the current list back to the grid-wide path
        const int q = L.qend;
        const int m = L.cnt[q % 3];
        for (int k = tid; k < m; k += blockDim.x) a.slist[q & 1][k] = L.list[q & 1][k];
        if (tid == 0) {
            __hip_atomic_store(&a.cnt[kCntStop], p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a.cnt[kCntS0 + q % 3] = m;
            a.cnt[kCntS0 + (q + 1) % 3] = 0;
            a.cnt[kCntNext] = q;
            a.cnt[kCntSecond] += second;
            a.cnt[kCntDet] += det;
        }
        return q;
    }

    // Grid-wide pass p over the nS cells of slist[p & 1]: all lanes [0, gsz).
    __device__ __forceinline__ void grid_pass(int p, int nS, int d0, int gtid, int gsz)
    {
        if (gtid == 0) {
            a.cnt[kCntS0 + (p + 2) % 3] = 0;  // last read at the start of pass p - 1
            a.cnt[kCntSecond] += nS;
            a.cnt[kCntNext] = p + 1;
        }
        run<__HIP_MEMORY_SCOPE_AGENT>(p, nS, d0, gtid, gsz, a.slist[p & 1], &a.cnt[kCntDet],
                                             GridMark{a.mark, a.slist[(p + 1) & 1], &a.cnt[kCntS0 + (p + 1) % 3],
                                               stamp_base + p + 1});
    }
};

// The stage's statistics once its MOOD loop ended at pass p (the first pass
// with an empty list): passes, detections and recomputed cells, and the mark
// stamp base of the next stage.
__device__ __forceinline__ void mood_finish(const MoodArgs &a, int p)
{
    *a.stamp = *a.stamp + p + 2;
    const long long det = (long long)a.cnt[kCntFail0] + a.cnt[kCntDet];
    DevStats *st = a.stats;
    st->passes[a.stage_idx] = p;
    if (p > st->max_passes) st->max_passes = p;
    st->detected_step += det;
    st->detected_total += det;
    st->second_pass_cells += a.cnt[kCntSecond];
    // every pass but the last lowered at least one cell
    if (a.nn_passes && p > 0) {
        st->nn_passes_step += p - 1;
        st->nn_passes_total += p - 1;
    }
}

// Rendezvous of an ordinary persistent launch (thread 0 of each workgroup),
// before any workgroup touches the stage's data.  Returns 0 when every
// workgroup of the grid has arrived: all are running, and a running
// workgroup stays resident until it returns, so grid barriers cannot wait on
// one that has not started.  When the others have not all arrived after
// rv_spins polls (the device is shared with other work), this workgroup
// sets kArriveAbort and returns 2: it takes every pass alone; the others
// return 1 and leave.  Arrivals are atomic adds on one word (no retries when
// all workgroups arrive at once); the abort is a compare-and-swap that sets
// the bit only while the count is below gridDim.x, and every reader checks
// the bit first, so all workgroups agree: either all see gridDim.x arrivals
// without the bit, or all but the one that set it see the bit.
__device__ __forceinline__ int mood_rendezvous(const MoodArgs &a)
{
    constexpr int AG = __HIP_MEMORY_SCOPE_AGENT;
    int *w = &a.cnt[kCntArrive];
    const int n = (int)gridDim.x;
    int v = __hip_atomic_fetch_add(w, 1, __ATOMIC_RELAXED, AG);
    if (v & kArriveAbort) return 1;
    if (a.rv_alone) return v == 0 ? 2 : 1;  // the first to arrive takes the passes alone
    for (int s = 0;; ++s) {
        v = load_relaxed(w);
        if (v & kArriveAbort) return 1;
        if (v >= n) return 0;
        if (s >= a.rv_spins) {
            if (__hip_atomic_compare_exchange_strong(w, &v, v | kArriveAbort, __ATOMIC_RELAXED, __ATOMIC_RELAXED, AG))
                return 2;
            continue;  // the word changed: an arrival or another abort
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Grid barrier of an ordinary persistent launch after its rendezvous: each
// workgroup adds one arrival to kCntBar (gridDim.x per barrier); the last to
// arrive publishes the barrier's number in kCntRel, on a cache line of its
// own, which the others poll (so that their polls do not queue behind the
// arrivals).  Every wavefront makes its own agent-scope release (the XCD's
// L2 written back after its stores) before the arrival and its own acquire
// (its caches invalidated) after the release: with the write-back made by
// one thread for the workgroup, after every wavefront's s_waitcnt, a 2048^2
// run with the workgroup-0 face-flux cache off read stale counters and
// stalled, so a wavefront's completed stores are not taken to be covered by
// another wavefront's write-back.  A wait that
// reaches its bound (a.bar_spins polls) counts a stall and lifts kCntRel to
// kBarStalled (an atomic maximum, which no later release lowers); a
// workgroup that finds it there returns false and leaves the kernel without
// starting another pass (the helpers of workgroup 0's passes stop on it too),
// so no wait blocks on a workgroup that has left, every pass a workgroup
// starts follows a completed barrier, and the host reports the stall.  ok is
// a workgroup-shared flag.
__device__ __forceinline__ bool ord_grid_sync(const MoodArgs &a, int &nbar, int *ok)
{
    constexpr int AG = __HIP_MEMORY_SCOPE_AGENT;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        int *rel = &a.cnt[kCntRel];
        const int gen = nbar + 1;
        const int old = __hip_atomic_fetch_add(&a.cnt[kCntBar], 1, __ATOMIC_RELAXED, AG);
        int v;
        if (old == gen * (int)gridDim.x - 1) {
            // the last arrival: every workgroup's writes are published
            v = __hip_atomic_fetch_max(rel, gen, __ATOMIC_RELEASE, AG);
            if (v < kBarStalled) v = gen;
        } else {
            for (int s = 0; (v = load_relaxed(rel)) < gen; ++s) {
                if (s >= a.bar_spins) {
                    atomicAdd(&a.stats->mood_stall, 1);
                    __hip_atomic_fetch_max(rel, kBarStalled, __ATOMIC_RELAXED, AG);
                    v = kBarStalled;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        *ok = v < kBarStalled;
    }
    ++nbar;
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return *ok != 0;
}

// The same between the passes of a workgroup that runs them alone (the
// passes read counters with agent-scope loads, and list entries that other
// wavefronts wrote, through the CU's caches).
__device__ __forceinline__ void alone_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// MOOD correction passes of one RK stage as a persistent kernel with one
// grid barrier per pass, from the first pass the wide launches before it
// left (kCntPass + 1; pass 1 when none ran); the loop ends on the device when
// a pass finds no cell to recompute.  a.sync says how it was launched:
// cooperatively (cooperative-groups barrier, the default), or as an ordinary
// launch that first meets at mood_rendezvous and then uses ord_grid_sync (it
// avoids the idle time the runtime puts around every cooperative launch, and
// the one-process resource those need).
template <int MORD, int NGP, int FLUX, bool GEN>
__global__ __launch_bounds__(kMoodThreads) void k_mood(MoodArgs a)
{
    using MP = MoodPass<MORD, NGP, FLUX, GEN>;
    __shared__ typename MP::Lds lds;
    __shared__ int s_rv, s_ok;
    MP mp(a, lds);
    const bool ord = a.sync == kMoodOrd;
    bool alone = false;
    if (ord) {
        if (threadIdx.x == 0) s_rv = mood_rendezvous(a);
        __syncthreads();
        if (s_rv == 1) return;
        alone = s_rv == 2;
        if (alone) {
            mp.g_on = false;  // no helper workgroups
            if (threadIdx.x == 0) atomicAdd(&a.stats->mood_alone, 1);
        }
    }
    mp.init_lds();
    cg::grid_group grid = cg::this_grid();
    const int ncell = a.B.NX * a.B.NY;
    const int gtid = alone ? (int)threadIdx.x : (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int gsz = alone ? (int)blockDim.x : (int)(gridDim.x * blockDim.x);
    int nbar = 0;
    // false: an ordinary launch's barrier stalled, and the workgroup leaves
    const auto sync = [&]() -> bool {
        if (alone) alone_sync();
        else if (ord) return ord_grid_sync(a, nbar, &s_ok);
        else grid.sync();
        return true;
    };
    int p = load_relaxed(&a.cnt[kCntPass]) + 1;
    for (;;) {
        // the pass's count and this lane's first list entry, in one round trip
        const int nS = load_relaxed(&a.cnt[kCntS0 + p % 3]);
        const int d0 = gtid / MP::G < ncell ? a.slist[p & 1][gtid / MP::G] : 0;
        if (nS == 0) break;
        if (gtid == 0 && p <= 2) atomicMax(&a.stats->long_list, nS);
        if (nS <= a.solo_max) {
            // workgroup 0 runs the short-list passes; the others fill the
            // device-wide face-flux cache meanwhile, then wait here
            if (alone || blockIdx.x == 0) mp.solo(p, nS);
            else mp.helper(p);
            if (!sync()) return;
            p = load_relaxed(&a.cnt[kCntNext]);
            continue;
        }
        mp.grid_pass(p, nS, d0, gtid, gsz);
        if (!sync()) return;
        ++p;
    }
    if (gtid == 0) mood_finish(a, p);
}

// The same passes without a persistent launch: each launch runs the current
// pass (all workgroups) or one run of workgroup-0 passes, and its last
// workgroup to finish publishes the next pass.  A launch whose pass has no
// cell returns at once, so the host launches in batches and checks
// kCntPending between them.
template <int MORD, int NGP, int FLUX, bool GEN>
__global__ __launch_bounds__(kMoodThreads) void k_mood_pass(MoodArgs a)
{
    using MP = MoodPass<MORD, NGP, FLUX, GEN>;
    __shared__ typename MP::Lds lds;
    __shared__ int s_p, s_n;
    __shared__ bool s_last;
    if (threadIdx.x == 0) {
        s_p = load_relaxed(&a.cnt[kCntPass]) + 1;
        s_n = load_relaxed(&a.cnt[kCntS0 + s_p % 3]);
    }
    __syncthreads();
    const int p = s_p, nS = s_n;
    if (nS == 0) return;
    MP mp(a, lds);
    mp.g_on = false;  // no other workgroups fill the device-wide cache here
    mp.init_lds();
    const int ncell = a.B.NX * a.B.NY;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    if (nS <= a.solo_max) {
        if (blockIdx.x == 0) mp.solo(p, nS);
    } else {
        const int d0 = gtid / MP::G < ncell ? a.slist[p & 1][gtid / MP::G] : 0;
        mp.grid_pass(p, nS, d0, gtid, gridDim.x * blockDim.x);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&a.cnt[kCntDone], 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last || threadIdx.x != 0) return;
    __threadfence();
    const int q = load_relaxed(&a.cnt[kCntNext]);
    a.cnt[kCntPass] = q - 1;
    a.cnt[kCntPending] = load_relaxed(&a.cnt[kCntS0 + q % 3]) > 0;
    a.cnt[kCntDone] = 0;
}

// A grid-wide pass over a long list in an ordinary launch of many small
// workgroups.  The persistent kernel's grid is kept small, because its
// per-pass grid barrier costs more with every workgroup, so it covers a list
// of thousands of cells in many trips, one after another.  Launched before
// it: the launch runs the current pass (kCntPass + 1) only when that pass
// lists more than wide_min cells, and then publishes the next pass as the
// per-pass launches do; otherwise every workgroup returns at once (all read
// the same counters, which no workgroup of the launch changes).
template <int MORD, int NGP, int FLUX, bool GEN>
__global__ __launch_bounds__(kWideThreads) void k_mood_wide(MoodArgs a)
{
    using MP = MoodPass<MORD, NGP, FLUX, GEN, true>;
    __shared__ typename MP::Lds lds;
    __shared__ int s_p, s_n;
    __shared__ bool s_last;
    if (threadIdx.x == 0) {
        s_p = load_relaxed(&a.cnt[kCntPass]) + 1;
        s_n = load_relaxed(&a.cnt[kCntS0 + s_p % 3]);
    }
    __syncthreads();
    const int p = s_p, nS = s_n;
    if (blockIdx.x == 0 && threadIdx.x == 0 && p <= 2) atomicMax(&a.stats->long_list, nS);
    if (nS <= a.wide_min) return;
    // a workgroup whose first trip has no cell has none in any trip
    if ((int)blockIdx.x * MP::CPB < nS) {
        MP mp(a, lds);
        mp.init_lds();
        const int ncell = a.B.NX * a.B.NY;
        const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
        const int d0 = gtid / MP::G < ncell ? a.slist[p & 1][gtid / MP::G] : 0;
        mp.grid_pass(p, nS, d0, gtid, gridDim.x * blockDim.x);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&a.cnt[kCntDone], 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last || threadIdx.x != 0) return;
    __threadfence();
    const int q = load_relaxed(&a.cnt[kCntNext]);
    a.cnt[kCntPass] = q - 1;
    a.cnt[kCntPending] = load_relaxed(&a.cnt[kCntS0 + q % 3]) > 0;
    a.cnt[kCntDone] = 0;
}

// Launch entry points of one (Mord, ngp) combination, for every numFlux and
// both boundary classes (gen: formula boundaries or a source term); the
// copy-boundary stage kernels are also specialised on the SSP-RK3 combination
// kinds.  Each combination is compiled in a translation unit of its own.
// The stage input through the ghost rule, for reconstructions at ghost cells.
struct BndAcc {
    Bnd B;
    const double4 *__restrict__ U;
    double xmin_top;
    __device__ __forceinline__ double4 operator()(int i, int j) const
    {
        return bnd_value(B, DevSrc{U, B.NX}, i, j, xmin_top);
    }
};

// Ghost ring after a step of a method that ends its MOOD loop without
// DETECTION (FOG, GP_SE, Unlim_POL; mod_FE.f90:160-162).  No Boundary_C
// refills the last stage's Uout: its ghost cells keep that stage's
// Boundary_C(Uin) values, and the face loop (mod_FE.f90:75-145, faces between
// cells -1..lf and -1..nf in 0-based indices, face states from the
// reconstruction of cells -1..lf, -1..nf) has updated those of the first
// layer, in the order of an interior cell's updates: + bottom, + left,
// - right, - top face.  The integrator's combination with the step-start state
// follows, as for the interior.  MORD 1: FOG's cell values as face states.
template <int MORD, int NGP, int FLUX>
__global__ __launch_bounds__(128) void k_ring_face(RingArgs a)
{
    const int NX = a.B.NX, NY = a.B.NY;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ring_size(NX, NY)) return;
    int i, j;
    ring_cell(r, NX, NY, i, j);
    const double *ts = a.tstate;
    const double t0 = ts[4], dt = ts[1], dtdx = ts[2], dtdy = ts[3];
    const double tb = a.integ == kRingRK3 ? t0 + 0.5 * dt : t0;
    const BndAcc acc{a.B, a.Bin, dmr_xmin_top(tb, a.B.sqrt3)};
    const double4 u = acc(i, j);
    double c[4] = {u.x, u.y, u.z, u.w};
    if (i >= -1 && i <= NX && j >= -1 && j <= NY) {
        // face (p, q) -> (p+1, q) or (p, q) -> (p, q+1) exists for p in [-1, NX-1], q in [-1, NY-1]
        auto exists = [&](int p, int q) { return p >= -1 && p <= NX - 1 && q >= -1 && q <= NY - 1; };
        auto flux = [&](int dir, int ai, int aj, int bi, int bj, double L[4]) {
            if (MORD == 1) {
                if (dir == 0) fog_flux<0, FLUX>(a.T, acc(ai, aj), acc(bi, bj), L);
                else fog_flux<1, FLUX>(a.T, acc(ai, aj), acc(bi, bj), L);
            } else if (dir == 0) {
                face_flux<MORD, NGP, 0, FLUX>(a.T, acc, ai, aj, bi, bj, L);
            } else {
                face_flux<MORD, NGP, 1, FLUX>(a.T, acc, ai, aj, bi, bj, L);
            }
        };
        double L[4];
        if (exists(i, j - 1)) {
            flux(1, i, j - 1, i, j, L);
            for (int k = 0; k < 4; ++k) c[k] = c[k] + dtdy * L[k];
        }
        if (exists(i - 1, j)) {
            flux(0, i - 1, j, i, j, L);
            for (int k = 0; k < 4; ++k) c[k] = c[k] + dtdx * L[k];
        }
        if (exists(i, j)) {
            flux(0, i, j, i + 1, j, L);
            for (int k = 0; k < 4; ++k) c[k] = c[k] - dtdx * L[k];
            flux(1, i, j, i, j + 1, L);
            for (int k = 0; k < 4; ++k) c[k] = c[k] - dtdy * L[k];
        }
    }
    if (a.gravity) gravity_source(c, dt * kGravity);
    if (a.integ == kRingFE) {
        a.ring[r] = make_double4(c[0], c[1], c[2], c[3]);
        return;
    }
    // the step-start state's ghosts as Boundary_C left them in the first stage
    const double4 un = bnd_value(a.B, DevSrc{a.A, NX}, i, j, dmr_xmin_top(t0, a.B.sqrt3));
    if (a.integ == kRingRK2)
        a.ring[r] = make_double4(0.5 * (un.x + c[0]), 0.5 * (un.y + c[1]), 0.5 * (un.z + c[2]), 0.5 * (un.w + c[3]));
    else
        a.ring[r] = make_double4((1.0 / 3) * un.x + (2.0 / 3) * c[0], (1.0 / 3) * un.y + (2.0 / 3) * c[1],
                                 (1.0 / 3) * un.z + (2.0 / 3) * c[2], (1.0 / 3) * un.w + (2.0 / 3) * c[3]);
}

#define GPM_INSTANTIATE_COMBO(M, G)                                                                          \
    template <int F, int C>                                                                                   \
    static void launch_stage_fc##M##G(bool gen, const StageArgs &a, int nblocks, hipStream_t s)              \
    {                                                                                                         \
        if (gen)                                                                                              \
            hipLaunchKernelGGL((k_stage<M, G, F, true, kCmbRuntime>), dim3(nblocks), dim3(kStageThreads), 0, s, a); \
        else hipLaunchKernelGGL((k_stage<M, G, F, false, C>), dim3(nblocks), dim3(kStageThreads), 0, s, a);  \
    }                                                                                                         \
    template <int F>                                                                                          \
    static void launch_stage_f##M##G(bool gen, const StageArgs &a, int nblocks, hipStream_t s)               \
    {                                                                                                         \
        switch (gen ? kCmbRuntime : a.C.kind) {                                                               \
        case kCmbCopy: launch_stage_fc##M##G<F, kCmbCopy>(gen, a, nblocks, s); break;                          \
        case kCmbRK3b: launch_stage_fc##M##G<F, kCmbRK3b>(gen, a, nblocks, s); break;                          \
        case kCmbRK3c: launch_stage_fc##M##G<F, kCmbRK3c>(gen, a, nblocks, s); break;                          \
        default: launch_stage_fc##M##G<F, kCmbRuntime>(gen, a, nblocks, s);                                   \
        }                                                                                                     \
    }                                                                                                         \
    template <int F>                                                                                          \
    static const void *mood_kernel_f##M##G(bool gen, bool coop)                                               \
    {                                                                                                         \
        if (coop)                                                                                             \
            return gen ? reinterpret_cast<const void *>(&k_mood<M, G, F, true>)                              \
                       : reinterpret_cast<const void *>(&k_mood<M, G, F, false>);                            \
        return gen ? reinterpret_cast<const void *>(&k_mood_pass<M, G, F, true>)                             \
                   : reinterpret_cast<const void *>(&k_mood_pass<M, G, F, false>);                           \
    }                                                                                                         \
    template <int F>                                                                                          \
    static const void *mood_wide_f##M##G(bool gen)                                                            \
    {                                                                                                         \
        return gen ? reinterpret_cast<const void *>(&k_mood_wide<M, G, F, true>)                             \
                   : reinterpret_cast<const void *>(&k_mood_wide<M, G, F, false>);                           \
    }                                                                                                         \
    template <>                                                                                               \
    const void *mood_wide_kernel_mg<M, G>(int flux, bool gen)                                                 \
    {                                                                                                         \
        if (flux == kHLL) return mood_wide_f##M##G<kHLL>(gen);                                               \
        if (flux == kLLF) return mood_wide_f##M##G<kLLF>(gen);                                               \
        if (flux == kUnsplit) return mood_wide_f##M##G<kUnsplit>(gen);                                       \
        return mood_wide_f##M##G<kHLLC>(gen);                                                                \
    }                                                                                                         \
    template <>                                                                                               \
    hipError_t launch_stage_mg<M, G>(int flux, bool gen, const StageArgs &a, int nblocks, hipStream_t s)     \
    {                                                                                                         \
        if (flux == kHLL) launch_stage_f##M##G<kHLL>(gen, a, nblocks, s);                                    \
        else if (flux == kLLF) launch_stage_f##M##G<kLLF>(gen, a, nblocks, s);                               \
        else if (flux == kUnsplit) launch_stage_f##M##G<kUnsplit>(gen, a, nblocks, s);                       \
        else launch_stage_f##M##G<kHLLC>(gen, a, nblocks, s);                                                \
        return hipGetLastError();                                                                             \
    }                                                                                                         \
    template <>                                                                                               \
    const void *mood_kernel_mg<M, G>(int flux, bool gen, bool coop)                                           \
    {                                                                                                         \
        if (flux == kHLL) return mood_kernel_f##M##G<kHLL>(gen, coop);                                       \
        if (flux == kLLF) return mood_kernel_f##M##G<kLLF>(gen, coop);                                       \
        if (flux == kUnsplit) return mood_kernel_f##M##G<kUnsplit>(gen, coop);                               \
        return mood_kernel_f##M##G<kHLLC>(gen, coop);                                                        \
    }                                                                                                         \
    template <>                                                                                               \
    hipError_t launch_ring_face_mg<M, G>(int flux, const RingArgs &a, hipStream_t s)                         \
    {                                                                                                         \
        const dim3 grid((ring_size(a.B.NX, a.B.NY) + 127) / 128), block(128);                                 \
        if (flux == kHLL) hipLaunchKernelGGL((k_ring_face<M, G, kHLL>), grid, block, 0, s, a);               \
        else if (flux == kLLF) hipLaunchKernelGGL((k_ring_face<M, G, kLLF>), grid, block, 0, s, a);          \
        else if (flux == kUnsplit) hipLaunchKernelGGL((k_ring_face<M, G, kUnsplit>), grid, block, 0, s, a);  \
        else hipLaunchKernelGGL((k_ring_face<M, G, kHLLC>), grid, block, 0, s, a);                           \
        return hipGetLastError();                                                                             \
    }

}  // namespace gpm
