// HIP kernels of the GP-MOOD SSP-RK3 step for gfx950 (MI355X).
//
// Per RK stage (Forward_Euler, mod_FE.f90:20-190):
//   k_stage  first MOOD pass over the whole grid.  A workgroup stages a tile of
//            the stage input with its halo in LDS, computes every face flux of
//            the tile once (GP reconstruction at the Gauss points + HLLC +
//            quadrature), applies the four face updates per cell in the
//            reference's order, runs the detection, writes the RK-combined
//            output and appends troubled cells to a work list (wavefront
//            ballot + one atomic per wavefront).
//   k_mood   the remaining MOOD passes, as one cooperative persistent kernel:
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

// Accessor of a stage input in HBM with the boundary rule applied.
struct GlobAcc {
    const double4 *__restrict__ U;
    int N, bc;
    __device__ __forceinline__ double4 operator()(int i, int j) const { return load_bc(U, N, bc, i, j); }
};

template <int STAGE>
__device__ __forceinline__ double4 rk_combine(const double4 *__restrict__ Un, size_t c, const double cand[4])
{
    if (STAGE == 1) return make_double4(cand[0], cand[1], cand[2], cand[3]);
    const double4 u = Un[c];
    if (STAGE == 2) {  // mod_SSP_RK3.f90:28: U2 = 0.75*Uin + 0.25*U2
        return make_double4(0.75 * u.x + 0.25 * cand[0], 0.75 * u.y + 0.25 * cand[1], 0.75 * u.z + 0.25 * cand[2],
                            0.75 * u.w + 0.25 * cand[3]);
    }
    // mod_SSP_RK3.f90:33: Uout = (1./3)*Uin + (2./3)*Uout
    const double a = 1.0 / 3, b = 2.0 / 3;
    return make_double4(a * u.x + b * cand[0], a * u.y + b * cand[1], a * u.z + b * cand[2], a * u.w + b * cand[3]);
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

template <int MORD, int NGP, int STAGE>
__global__ __launch_bounds__(kStageThreads) void k_stage(StageArgs a)
{
    constexpr int H = kHalo<MORD>;
    constexpr int PW = kTX + 2 * H, PH = kTY + 2 * H;
    __shared__ double4 sU[PW * PH];
    __shared__ double4 sFx[kTY * (kTX + 1)];
    __shared__ double4 sFy[(kTY + 1) * kTX];

    const int N = a.N;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile < 0 || tile >= a.ntiles) return;
    const int x0 = (tile % a.tiles_x) * kTX, y0 = (tile / a.tiles_x) * kTY;
    const int tid = threadIdx.x;

    for (int idx = tid; idx < PW * PH; idx += kStageThreads) {
        const int ti = idx % PW, tj = idx / PW;
        sU[idx] = load_bc(a.Uin, N, a.bc, x0 - H + ti, y0 - H + tj);
    }
    __syncthreads();

    const DevTables *__restrict__ T = a.T;
    const LdsAcc acc{sU, PW};
    constexpr int NFX = (kTX + 1) * kTY, NFY = kTX * (kTY + 1);
    for (int f = tid; f < NFX + NFY; f += kStageThreads) {
        double L[4];
        if (f < NFX) {
            const int fy = f / (kTX + 1), fx = f % (kTX + 1);
            face_flux<MORD, NGP, 0>(T, acc, fx - 1 + H, fy + H, fx + H, fy + H, L);
            sFx[f] = make_double4(L[0], L[1], L[2], L[3]);
        } else {
            const int g = f - NFX;
            const int gy = g / kTX, gx = g % kTX;
            face_flux<MORD, NGP, 1>(T, acc, gx + H, gy - 1 + H, gx + H, gy + H, L);
            sFy[g] = make_double4(L[0], L[1], L[2], L[3]);
        }
    }
    __syncthreads();

    const double dtdx = a.tstate[2], dtdy = a.tstate[3];
    const MoodParams P{a.dmp, a.u2, a.u2tol, a.dx, a.dy};
    for (int c = tid; c < kTX * kTY; c += kStageThreads) {
        const int cx = c % kTX, cy = c / kTX;
        const int gi = x0 + cx, gj = y0 + cy;
        const bool inside = gi < N && gj < N;
        bool dec = false;
        if (inside) {
            const double4 u = sU[(cy + H) * PW + cx + H];
            const double4 b = sFy[cy * kTX + cx], l = sFx[cy * (kTX + 1) + cx];
            const double4 r = sFx[cy * (kTX + 1) + cx + 1], t = sFy[(cy + 1) * kTX + cx];
            const double fyb[4] = {b.x, b.y, b.z, b.w}, fxl[4] = {l.x, l.y, l.z, l.w};
            const double fxr[4] = {r.x, r.y, r.z, r.w}, fyt[4] = {t.x, t.y, t.z, t.w};
            double cand[4];
            apply_fluxes(u, fyb, fxl, fxr, fyt, dtdx, dtdy, cand);
            if (MORD > 1) dec = decide(P, T, acc, cx + H, cy + H, cand);
            const size_t gc = (size_t)gj * N + gi;
            a.Uout[gc] = rk_combine<STAGE>(a.Un, gc, cand);
            a.gpo[gc] = (int8_t)MORD;
        }
        wave_append(dec, gj * N + gi, a.flist, a.cnt_f);
    }
}

// Recompute one cell with the current orders (all four faces) and decide.
template <int MORD, int NGP, int STAGE>
__device__ __forceinline__ bool recompute_cell(const MoodArgs &a, int c)
{
    const int N = a.N;
    const int i = c % N, j = c / N;
    const GlobAcc acc{a.Uin, N, a.bc};
    const int8_t *__restrict__ g = a.gpo;
    const int me = g[c];
    auto ord_of = [&](int ii, int jj) -> int {
        return (ii < 0 || ii >= N || jj < 0 || jj >= N) ? MORD : (int)g[(size_t)jj * N + ii];
    };
    double fyb[4], fxl[4], fxr[4], fyt[4];
    {
        const int o = min(ord_of(i, j - 1), me);
        face_flux_dyn<MORD, NGP, 1>(a.T, acc, o, i, j - 1, i, j, fyb);
    }
    {
        const int o = min(ord_of(i - 1, j), me);
        face_flux_dyn<MORD, NGP, 0>(a.T, acc, o, i - 1, j, i, j, fxl);
    }
    {
        const int o = min(ord_of(i + 1, j), me);
        face_flux_dyn<MORD, NGP, 0>(a.T, acc, o, i, j, i + 1, j, fxr);
    }
    {
        const int o = min(ord_of(i, j + 1), me);
        face_flux_dyn<MORD, NGP, 1>(a.T, acc, o, i, j, i, j + 1, fyt);
    }
    double cand[4];
    apply_fluxes(a.Uin[c], fyb, fxl, fxr, fyt, a.tstate[2], a.tstate[3], cand);
    bool dec = false;
    if (me > 1) {
        const MoodParams P{a.dmp, a.u2, a.u2tol, a.dx, a.dy};
        dec = decide(P, a.T, acc, i, j, cand);
    }
    a.Uout[c] = rk_combine<STAGE>(a.Un, c, cand);
    return dec;
}

__device__ __forceinline__ int load_relaxed(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int MORD, int NGP, int STAGE>
__global__ __launch_bounds__(kMoodThreads) void k_mood(MoodArgs a)
{
    cg::grid_group grid = cg::this_grid();
    const int N = a.N;
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gsz = gridDim.x * blockDim.x;
    const int stamp0 = *a.stamp;
    int p = 0;
    long long second = 0;
    for (;;) {
        const int nF = load_relaxed(&a.cnt_f[p]);
        if (nF == 0) break;
        if (p + 1 >= a.pmax) {
            if (gtid == 0) atomicOr(a.err, 1);
            break;
        }
        const int *__restrict__ F = (p & 1) ? a.flist_b : a.flist_a;
        int *__restrict__ Fn = (p & 1) ? a.flist_a : a.flist_b;
        const int stamp = stamp0 + p;
        // phase A: cascade and neighbourhood marking (mod_subroutine_MOOD.f90:148-185)
        for (int t = gtid; t < nF; t += gsz) {
            const int c = F[t];
            const int i = c % N, j = c / N;
            a.gpo[c] = (int8_t)lower_order(a.gpo[c], a.cascade);
            const int ni[5] = {i, i - 1, i, i + 1, i}, nj[5] = {j - 1, j, j, j, j + 1};
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                if (ni[k] < 0 || ni[k] >= N || nj[k] < 0 || nj[k] >= N) continue;
                const int d = nj[k] * N + ni[k];
                if (atomicExch(&a.mark[d], stamp) != stamp) {
                    const int slot = atomicAdd(&a.cnt_s[p], 1);
                    a.slist[slot] = d;
                }
            }
        }
        grid.sync();
        // phase B: recompute the marked cells and detect
        const int nS = load_relaxed(&a.cnt_s[p]);
        second += nS;
        const int nloop = (nS + gsz - 1) / gsz * gsz;
        for (int t = gtid; t < nloop; t += gsz) {
            bool dec = false;
            int d = 0;
            if (t < nS) {
                d = a.slist[t];
                dec = recompute_cell<MORD, NGP, STAGE>(a, d);
            }
            wave_append(dec, d, Fn, &a.cnt_f[p + 1]);
        }
        grid.sync();
        ++p;
    }
    if (gtid == 0) {
        *a.stamp = stamp0 + p + 1;
        long long det = 0;
        for (int q = 0; q <= p; ++q) det += a.cnt_f[q];
        DevStats *st = a.stats;
        st->passes[a.stage_idx] = p + 1;
        if (p + 1 > st->max_passes) st->max_passes = p + 1;
        st->detected_step += det;
        st->detected_total += det;
        st->second_pass_cells += second;
    }
}

}  // namespace gpm
