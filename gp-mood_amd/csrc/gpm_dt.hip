// Setdt reduction kernel and per-step bookkeeping.
#include "gpm_kernels_impl.h"

namespace gpm {

constexpr int kDtThreads = 1024;

// Minimum over the workgroup (fmin2 at every level): wavefront shuffles, then
// one value per wavefront through LDS.  Valid on every lane of wavefront 0.
__device__ __forceinline__ double block_min(double m, double *red)
{
    for (int o = 32; o > 0; o >>= 1) m = fmin2(m, __shfl_xor(m, o));
    const int w = threadIdx.x / 64, nw = blockDim.x / 64;
    __syncthreads();  // red is reused across calls
    if (__lane_id() == 0) red[w] = m;
    __syncthreads();
    m = 1.0e30;
    if (w == 0) {
        if (__lane_id() < nw) m = red[__lane_id()];
        for (int o = 32; o > 0; o >>= 1) m = fmin2(m, __shfl_xor(m, o));
    }
    return m;
}

// vx = mx / rho, vy = my / rho, the pressure (its kinetic energy over rho) and
// cs = sqrt(gamma p / rho) as c2p_shared forms them for a Riemann state: the
// four divisions by the density share its reciprocal, the same bits as four
// ddiv calls.  The kernel is bound by these correctly rounded FP64 divisions
// (nine per cell, five after sharing), not by its 32 B per cell.
__device__ __forceinline__ double dt_of_cell(const double4 &u, double dx, double dy)
{
    const double uu[4] = {u.x, u.y, u.z, u.w};
    const Prim s = c2p_shared(uu);
    return dv(1.0, dv(1.0, dv(dx, fabs(s.vx) + s.c)) + dv(1.0, dv(dy, fabs(s.vy) + s.c)));
}

// Setdt: per-cell 1/(1/(dx/(|vx|+c)) + 1/(dy/(|vy|+c))), minimum over the
// grid and its ghost cells, times CFL, clipped to dtfinal (physics.f90:98-124).
// Ghost cells that copy interior cells (index-map boundaries) add nothing to
// the minimum; otherwise the ring holds them.  The last workgroup to finish
// reduces the per-workgroup minima, records the step-start time, advances t
// and zeroes the step counters.  Few large workgroups: every workgroup ends
// with a fence and an atomic on one counter, which serialise across the device.
__global__ __launch_bounds__(kDtThreads) void k_dt(DtArgs a)
{
    __shared__ double red[kDtThreads / 64];
    __shared__ bool last;
    const int tid = threadIdx.x;
    double m = 1.0e30;
    if (a.mode == kDtCompute) {
        // four coalesced cells per thread and trip, so that loads and division
        // chains of independent cells overlap
        const size_t ncell = (size_t)a.NX * a.NY;
        const size_t gs = (size_t)gridDim.x * blockDim.x;
        for (size_t c0 = (size_t)blockIdx.x * blockDim.x + tid; c0 < ncell; c0 += 4 * gs) {
            double4 u[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const size_t c = c0 + q * gs;
                u[q] = c < ncell ? a.U[c] : make_double4(1.0, 0.0, 0.0, 1.0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double v = dt_of_cell(u[q], a.dx, a.dy);
                if (c0 + q * gs < ncell) m = fmin2(m, v);
            }
        }
        if (a.ring) {
            const size_t nr = (size_t)ring_size(a.NX, a.NY);
            for (size_t r = (size_t)blockIdx.x * blockDim.x + tid; r < nr; r += gs)
                m = fmin2(m, dt_of_cell(a.ring[r], a.dx, a.dy));
        }
    }
    m = block_min(m, red);
    if (tid == 0) {
        a.block_min[blockIdx.x] = m;
        __threadfence();
        const unsigned prev = atomicAdd(a.done, 1u);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double mm = 1.0e30;
    for (int b = tid; b < (int)gridDim.x; b += blockDim.x)
        mm = fmin2(mm, __hip_atomic_load(&a.block_min[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    mm = block_min(mm, red);
    // a new step zeroes the MOOD pass counters of its three stages
    if (a.new_step)
        for (int k = tid; k < a.ncounters; k += blockDim.x) a.counters[k] = 0;
    if (tid == 0) {
        double *ts = a.tstate;
        const double t = ts[0];
        double dt;
        if (a.mode == kDtCompute) {
            dt = mm * a.cfl;
            const double dtfinal = a.use_tmax ? a.tmax - t : a.dtfinal;
            if (dt > dtfinal) dt = dtfinal;
        } else {
            dt = a.dt_given;
        }
        ts[1] = dt;
        ts[2] = dv(dt, a.dx);
        ts[3] = dv(dt, a.dy);
        ts[4] = t;
        if (a.advance_t) ts[0] = t + dt;
        if (a.dt_log) a.dt_log[a.step_idx] = dt;
        if (a.new_step) {
            a.stats->detected_step = 0;
            a.stats->apriori_step = 0;
            a.stats->nn_passes_step = 0;
            a.stats->steps += 1;
        }
        *a.done = 0u;
    }
}


// The step's FP64 division on n operand pairs, for tests against the host's
// IEEE division: ddiv itself (shared = false) or the shared-reciprocal form
// qdiv(a, recip(b)) the Riemann solvers use.
__global__ void k_ddiv(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ q, long long n,
                       bool shared)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        q[i] = shared ? qdiv(a[i], recip(b[i])) : ddiv(a[i], b[i]);
}

hipError_t launch_ddiv(const double *a, const double *b, double *q, long long n, bool shared, hipStream_t s)
{
    hipLaunchKernelGGL(k_ddiv, dim3(1024), dim3(256), 0, s, a, b, q, n, shared);
    return hipGetLastError();
}

// Statistics of a stage whose MOOD passes ran as per-pass launches.
__global__ void k_mood_finish(MoodArgs a) { mood_finish(a, a.cnt[kCntPass] + 1); }

hipError_t launch_mood_finish(const MoodArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_mood_finish, dim3(1), dim3(1), 0, s, a);
    return hipGetLastError();
}

// Start of a run on buffers that keep an earlier run's contents (the
// allocation-free entry point): marks 0, orders Mord, counters, statistics
// and the completion counter 0, the mark stamp 1, time t.
__global__ void k_run_reset(int *__restrict__ mark, order_t *__restrict__ gpo, size_t n, order_t g0,
                            int *__restrict__ counters, int ncounters, DevStats *__restrict__ stats,
                            double *__restrict__ tstate, double t, unsigned *__restrict__ done, int *__restrict__ stamp)
{
    const size_t gs = (size_t)gridDim.x * blockDim.x;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gs) {
        mark[c] = 0;
        gpo[c] = g0;
    }
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < ncounters; k += blockDim.x) counters[k] = 0;
        if (threadIdx.x == 0) {
            *stats = DevStats{};
            for (int k = 0; k < 8; ++k) tstate[k] = 0.0;
            tstate[0] = t;
            tstate[4] = t;
            *done = 0u;
            *stamp = 1;
        }
    }
}

hipError_t launch_run_reset(int *mark, order_t *gpo, size_t n, order_t g0, int *counters, int ncounters,
                            DevStats *stats, double *tstate, double t, unsigned *done, int *stamp, hipStream_t s)
{
    hipLaunchKernelGGL(k_run_reset, dim3(512), dim3(256), 0, s, mark, gpo, n, g0, counters, ncounters, stats, tstate, t,
                       done, stamp);
    return hipGetLastError();
}

// Order words -> plain CellGPO bytes.
__global__ void k_unpack_orders(const order_t *__restrict__ g, int8_t *__restrict__ out, size_t n)
{
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        out[c] = (int8_t)order_of_word(g[c]);
}

// Order words from a priori orders (one per cell, version 0).
__global__ void k_orders_from_pred(const int8_t *__restrict__ pred, order_t *__restrict__ g, size_t n)
{
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x)
        g[c] = pack_order(pred[c], pred[c], 0);
}

hipError_t launch_orders_from_pred(const int8_t *pred, order_t *g, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_orders_from_pred, dim3(1024), dim3(256), 0, s, pred, g, n);
    return hipGetLastError();
}

hipError_t launch_unpack_orders(const order_t *g, int8_t *out, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_unpack_orders, dim3(1024), dim3(256), 0, s, g, out, n);
    return hipGetLastError();
}

// Ghost ring of a state (RingArgs): one thread per ghost cell.
__global__ void k_ring(RingArgs a)
{
    const int NX = a.B.NX, NY = a.B.NY;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ring_size(NX, NY)) return;
    int i, j;
    ring_cell(r, NX, NY, i, j);
    const double *ts = a.tstate;
    const double sq = a.B.sqrt3;
    if (a.mode == kRingFill) {
        a.ring[r] = bnd_value(a.B, DevSrc{a.A, NX}, i, j, dmr_xmin_top(ts[0], sq));
        return;
    }
    const double t0 = ts[4], dt = ts[1];
    if (a.mode == kRingRK4) {  // mod_SSP_RK4.f90:34 over the ghost cells
        const double xm = dmr_xmin_top(t0, sq);
        const double4 q2 = bnd_value(a.B, DevSrc{a.X2, NX}, i, j, xm), q3 = bnd_value(a.B, DevSrc{a.X3, NX}, i, j, xm);
        const double4 q4 = bnd_value(a.B, DevSrc{a.X4, NX}, i, j, xm);
        const double4 e3 = bnd_value(a.B, DevSrc{a.band3, NX}, i, j, xm), e4 = bnd_value(a.B, DevSrc{a.band, NX}, i, j, xm);
        double g3[4] = {e3.x, e3.y, e3.z, e3.w}, g4[4] = {e4.x, e4.y, e4.z, e4.w};
        if (a.gravity) {
            gravity_source(g3, dt * kGravity);
            gravity_source(g4, dt * kGravity);
        }
        const double x2[4] = {q2.x, q2.y, q2.z, q2.w}, x3[4] = {q3.x, q3.y, q3.z, q3.w}, x4[4] = {q4.x, q4.y, q4.z, q4.w};
        double o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double fu3 = g3[k] - x3[k], fu4 = g4[k] - x4[k];
            o[k] = kRK4f2 * x2[k] + kRK4f3 * x3[k] + kRK4ff3 * fu3 + kRK4f4 * x4[k] + kRK4ff4 * fu4;
        }
        a.ring[r] = make_double4(o[0], o[1], o[2], o[3]);
        return;
    }
    // the last stage's Forward_Euler output: Boundary_C at its t_rk, then the source
    const double tb = a.mode == kRingRK3 ? t0 + 0.5 * dt : t0;
    const double4 b = bnd_value(a.B, DevSrc{a.band, NX}, i, j, dmr_xmin_top(tb, sq));
    double c[4] = {b.x, b.y, b.z, b.w};
    if (a.gravity) gravity_source(c, dt * kGravity);
    if (a.mode == kRingFE) {
        a.ring[r] = make_double4(c[0], c[1], c[2], c[3]);
        return;
    }
    // the step-start state's ghosts as Boundary_C left them in the first stage
    const double4 u = bnd_value(a.B, DevSrc{a.A, NX}, i, j, dmr_xmin_top(t0, sq));
    if (a.mode == kRingRK2)
        a.ring[r] = make_double4(0.5 * (u.x + c[0]), 0.5 * (u.y + c[1]), 0.5 * (u.z + c[2]), 0.5 * (u.w + c[3]));
    else
        a.ring[r] = make_double4((1.0 / 3) * u.x + (2.0 / 3) * c[0], (1.0 / 3) * u.y + (2.0 / 3) * c[1],
                                 (1.0 / 3) * u.z + (2.0 / 3) * c[2], (1.0 / 3) * u.w + (2.0 / 3) * c[3]);
}

hipError_t launch_ring(const RingArgs &a, hipStream_t s)
{
    const int n = ring_size(a.B.NX, a.B.NY);
    hipLaunchKernelGGL(k_ring, dim3((n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Ghosted state [NY+8][NX+8] of the interior U: ghost cells from the ring, or
// by the boundary rule (copy boundaries).
__global__ void k_pack_ghosted(const double4 *__restrict__ U, const double4 *__restrict__ ring, Bnd B,
                               double4 *__restrict__ out)
{
    const int NX = B.NX, NY = B.NY, W = NX + 8;
    const long long n = (long long)W * (NY + 8);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % W) - 4, j = (int)(c / W) - 4;
        double4 v;
        if (i >= 0 && i < NX && j >= 0 && j < NY) v = U[(size_t)j * NX + i];
        else if (ring) v = ring[ring_index(i, j, NX, NY)];
        else v = bnd_value(B, DevSrc{U, NX}, i, j, 0.0);
        out[c] = v;
    }
}

// Interior (and, with ring, the ghost cells) of a ghosted state.
__global__ void k_unpack_ghosted(const double4 *__restrict__ in, double4 *__restrict__ U, double4 *__restrict__ ring,
                                 int NX, int NY)
{
    const int W = NX + 8;
    const long long n = (long long)W * (NY + 8);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % W) - 4, j = (int)(c / W) - 4;
        if (i >= 0 && i < NX && j >= 0 && j < NY) {
            if (U) U[(size_t)j * NX + i] = in[c];
        } else if (ring) {
            ring[ring_index(i, j, NX, NY)] = in[c];
        }
    }
}

// CellGPO(lb:le, nb:ne) as int32 from the order words; ghost cells (and every
// cell when `all_mord`) hold Mord.
__global__ void k_cellgpo_ghosted(const order_t *__restrict__ g, int32_t *__restrict__ out, int NX, int NY, int mord,
                                  int all_mord)
{
    const int W = NX + 8;
    const long long n = (long long)W * (NY + 8);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % W) - 4, j = (int)(c / W) - 4;
        const bool in = i >= 0 && i < NX && j >= 0 && j < NY;
        out[c] = (!in || all_mord) ? mord : order_of_word(g[(size_t)j * NX + i]);
    }
}

hipError_t launch_pack_ghosted(const double4 *U, const double4 *ring, const Bnd &B, double4 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_ghosted, dim3(2048), dim3(256), 0, s, U, ring, B, out);
    return hipGetLastError();
}

hipError_t launch_unpack_ghosted(const double4 *in, double4 *U, double4 *ring, int NX, int NY, hipStream_t s)
{
    hipLaunchKernelGGL(k_unpack_ghosted, dim3(2048), dim3(256), 0, s, in, U, ring, NX, NY);
    return hipGetLastError();
}

hipError_t launch_cellgpo_ghosted(const order_t *g, int32_t *out, int NX, int NY, int mord, int all_mord,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(k_cellgpo_ghosted, dim3(2048), dim3(256), 0, s, g, out, NX, NY, mord, all_mord);
    return hipGetLastError();
}

hipError_t launch_dt(const DtArgs &a, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_dt, dim3(nblocks), dim3(kDtThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace gpm
