// The reference's own a priori network (method NN_GP_MOOD / NN_GP_MOOD_CC
// with a network in the reference's text format): compute_CellGPO_with_NN,
// mod_NN.f90:14-58, with format_input (mod_append_NN_dataset.f90:81-126) and
// forward (mod_NN.f90:165-181), in FP32 with the reference's rounding:
//   * the R+1 stencil in real(4), variable fastest, then -666 up to L inputs;
//   * matmul: each row summed in real(4) from 0 in column order, then + bias
//     (unfused multiply and add; the Fortran runtime's order);
//   * sigmoid one/(one + exp(-x)) with the expf of nn_expf below;
//   * r(1) > r(2) lowers the cell to order 1, anything else gives order 3.
#include "gpm_kernels.h"

#include <cstring>

namespace gpm {

// exp of a float by the table method of Szabolcs Nagy's expf (ARM
// optimized-routines, the expf of glibc 2.28 and later), as it evaluates on an
// x86-64 CPU with FMA: x*N/ln2 = k + r, 2^(k/N) from a 32-entry table, 2^(r/N)
// by a cubic, all in double with the fused operations of that build, then one
// rounding to float.  Checked equal to the host C library's expf for every
// float in [-0x1.9fe368p6, 0x1.62e42ep6] (the whole range where neither
// overflow nor underflow applies).
//
// Upstream: ARM optimized-routines, math/expf.c with its table in
// math/exp2f_data.c (licence: MIT OR Apache-2.0 WITH LLVM-exception), which
// glibc carries as sysdeps/ieee754/flt-32/e_expf.c and e_exp2f_data.c
// (LGPL-2.1-or-later).  This is a restatement of the published method: the
// table entries below are the values 2^(i/32) rounded to double, and the
// cubic's coefficients are the method's; no upstream source is copied.
//
// kExp2Tab[i] = bits of 2^(i/32) (rounded to double) minus i << 47, so that
// adding k << 47 puts k/32's integer part into the exponent.
constexpr unsigned long long kExp2Tab[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
};

__host__ __device__ __forceinline__ unsigned f32_bits(float x)
{
    unsigned u;
    std::memcpy(&u, &x, 4);
    return u;
}
__host__ __device__ __forceinline__ unsigned long long f64_bits(double x)
{
    unsigned long long u;
    std::memcpy(&u, &x, 8);
    return u;
}
__host__ __device__ __forceinline__ double f64_of(unsigned long long u)
{
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

__host__ __device__ __forceinline__ float nn_expf(float x)
{
    constexpr double kN = 32.0;
    constexpr double kInvLn2N = 0x1.71547652b82fep+0 * kN;
    constexpr double kShift = 0x1.8p+52;
    // 2^(r/N) ~ C0 r^3 + C1 r^2 + C2 r + 1 (the exp2 cubic, scaled for r in units of 1/N)
    constexpr double kC0 = 0x1.c6af84b912394p-5 / kN / kN / kN;
    constexpr double kC1 = 0x1.ebfce50fac4f3p-3 / kN / kN;
    constexpr double kC2 = 0x1.62e42ff0c52d6p-1 / kN;
    const unsigned ux = f32_bits(x);
    const unsigned abstop = (ux >> 20) & 0x7ffu;
    if (abstop >= 0x42bu) {  // |x| >= 88, inf or nan
        if (ux == 0xff800000u) return 0.0f;
        if (abstop >= 0x7f8u) return x + x;
        if (x > 0x1.62e42ep6f) return __builtin_huge_valf();
        if (x < -0x1.9fe368p6f) return 0.0f;
    }
    const double xd = (double)x;
    double kd = __builtin_fma(kInvLn2N, xd, kShift);
    const unsigned long long ki = f64_bits(kd);
    kd -= kShift;
    const double r = __builtin_fma(kInvLn2N, xd, -kd);
    const unsigned long long t = kExp2Tab[ki & 31] + (ki << 47);
    const double s = f64_of(t);
    const double z = __builtin_fma(kC0, r, kC1);
    const double r2 = r * r;
    double y = __builtin_fma(kC2, r, 1.0);
    y = __builtin_fma(z, r2, y);
    y = y * s;
    return (float)y;
}

__device__ __forceinline__ float nn_sigmoid(float x) { return 1.0f / (1.0f + nn_expf(-x)); }

// ixiy_sp1 (GP_init.f90:273-343): the R+1 sphere stencil in the reference's order
__constant__ int kSp1X[25] = {0, -1, 0, 1, 0, -1, -1, 1, 1, 0, -2, 0, 2, -3, -2, -1, 0, 1, 2, 3, 2, 1, 0, -1, -2};
__constant__ int kSp1Y[25] = {0, 0, 1, 0, -1, -1, 1, 1, -1, -2, 0, 2, 0, 0, 1, 2, 3, 2, 1, 0, -1, -2, -3, -2, -1};

// One thread per interior cell, one wavefront per workgroup.  The weights
// are zero-padded to the compile-time widths KP (inputs) and kRefNnMaxLen
// (hidden units) in the device copy: row i of layer 1 is w0p[i*KP .. i*KP+KP-1]
// and of layer 2 w1p[i*HP .. ].  Each output sums its row from 0 in input
// order, so the real terms come first and in the reference's order; the
// padding terms add +0*0, which leaves every sum unchanged except that a
// -0 sum becomes +0, and no output depends on the sign of a zero sum (the sum
// only enters sigmoid(sum + bias), r(1) > r(2) and exp).  The inputs live in
// registers; the weight rows stream through the scalar cache; the first hidden
// layer goes through LDS once to be held in registers for layer 2; layer 3 is
// summed as layer 2 produces its outputs, in the same order.
// LX, LEN > 0: a build for one network shape (L = LX, length = LEN), whose
// loops stop at the real inputs and hidden units and skip the padding.
template <int S, int KP, int LX, int LEN>
__global__ __launch_bounds__(kRefNnThreads) void k_refnn(RefNnArgs a)
{
    constexpr int HP = kRefNnMaxLen;
    static_assert(LX <= KP && LEN <= HP, "shape within the padded widths");
    constexpr int KU = LX > 0 ? LX : KP;    // inputs summed per hidden unit
    constexpr int HU = LEN > 0 ? LEN : HP;  // hidden units evaluated
    __shared__ float hs[HP * kRefNnThreads];
    const int NX = a.B.NX, NY = a.B.NY;
    const int lane = threadIdx.x;
    const long long cell = (long long)blockIdx.x * kRefNnThreads + lane;
    const bool valid = cell < (long long)NX * NY;
    const int i = valid ? (int)(cell % NX) : 0, j = valid ? (int)(cell / NX) : 0;
    const int L = LX > 0 ? LX : a.L, len = LEN > 0 ? LEN : a.len;
    // format_input: stencil values, -666 up to L, then the padding's zeros;
    // per-variable max/min skipping NaN
    float x[KP];
    float mx[4], mn[4];
    bool any[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) { mx[v] = -__builtin_inff(); mn[v] = __builtin_inff(); any[v] = false; }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double4 u = valid ? load_copy(a.U, NX, NY, a.B.bc, i + kSp1X[s], j + kSp1Y[s])
                                : make_double4(1.0, 0.0, 0.0, 1.0);
        const float f[4] = {(float)u.x, (float)u.y, (float)u.z, (float)u.w};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            x[4 * s + v] = f[v];
            if (!__builtin_isnan(f[v])) {
                any[v] = true;
                if (f[v] > mx[v]) mx[v] = f[v];
                if (f[v] < mn[v]) mn[v] = f[v];
            }
        }
    }
#pragma unroll
    for (int k = 4 * S; k < KP; ++k) x[k] = k < L ? -666.0f : 0.0f;
    bool cst = true;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const float d = any[v] ? mx[v] - mn[v] : __builtin_nanf("");
        if (!((double)d < 1e-10)) cst = false;
    }
    const float *__restrict__ w0 = a.w;                      // [len][KP]
    const float *__restrict__ b0 = w0 + (size_t)len * KP;     // [len]
    const float *__restrict__ w1 = b0 + len;                  // [len][HP]
    const float *__restrict__ b1 = w1 + (size_t)len * HP;     // [len]
    const float *__restrict__ w2 = b1 + len;                  // weight2(2, length), Fortran layout
    const float *__restrict__ b2 = w2 + 2 * len;
    // every lane runs the network (cst lanes discard it): uniform loops
    for (int q = 0; q < HU; ++q) {
        float h = 0.0f;
        if (LEN > 0 || q < len) {
            const float *__restrict__ w = w0 + (size_t)q * KP;
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < KU; ++k) acc = acc + w[k] * x[k];
            h = nn_sigmoid(acc + b0[q]);
        }
        hs[q * kRefNnThreads + lane] = h;
    }
    float h0[HU];
#pragma unroll
    for (int k = 0; k < HU; ++k) h0[k] = hs[k * kRefNnThreads + lane];
    float r0 = 0.0f, r1 = 0.0f;
    for (int q = 0; q < len; ++q) {
        const float *__restrict__ w = w1 + (size_t)q * HP;
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < HU; ++k) acc = acc + w[k] * h0[k];
        const float h = nn_sigmoid(acc + b1[q]);
        r0 = r0 + w2[2 * q] * h;
        r1 = r1 + w2[2 * q + 1] * h;
    }
    r0 = r0 + b2[0];
    r1 = r1 + b2[1];
    if (a.softmax) {  // softmax, mod_NN.f90:133-147
        r0 = nn_expf(r0);
        r1 = nn_expf(r1);
        const float den = r0 + r1;
        r0 = r0 / den;
        r1 = r1 / den;
    }
    bool low = false;
    if (valid) {
        int8_t ord = 3, pri = (int8_t)a.mord;
        if (cst) {
            pri = 3;
            r0 = 0.0f;
            r1 = 1.0f;
        } else if (r0 > r1) {
            ord = 1;
            pri = 1;
            low = true;
        }
        a.pred[cell] = ord;
        a.priori[cell] = pri;
        if (a.rowv && i == NX - 1) {
            a.rowv[2 * j] = r0;
            a.rowv[2 * j + 1] = r1;
        }
    }
    const unsigned long long m = __ballot(low);
    if (lane == 0 && m) {
        const long long c = (long long)__popcll(m);
        atomicAdd(reinterpret_cast<unsigned long long *>(&a.stats->apriori_step), (unsigned long long)c);
        atomicAdd(reinterpret_cast<unsigned long long *>(&a.stats->detected_step), (unsigned long long)c);
        atomicAdd(reinterpret_cast<unsigned long long *>(&a.stats->detected_total), (unsigned long long)c);
        atomicAdd(reinterpret_cast<unsigned long long *>(&a.stats->a_priori_lowered), (unsigned long long)c);
    }
}

int refnn_input_pad(int L) { return L <= 64 ? 64 : 128; }

// 64-bit mix of (seed, step, cell) to a uniform double in [0, 1)
__device__ __forceinline__ double dump_uniform(unsigned long long seed, long long step, long long cell)
{
    unsigned long long z = seed ^ (0x9E3779B97F4A7C15ull * (unsigned long long)(step + 1)) ^
                           (0xBF58476D1CE4E5B9ull * (unsigned long long)(cell + 1));
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

// append_to_NN_datatset (mod_append_NN_dataset.f90:11-79) after the first RK
// stage: for each interior cell, the stage input's R+1 stencil through
// format_input (the network's inputs) and the stage's final order as the label,
// (0, 1) for order 3 and (1, 0) otherwise.  Constant stencils are skipped;
// order-3 samples are kept with probability freq_R0 = NR0 / (NR0 + NR1) (0.5
// before any sample), from the counts at the start of this stage.  Kept samples
// go to row atomicAdd(count) while it is below capacity.
template <int S>
__global__ __launch_bounds__(256) void k_refnn_dump(RefDumpArgs a)
{
    const int NX = a.B.NX, NY = a.B.NY;
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (long long)NX * NY) return;
    const int i = (int)(cell % NX), j = (int)(cell / NX);
    const double xt = dmr_xmin_top(a.tstate[4], a.B.sqrt3);
    float x[4 * S];
    bool cst = true;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        float mx = -__builtin_inff(), mn = __builtin_inff();
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double4 u = bnd_value(a.B, DevSrc{a.U, NX}, i + kSp1X[s], j + kSp1Y[s], xt);
            const float f = (float)(v == 0 ? u.x : v == 1 ? u.y : v == 2 ? u.z : u.w);
            x[4 * s + v] = f;
            if (!__builtin_isnan(f)) {
                any = true;
                if (f > mx) mx = f;
                if (f < mn) mn = f;
            }
        }
        const float d = any ? mx - mn : __builtin_nanf("");
        if (!((double)d < 1e-10)) cst = false;
    }
    if (cst) return;
    const int ord = order_of_word(a.gpo[cell]);
    if (ord == 3) {
        const long long n0 = a.nr_start[0], n1 = a.nr_start[1];
        const double freq = (n0 + n1) > 0 ? (double)n0 / (double)(n0 + n1) : 0.5;
        if (dump_uniform(a.seed, a.step, cell) > freq) return;
    }
    const unsigned long long row = atomicAdd(a.count, 1ull);
    if (row >= (unsigned long long)a.capacity) return;
    float *dst = a.inputs + row * (size_t)a.L;
    for (int k = 0; k < a.L; ++k) dst[k] = k < 4 * S ? x[k] : -666.0f;
    a.labels[2 * row] = ord == 3 ? 0.0f : 1.0f;
    a.labels[2 * row + 1] = ord == 3 ? 1.0f : 0.0f;
    if (a.cells) a.cells[row] = (long long)cell;
    atomicAdd(reinterpret_cast<unsigned long long *>(&a.nr[ord == 3 ? 1 : 0]), 1ull);
}

__global__ void k_refnn_dump_begin(const long long *nr, long long *nr_start)
{
    if (threadIdx.x < 2) nr_start[threadIdx.x] = nr[threadIdx.x];
}

hipError_t launch_refnn_dump(const RefDumpArgs &a, int mord, hipStream_t s)
{
    hipLaunchKernelGGL(k_refnn_dump_begin, dim3(1), dim3(64), 0, s, a.nr, a.nr_start);
    const long long n = (long long)a.B.NX * a.B.NY;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (mord == 3) hipLaunchKernelGGL(k_refnn_dump<13>, grid, block, 0, s, a);
    else hipLaunchKernelGGL(k_refnn_dump<25>, grid, block, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_refnn(const RefNnArgs &a, hipStream_t s)
{
    const long long n = (long long)a.B.NX * a.B.NY;
    const dim3 grid((unsigned)((n + kRefNnThreads - 1) / kRefNnThreads)), block(kRefNnThreads);
    const int kp = refnn_input_pad(a.L);
    // the problem files' shapes (L 57, length 60; L 52 is the bare stencil) get builds of their own
    if (a.mord == 3 && a.L == 57 && a.len == 60) hipLaunchKernelGGL((k_refnn<13, 64, 57, 60>), grid, block, 0, s, a);
    else if (a.mord == 3 && a.L == 52 && a.len == 60) hipLaunchKernelGGL((k_refnn<13, 64, 52, 60>), grid, block, 0, s, a);
    else if (a.mord == 3 && kp == 64) hipLaunchKernelGGL((k_refnn<13, 64, 0, 0>), grid, block, 0, s, a);
    else if (a.mord == 3) hipLaunchKernelGGL((k_refnn<13, 128, 0, 0>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_refnn<25, 128, 0, 0>), grid, block, 0, s, a);
    return hipGetLastError();
}

// values_NN := (0, 1) over the whole array (mod_FE.f90:45-46)
__global__ void k_vnn_reset(float *vnn, long long plane)
{
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < plane; c += (long long)gridDim.x * blockDim.x) {
        vnn[c] = 0.0f;
        vnn[c + plane] = 1.0f;
    }
}

// values_NN(L, n, :) = the row's last values, rows in order; the flat index
// follows the Fortran storage of dimension(-1:lf+2, -1:nf+2, 2), so an L past
// the row lands in a later row, as in the reference.
__global__ void k_vnn_scatter(float *vnn, const float *rowv, int NX, int NY, int L)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const long long W = NX + 4, plane = W * (NY + 4);
    for (int n = 1; n <= NY; ++n) {
        const long long base = (long long)(L + 1) + (long long)(n + 1) * W;
        vnn[base] = rowv[2 * (n - 1)];
        vnn[base + plane] = rowv[2 * (n - 1) + 1];
    }
}

hipError_t launch_refnn_values(float *vnn, const float *rowv, int NX, int NY, int L, bool scatter, hipStream_t s)
{
    const long long plane = (long long)(NX + 4) * (NY + 4);
    hipLaunchKernelGGL(k_vnn_reset, dim3(256), dim3(256), 0, s, vnn, plane);
    if (scatter) hipLaunchKernelGGL(k_vnn_scatter, dim3(1), dim3(64), 0, s, vnn, rowv, NX, NY, L);
    return hipGetLastError();
}

__host__ __device__ __forceinline__ float f32_of(unsigned u)
{
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

__global__ void k_nn_expf_bits(unsigned first, long long n, float *out)
{
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x)
        out[c] = nn_expf(f32_of(first + (unsigned)c));
}

// The same function compiled for the host CPU.
void nn_expf_bits_host(uint32_t first, long long n, float *out)
{
    for (long long c = 0; c < n; ++c) out[c] = nn_expf(f32_of(first + (unsigned)c));
}

hipError_t launch_nn_expf_bits(uint32_t first, long long n, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_nn_expf_bits, dim3(4096), dim3(256), 0, s, first, n, out);
    return hipGetLastError();
}

}  // namespace gpm
