// Learned a priori order predictor: an MLP  in -> 128 -> 128 -> 64 -> C
// (ReLU, BF16 weights and activations, FP32 accumulation) with hand-written
// forward, backward and AdamW kernels on the gfx950 matrix cores, and the
// feature extraction shared by the solver's data dump and its fused inference.
//
// Matrix-core fragments (wavefront of 64 lanes, lane l), gfx950's
// v_mfma_f32_16x16x32_bf16 for every reduction of a multiple of 32:
//   A (16 x 32, rows x reduction): lane l holds A[l % 16][8*(l/16) + 0..7]
//   B (32 x 16, reduction x cols): lane l holds B[8*(l/16) + 0..7][l % 16]
//   D (16 x 16): lane l holds D[4*(l/16) + r][l % 16], r = 0..3
// and v_mfma_f32_16x16x16_bf16 (4 bf16 per lane, 4*(l/16) + 0..3) for the one
// reduction of 16 (the last layer's backward, over the 16 padded classes).
// Weights are stored [out][in] (PyTorch Linear order), so a B fragment of
// X W^T is consecutive bf16 of one weight row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "gpm_device.h"
#include "gpm_mlp.h"
#include "../../include/gpmood.h"

namespace gpm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                      // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x4 load_bf4(const uint16_t *p)
{
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xffff); r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffff); r[3] = (short)(v.y >> 16);
    return r;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// eight consecutive bf16 (16-byte aligned) as an x32 fragment
__device__ __forceinline__ bf16x8 load_bf8(const uint16_t *p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// One wavefront: Y[16 x NP] = act(X[16 x KP] W[NP x KP]^T + b).  X and W are
// bf16 with row strides ldx and KP; rows of W past n_valid and bias entries
// past n_valid read as zero.  Y goes to bf16 (ldy) or, with OUTF, to fp32.
template <int KP, int NP, bool RELU, bool OUTF>
__device__ __forceinline__ void dense16(const uint16_t *X, int ldx, const uint16_t *__restrict__ W,
                                        const float *__restrict__ b, void *Y, int ldy, int lane)
{
    f32x4 acc[NP / 16];
#pragma unroll
    for (int t = 0; t < NP / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, q = lane >> 4;
    static_assert(KP % 32 == 0, "the x32 fragments tile the reduction");
#pragma unroll 2
    for (int ks = 0; ks < KP / 32; ++ks) {
        const bf16x8 a = load_bf8(X + r16 * ldx + ks * 32 + 8 * q);
#pragma unroll
        for (int t = 0; t < NP / 16; ++t) {
            const bf16x8 w = load_bf8(W + (size_t)(t * 16 + r16) * KP + ks * 32 + 8 * q);
            acc[t] = mfma32(a, w, acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < NP / 16; ++t) {
        const int col = t * 16 + r16;
        const float bb = b ? b[col] : 0.f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * q + r;
            float v = acc[t][r] + bb;
            if (RELU) v = v > 0.f ? v : 0.f;
            if (OUTF) reinterpret_cast<float *>(Y)[row * ldy + col] = v;
            else reinterpret_cast<uint16_t *>(Y)[row * ldy + col] = f2bf(v);
        }
    }
}

// ---------------------------------------------------------------------------
// features: the (2R+1)^2 box stencil of a cell, 4 conserved variables each,
// normalised per variable by the stencil's max |value| (FP32), flattened with
// the variable fastest: f[4*(py*S + px) + k].

// Per-variable max |value| over the stencil points handled by this lane
// (points s = first, first+step, ...), and the normalised features of those
// points passed to out(index, value).
template <int S>
__device__ __forceinline__ void stencil_absmax(const double4 *__restrict__ U, int N, int bc, int i, int j, int first,
                                               int step, float mx[4])
{
    constexpr int R = S / 2;
    mx[0] = mx[1] = mx[2] = mx[3] = 0.f;
    for (int s = first; s < S * S; s += step) {
        const double4 u = load_bc(U, N, bc, i + s % S - R, j + s / S - R);
        mx[0] = fmaxf(mx[0], fabsf((float)u.x));
        mx[1] = fmaxf(mx[1], fabsf((float)u.y));
        mx[2] = fmaxf(mx[2], fabsf((float)u.z));
        mx[3] = fmaxf(mx[3], fabsf((float)u.w));
    }
}

template <int S, class Out>
__device__ __forceinline__ void stencil_write(const double4 *__restrict__ U, int N, int bc, int i, int j, int first,
                                              int step, const float mx[4], Out out)
{
    constexpr int R = S / 2;
    float inv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) inv[k] = mx[k] > 0.f ? 1.0f / mx[k] : 0.f;
    for (int s = first; s < S * S; s += step) {
        const double4 u = load_bc(U, N, bc, i + s % S - R, j + s / S - R);
        out(4 * s + 0, (float)u.x * inv[0]);
        out(4 * s + 1, (float)u.y * inv[1]);
        out(4 * s + 2, (float)u.z * inv[2]);
        out(4 * s + 3, (float)u.w * inv[3]);
    }
}

template <int S, class Out>
__device__ __forceinline__ void cell_features(const double4 *__restrict__ U, int N, int bc, int i, int j, Out out)
{
    float mx[4];
    stencil_absmax<S>(U, N, bc, i, j, 0, 1, mx);
    stencil_write<S>(U, N, bc, i, j, 0, 1, mx, out);
}

// ---------------------------------------------------------------------------
// training kernels (batch rows in blocks of 64 = 4 wavefronts x 16 rows)

// Layer 0 forward from FP32 features: X fp32 [B][in] -> H0 bf16 [B][128].
template <int KP>
__global__ __launch_bounds__(256) void k_fwd_in(const float *__restrict__ X, int in, int B,
                                                const uint16_t *__restrict__ W, const float *__restrict__ b,
                                                uint16_t *__restrict__ H)
{
    __shared__ uint16_t xs[4][16 * KP];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = blockIdx.x * 64 + wv * 16;
    for (int e = lane; e < 16 * KP; e += 64) {
        const int r = e / KP, c = e % KP;
        const int gr = row0 + r;
        xs[wv][e] = (gr < B && c < in) ? f2bf(X[(size_t)gr * in + c]) : (uint16_t)0;
    }
    __syncthreads();
    if (row0 >= B) return;
    // rows past B are computed from zeros and written to a scratch tail of H
    dense16<KP, 128, true, false>(xs[wv], KP, W, b, H + (size_t)row0 * 128, 128, lane);
}

// Hidden layers: H_in bf16 [B][KI] -> H_out bf16 [B][NO] (ReLU) or logits fp32.
template <int KI, int NO, bool RELU, bool OUTF>
__global__ __launch_bounds__(256) void k_fwd_hidden(const uint16_t *__restrict__ Hin, int B,
                                                    const uint16_t *__restrict__ W, const float *__restrict__ b,
                                                    void *Hout)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = blockIdx.x * 64 + wv * 16;
    if (row0 >= B) return;
    if (OUTF)
        dense16<KI, NO, RELU, true>(Hin + (size_t)row0 * KI, KI, W, b,
                                    reinterpret_cast<float *>(Hout) + (size_t)row0 * NO, NO, lane);
    else
        dense16<KI, NO, RELU, false>(Hin + (size_t)row0 * KI, KI, W, b,
                                     reinterpret_cast<uint16_t *>(Hout) + (size_t)row0 * NO, NO, lane);
}

// Softmax cross-entropy over C classes (logits fp32 [B][16]); writes the
// gradient of the mean loss (bf16 [B][16], zero past C), and per-block sums
// of the loss and of correct predictions.
__global__ __launch_bounds__(256) void k_softmax_ce(const float *__restrict__ logits, const int *__restrict__ y, int B,
                                                    int C, uint16_t *__restrict__ dlog, float *__restrict__ part)
{
    __shared__ float sl[256], sc[256];
    const int m = blockIdx.x * 256 + threadIdx.x;
    float loss = 0.f, corr = 0.f;
    if (m < B) {
        const float *z = logits + (size_t)m * 16;
        float zmax = z[0];
        int am = 0;
        for (int c = 1; c < C; ++c)
            if (z[c] > zmax) { zmax = z[c]; am = c; }
        float se = 0.f;
        for (int c = 0; c < C; ++c) se += expf(z[c] - zmax);
        const float lse = zmax + logf(se);
        const int lab = y[m];
        loss = lse - z[lab];
        corr = (am == lab) ? 1.f : 0.f;
        const float invB = 1.0f / (float)B;
        for (int c = 0; c < 16; ++c) {
            float g = 0.f;
            if (c < C) g = (expf(z[c] - lse) - (c == lab ? 1.f : 0.f)) * invB;
            dlog[(size_t)m * 16 + c] = f2bf(g);
        }
    }
    sl[threadIdx.x] = loss;
    sc[threadIdx.x] = corr;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) { sl[threadIdx.x] += sl[threadIdx.x + s]; sc[threadIdx.x] += sc[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = sl[0]; part[2 * blockIdx.x + 1] = sc[0]; }
}

// Backward through a layer's weights: dX[B][KO] = (dY[B][NI] W[NI][KO]) * (Hprev > 0),
// with WT = W^T stored [KO][NI].
template <int NI, int KO>
__global__ __launch_bounds__(256) void k_bwd_data(const uint16_t *__restrict__ dY, int B, const uint16_t *__restrict__ WT,
                                                  const uint16_t *__restrict__ Hprev, uint16_t *__restrict__ dX)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = blockIdx.x * 64 + wv * 16;
    if (row0 >= B) return;
    f32x4 acc[KO / 16];
#pragma unroll
    for (int t = 0; t < KO / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int r16 = lane & 15, q = lane >> 4;
    const uint16_t *Yr = dY + (size_t)row0 * NI;
    if constexpr (NI % 32 == 0) {
#pragma unroll
        for (int ks = 0; ks < NI / 32; ++ks) {
            const bf16x8 a = load_bf8(Yr + r16 * NI + ks * 32 + 8 * q);
#pragma unroll
            for (int t = 0; t < KO / 16; ++t)
                acc[t] = mfma32(a, load_bf8(WT + (size_t)(t * 16 + r16) * NI + ks * 32 + 8 * q), acc[t]);
        }
    } else {
#pragma unroll
        for (int ks = 0; ks < NI / 16; ++ks) {
            const bf16x4 a = load_bf4(Yr + r16 * NI + ks * 16 + 4 * q);
#pragma unroll
            for (int t = 0; t < KO / 16; ++t)
                acc[t] = mfma16(a, load_bf4(WT + (size_t)(t * 16 + r16) * NI + ks * 16 + 4 * q), acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < KO / 16; ++t) {
        const int col = t * 16 + r16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const size_t o = (size_t)(row0 + 4 * q + r) * KO + col;
            const float h = bf2f(Hprev[o]);
            dX[o] = f2bf(h > 0.f ? acc[t][r] : 0.f);
        }
    }
}

// Weight gradient partials: for a chunk of CH batch rows per workgroup,
// part[blk][n][k] = sum_m dY[m][n] X[m][k] and partb[blk][n] = sum_m dY[m][n].
// X is bf16 [B][KP] (or fp32 features [B][in] when XF, converted to bf16).
template <int NO, int KP, bool XF>
__global__ __launch_bounds__(256) void k_bwd_weight(const uint16_t *__restrict__ dY, const void *__restrict__ Xv, int in,
                                                    int B, float *__restrict__ part, float *__restrict__ partb)
{
    constexpr int CH = 64;
    __shared__ uint16_t yT[NO][CH + 4];
    __shared__ uint16_t xT[KP][CH + 4];
    const int m0 = blockIdx.x * CH;
    for (int e = threadIdx.x; e < CH * NO; e += 256) {
        const int m = e / NO, n = e % NO;
        yT[n][m] = (m0 + m < B) ? dY[(size_t)(m0 + m) * NO + n] : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < CH * KP; e += 256) {
        const int m = e / KP, k = e % KP;
        uint16_t v = 0;
        if (m0 + m < B) {
            if (XF) v = (k < in) ? f2bf(reinterpret_cast<const float *>(Xv)[(size_t)(m0 + m) * in + k]) : (uint16_t)0;
            else v = reinterpret_cast<const uint16_t *>(Xv)[(size_t)(m0 + m) * KP + k];
        }
        xT[k][m] = v;
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, q = lane >> 4;
    constexpr int TN = NO / 16, TK = KP / 16, NT = TN * TK;
    for (int tile = wv; tile < NT; tile += 4) {
        const int tn = tile / TK, tk = tile % TK;
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        static_assert(CH % 32 == 0, "the x32 fragments tile the batch chunk");
#pragma unroll
        for (int ms = 0; ms < CH / 32; ++ms) {
            uint16_t a[8], bb[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                a[c] = yT[tn * 16 + r16][ms * 32 + 8 * q + c];
                bb[c] = xT[tk * 16 + r16][ms * 32 + 8 * q + c];
            }
            acc = mfma32(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, bb), acc);
        }
        float *P = part + (size_t)blockIdx.x * NO * KP;
#pragma unroll
        for (int r = 0; r < 4; ++r) P[(size_t)(tn * 16 + 4 * q + r) * KP + tk * 16 + r16] = acc[r];
    }
    for (int n = threadIdx.x; n < NO; n += 256) {
        float sb = 0.f;
        for (int m = 0; m < CH; ++m) sb += bf2f(yT[n][m]);
        partb[(size_t)blockIdx.x * NO + n] = sb;
    }
}

// Deterministic sum of the per-workgroup partials: 64 outputs per workgroup,
// wavefront g sums partials g, g+4, g+8, ... of its output in order, and the
// four sums are combined in a fixed order.
__global__ __launch_bounds__(256) void k_reduce_parts(const float *__restrict__ part, int nblk, int len,
                                                      float *__restrict__ out)
{
    __shared__ float red[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + c;
    float s = 0.f;
    if (i < len)
        for (int b = g; b < nblk; b += 4) s += part[(size_t)b * len + i];
    red[g][c] = s;
    __syncthreads();
    if (g == 0 && i < len) out[i] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// AdamW (decoupled weight decay, bias-corrected moments) on the FP32 master
// parameters, as torch.optim.AdamW computes it, over all eight tensors of the
// network in one launch; each weight's bf16 copies (W [NO][KP] and its
// transpose) used by the kernels are rewritten from the new value.
struct AdamSet {
    float *p[8], *g[8], *m[8], *v[8];
    uint16_t *W[8], *WT[8];   // null for the biases
    int off[9];               // tensor t has off[t+1] - off[t] elements
    int kp[8], no[8];
};

__global__ __launch_bounds__(256) void k_adamw(const AdamSet a, float lr, float b1, float b2, float eps, float wd,
                                               float bc1, float bc2sqrt)
{
    const int t = blockIdx.y;  // one tensor per grid row: uniform index into the argument arrays
    const int n = a.off[t + 1] - a.off[t];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float pi = a.p[t][i] * (1.0f - lr * wd);
        const float gi = a.g[t][i];
        const float mi = b1 * a.m[t][i] + (1.0f - b1) * gi;
        const float vi = b2 * a.v[t][i] + (1.0f - b2) * gi * gi;
        a.m[t][i] = mi;
        a.v[t][i] = vi;
        const float denom = sqrtf(vi) / bc2sqrt + eps;
        pi = pi - (lr / bc1) * (mi / denom);
        a.p[t][i] = pi;
        if (a.W[t]) {
            const uint16_t h = f2bf(pi);
            a.W[t][i] = h;
            a.WT[t][(size_t)(i % a.kp[t]) * a.no[t] + i / a.kp[t]] = h;
        }
    }
}

// FP32 master [NO][KP] -> bf16 W [NO][KP] and W^T [KP][NO].
__global__ void k_to_bf16(const float *__restrict__ p, int no, int kp, uint16_t *__restrict__ W, uint16_t *__restrict__ WT)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < no * kp; i += gridDim.x * blockDim.x) {
        const int n = i / kp, k = i % kp;
        const uint16_t h = f2bf(p[i]);
        W[i] = h;
        if (WT) WT[(size_t)k * no + n] = h;
    }
}

// ---------------------------------------------------------------------------
// inference inside the step

// LDS written by the lanes of one wavefront, read by other lanes of it.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NE>
__device__ __forceinline__ void stage_lds(uint16_t *dst, const uint16_t *__restrict__ src)
{
    static_assert(NE % 8 == 0, "16-byte chunks");
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    for (int e = threadIdx.x; e < NE / 8; e += blockDim.x) d4[e] = s4[e];
}

// Predicted order per listed cell: features from the stage input, the four
// layers on the matrix cores, argmax -> order 2c+1.  The workgroup copies all
// weights to LDS once; each wavefront then takes 16 cells per trip on its own
// activation buffers (four lanes per cell for the features, max |value|
// combined across the four lanes).  The cell list and its count live on the
// device.
template <int S, int KP>
__global__ __launch_bounds__(256) void k_mlp_predict(const double4 *__restrict__ U, int N, int bc, const int *__restrict__ list,
                                                     const int *__restrict__ count, const MlpDev w, int C, int mord,
                                                     order_t *__restrict__ gpo, int8_t *__restrict__ pred,
                                                     long long *__restrict__ lowered)
{
    constexpr int IN = 4 * S * S;
    constexpr int KX = KP > 128 ? KP : 128;
    __shared__ __align__(16) uint16_t sW0[128 * KP];
    __shared__ __align__(16) uint16_t sW1[128 * 128];
    __shared__ __align__(16) uint16_t sW2[64 * 128];
    __shared__ __align__(16) uint16_t sW3[16 * 64];
    __shared__ __align__(16) uint16_t xa[4][16 * KX];
    __shared__ __align__(16) uint16_t xb[4][16 * 128];
    __shared__ float lg[4][16 * 16];
    const int n = *count;
    if (blockIdx.x * 64 >= n) return;  // the same for the whole workgroup
    stage_lds<128 * KP>(sW0, w.W[0]);
    stage_lds<128 * 128>(sW1, w.W[1]);
    stage_lds<64 * 128>(sW2, w.W[2]);
    stage_lds<16 * 64>(sW3, w.W[3]);
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, q = lane >> 4;
    for (int bb = blockIdx.x * 64; bb < n; bb += gridDim.x * 64) {
        const int base = bb + wv * 16;
        if (base >= n) break;  // per wavefront: no workgroup barrier below
        const int t = base + m;
        const int cell = t < n ? list[t] : 0;
        const int ci = cell % N, cj = cell / N;
        float mx[4];
        stencil_absmax<S>(U, N, bc, ci, cj, q, 4, mx);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], 16));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], 32));
        }
        uint16_t *row = &xa[wv][m * KP];
        if (t < n) {
            stencil_write<S>(U, N, bc, ci, cj, q, 4, mx, [&](int e, float v) { row[e] = f2bf(v); });
        } else {
            for (int e = q; e < IN; e += 4) row[e] = 0;
        }
        for (int e = IN + q; e < KP; e += 4) row[e] = 0;
        wave_sync();
        dense16<KP, 128, true, false>(xa[wv], KP, sW0, w.b[0], xb[wv], 128, lane);
        wave_sync();
        dense16<128, 128, true, false>(xb[wv], 128, sW1, w.b[1], xa[wv], 128, lane);
        wave_sync();
        dense16<128, 64, true, false>(xa[wv], 128, sW2, w.b[2], xb[wv], 64, lane);
        wave_sync();
        dense16<64, 16, false, true>(xb[wv], 64, sW3, w.b[3], lg[wv], 16, lane);
        wave_sync();
        if (q == 0 && t < n) {
            const float *z = &lg[wv][m * 16];
            int am = 0;
            float zm = z[0];
            for (int c = 1; c < C; ++c)
                if (z[c] > zm) { zm = z[c]; am = c; }
            const int ord = 2 * am + 1 <= mord ? 2 * am + 1 : mord;
            gpo[cell] = pack_order(ord, ord, 0);
            pred[cell] = (int8_t)ord;
            if (ord < mord) atomicAdd((unsigned long long *)lowered, 1ull);
        }
        wave_sync();  // the next trip rewrites xa
    }
}

// Gate of the a priori prediction: only cells where the a posteriori check
// can fail are sent to the network -- cells the shock detector flags (the DMP
// runs only there, mod_subroutine_MOOD.f90:53-55) and near-vacuum cells
// (positivity); every other cell starts at Mord.  With gate off every cell
// is predicted.
__global__ __launch_bounds__(256) void k_mlp_gate(const double4 *__restrict__ U, int N, int bc, double dx, double dy,
                                                  int gate, int mord, order_t *__restrict__ gpo, int8_t *__restrict__ pred,
                                                  int *__restrict__ list, int *__restrict__ count)
{
    // 32x8 cells per workgroup; the tile and a one-cell ring of the stage
    // input, and each of their pressures, are staged in LDS once
    constexpr int TX = 32, TY = 8, PW = TX + 2, PH = TY + 2;
    __shared__ double4 sU[PW * PH];
    __shared__ double sP[PW * PH];
    const int tiles_x = (N + TX - 1) / TX;
    const int x0 = (blockIdx.x % tiles_x) * TX, y0 = (blockIdx.x / tiles_x) * TY;
    const int tid = threadIdx.x;
    if (gate) {
        for (int idx = tid; idx < PW * PH; idx += 256) {
            const double4 u = load_bc(U, N, bc, x0 - 1 + idx % PW, y0 - 1 + idx / PW);
            sU[idx] = u;
            sP[idx] = pressure(u.x, u.y, u.z, u.w);
        }
        __syncthreads();
    }
    const int lx = tid % TX, ly = tid / TX;
    const int i = x0 + lx, j = y0 + ly;
    bool sel = false;
    if (i < N && j < N) {
        const size_t c = (size_t)j * N + i;
        if (!gate) {
            sel = true;
        } else {
            // shock_detector with the pressures from LDS
            const int q = (ly + 1) * PW + lx + 1;
            const Cell cc = to_cell(sU[q]);
            const double pc = sP[q];
            sel = shock_mach(cc, pc) &&
                  shock_compression(cc, to_cell(sU[q + 1]), to_cell(sU[q - 1]), to_cell(sU[q + PW]),
                                    to_cell(sU[q - PW]), dx, dy) &&
                  shock_pressure_jump(sP[q - 1], sP[q + 1], sP[q - PW], sP[q + PW], dx, dy);
            // near-vacuum cells (positivity)
            if (!sel) sel = !(cc.r > 1e-6) || !(pc > 1e-9);
        }
        if (!sel) {
            gpo[c] = pack_order(mord, mord, 0);
            pred[c] = (int8_t)mord;
        }
    }
    const unsigned long long msk = __ballot(sel);
    if (msk) {
        const int lane = __lane_id();
        const int leader = __ffsll((long long)msk) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(count, __popcll(msk));
        base = __shfl(base, leader);
        if (sel) list[base + __popcll(msk & ((1ull << lane) - 1ull))] = j * N + i;
    }
}

// Training-data dump after an RK stage: for sampled cells, the features of the
// stage input and the stage's final order (label class (ord-1)/2).  Cells
// below Mord are always kept; cells at Mord with probability keep_mord (a
// counter-based hash of (seed, cell) decides).  Only cells the gate would
// send to the network are considered when gated_only is set.
template <int S>
__global__ __launch_bounds__(256) void k_dump(const double4 *__restrict__ U, int N, int bc, double dx, double dy,
                                              const order_t *__restrict__ gpo, int mord, float keep_mord,
                                              unsigned long long seed, int gated_only, float *__restrict__ X,
                                              int *__restrict__ Y, long long cap, unsigned long long *__restrict__ count)
{
    const size_t ncell = (size_t)N * N;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % N), j = (int)(c / N);
        const int ord = order_of_word(gpo[c]);
        if (gated_only) {
            struct GA {
                const double4 *U;
                int N, bc;
                __device__ double4 operator()(int a, int b) const { return load_bc(U, N, bc, a, b); }
            } acc{U, N, bc};
            const MoodParams P{1, 1, 1, dx, dy};
            bool sel = shock_detector(P, acc, i, j);
            if (!sel) {
                const double4 u = U[c];
                const double p = pressure(u.x, u.y, u.z, u.w);
                sel = !(u.x > 1e-6) || !(p > 1e-9);
            }
            if (!sel) continue;
        }
        if (ord == mord) {
            unsigned long long h = seed ^ (0x9E3779B97F4A7C15ull * (c + 1));
            h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull; h ^= h >> 33;
            const float u01 = (float)(h >> 40) * (1.0f / 16777216.0f);
            if (u01 >= keep_mord) continue;
        }
        const unsigned long long slot = atomicAdd(count, 1ull);
        if ((long long)slot >= cap) continue;
        float *row = X + slot * (4 * S * S);
        cell_features<S>(U, N, bc, i, j, [&](int e, float v) { row[e] = v; });
        Y[slot] = (ord - 1) / 2;
    }
}

// Features of listed cells (for offline evaluation and tests).
template <int S>
__global__ void k_features(const double4 *__restrict__ U, int N, int bc, const int *__restrict__ cells, int n,
                           float *__restrict__ X)
{
    constexpr int IN = 4 * S * S;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        float *row = X + (size_t)t * IN;
        cell_features<S>(U, N, bc, cells[t] % N, cells[t] / N, [&](int e, float v) { row[e] = v; });
    }
}

// ---------------------------------------------------------------------------
// host side

hipError_t mlp_launch_gate(const double4 *U, int N, int bc, double dx, double dy, int gate, int mord, order_t *gpo,
                           int8_t *pred, int *list, int *count, hipStream_t s)
{
    const int tiles = ((N + 31) / 32) * ((N + 7) / 8);
    hipLaunchKernelGGL(k_mlp_gate, dim3(tiles), dim3(256), 0, s, U, N, bc, dx, dy, gate, mord, gpo, pred, list, count);
    return hipGetLastError();
}

// one workgroup per CU: each holds all weights in LDS
constexpr int kPredictBlocks = 256;

hipError_t mlp_launch_predict(const double4 *U, int N, int bc, const int *list, const int *count, const MlpDev &w,
                              int stencil, int C, int mord, order_t *gpo, int8_t *pred, long long *lowered,
                              hipStream_t s)
{
    if (stencil == 5)
        hipLaunchKernelGGL((k_mlp_predict<5, 128>), dim3(kPredictBlocks), dim3(256), 0, s, U, N, bc, list, count, w, C, mord, gpo,
                           pred, lowered);
    else if (stencil == 7)
        hipLaunchKernelGGL((k_mlp_predict<7, 224>), dim3(kPredictBlocks), dim3(256), 0, s, U, N, bc, list, count, w, C, mord, gpo,
                           pred, lowered);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t mlp_launch_dump(const double4 *U, int N, int bc, double dx, double dy, const order_t *gpo, int mord,
                           int stencil, float keep_mord, unsigned long long seed, int gated_only, float *X, int *Y,
                           long long cap, unsigned long long *count, hipStream_t s)
{
    if (stencil == 5)
        hipLaunchKernelGGL(k_dump<5>, dim3(2048), dim3(256), 0, s, U, N, bc, dx, dy, gpo, mord, keep_mord, seed, gated_only, X, Y, cap, count);
    else if (stencil == 7)
        hipLaunchKernelGGL(k_dump<7>, dim3(2048), dim3(256), 0, s, U, N, bc, dx, dy, gpo, mord, keep_mord, seed, gated_only, X, Y, cap, count);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t mlp_launch_features(const double4 *U, int N, int bc, int stencil, const int *cells, int n, float *X, hipStream_t s)
{
    if (stencil == 5) hipLaunchKernelGGL(k_features<5>, dim3(256), dim3(256), 0, s, U, N, bc, cells, n, X);
    else if (stencil == 7) hipLaunchKernelGGL(k_features<7>, dim3(256), dim3(256), 0, s, U, N, bc, cells, n, X);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace gpm

// ---------------------------------------------------------------------------
// C ABI of the predictor (declared in include/gpmood.h)

using namespace gpm;

struct gpm_mlp {
    int in = 0, kp = 0, C = 0, stencil = 0, max_batch = 0;
    hipStream_t stream = nullptr;
    // layer l: NO[l] x KPL[l]
    int NO[4] = {128, 128, 64, 16}, KPL[4] = {0, 128, 128, 64};
    float *p[4] = {}, *pb[4] = {};       // fp32 master weights [NO][KPL], biases [NO]
    float *g[4] = {}, *gb[4] = {};       // gradients
    float *m[4] = {}, *v[4] = {}, *mb[4] = {}, *vb[4] = {};
    uint16_t *W[4] = {}, *WT[4] = {};    // bf16 copies
    uint16_t *H[3] = {};                 // activations [max_batch][NO]
    float *logits = nullptr;             // [max_batch][16]
    uint16_t *dH[4] = {};                // gradients of layer outputs (bf16)
    float *part = nullptr, *partb = nullptr, *lossp = nullptr;
    long long step = 0;
    std::string msg;
};

namespace {
int hchk(gpm_mlp *m, hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    if (m) m->msg = std::string(what) + ": " + hipGetErrorString(e);
    return -2;
}
#define MCHK(m, expr) do { int r_ = hchk((m), (expr), #expr); if (r_) return r_; } while (0)

int mlp_in_kp(int in) { return (in + 31) / 32 * 32; }  // x32 fragments tile the input

template <class T>
hipError_t zalloc(T **p, size_t n)
{
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
    if (e == hipSuccess) e = hipMemset(*p, 0, n * sizeof(T));
    return e;
}

int refresh_bf16(gpm_mlp *m)
{
    for (int l = 0; l < 4; ++l) {
        hipLaunchKernelGGL(k_to_bf16, dim3(128), dim3(256), 0, m->stream, m->p[l], m->NO[l], m->KPL[l], m->W[l], m->WT[l]);
        MCHK(m, hipGetLastError());
    }
    return 0;
}
}  // namespace

namespace gpm {
MlpDev mlp_device_view(const gpm_mlp *m)
{
    MlpDev d{};
    for (int l = 0; l < 4; ++l) { d.W[l] = m->W[l]; d.b[l] = m->pb[l]; }
    return d;
}
int mlp_stencil(const gpm_mlp *m) { return m->stencil; }
int mlp_classes(const gpm_mlp *m) { return m->C; }
}  // namespace gpm

extern "C" {

int gpm_mlp_create(int32_t in_features, int32_t n_classes, int32_t max_batch, gpm_mlp **out)
{
    if (!out) return -1;
    *out = nullptr;
    int stencil = 0;
    if (in_features == 100) stencil = 5;
    else if (in_features == 196) stencil = 7;
    else return -4;
    if (n_classes < 2 || n_classes > 16 || max_batch < 1) return -1;
    gpm_mlp *m = new gpm_mlp();
    m->in = in_features;
    m->kp = mlp_in_kp(in_features);
    m->C = n_classes;
    m->stencil = stencil;
    m->max_batch = (max_batch + 63) / 64 * 64;
    m->KPL[0] = m->kp;
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    const size_t B = m->max_batch;
    for (int l = 0; l < 4 && e == hipSuccess; ++l) {
        const size_t nw = (size_t)m->NO[l] * m->KPL[l];
        for (float **q : {&m->p[l], &m->g[l], &m->m[l], &m->v[l]})
            if (e == hipSuccess) e = zalloc(q, nw);
        for (float **q : {&m->pb[l], &m->gb[l], &m->mb[l], &m->vb[l]})
            if (e == hipSuccess) e = zalloc(q, (size_t)m->NO[l]);
        if (e == hipSuccess) e = zalloc(&m->W[l], nw);
        if (e == hipSuccess) e = zalloc(&m->WT[l], nw);
        if (e == hipSuccess) e = zalloc(&m->dH[l], B * m->NO[l]);
        if (l < 3 && e == hipSuccess) e = zalloc(&m->H[l], B * m->NO[l]);
    }
    if (e == hipSuccess) e = zalloc(&m->logits, B * 16);
    if (e == hipSuccess) e = zalloc(&m->part, (B / 64) * (size_t)128 * 224);
    if (e == hipSuccess) e = zalloc(&m->partb, (B / 64) * (size_t)128);
    if (e == hipSuccess) e = zalloc(&m->lossp, 2 * (B / 256 + 1));
    if (e != hipSuccess) {
        gpm_mlp_destroy(m);
        return -3;
    }
    *out = m;
    return 0;
}

int gpm_mlp_destroy(gpm_mlp *m)
{
    if (!m) return 0;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (int l = 0; l < 4; ++l) {
        void *bufs[] = {m->p[l], m->pb[l], m->g[l], m->gb[l], m->m[l], m->v[l], m->mb[l], m->vb[l], m->W[l], m->WT[l],
                        m->dH[l], l < 3 ? m->H[l] : nullptr};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }
    void *bufs[] = {m->logits, m->part, m->partb, m->lossp};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

const char *gpm_mlp_last_error(const gpm_mlp *m) { return m ? m->msg.c_str() : "null"; }

int64_t gpm_mlp_num_params(const gpm_mlp *m)
{
    if (!m) return -1;
    const int real_in[4] = {m->in, 128, 128, 64}, real_out[4] = {128, 128, 64, m->C};
    int64_t n = 0;
    for (int l = 0; l < 4; ++l) n += (int64_t)real_out[l] * real_in[l] + real_out[l];
    return n;
}

// params in PyTorch order: W0 [128][in], b0, W1 [128][128], b1, W2 [64][128], b2, W3 [C][64], b3
static int xfer_params(gpm_mlp *m, float *host, bool to_dev, bool grads)
{
    const int real_in[4] = {m->in, 128, 128, 64}, real_out[4] = {128, 128, 64, m->C};
    size_t off = 0;
    for (int l = 0; l < 4; ++l) {
        float *dw = grads ? m->g[l] : m->p[l];
        float *db = grads ? m->gb[l] : m->pb[l];
        std::vector<float> tmp((size_t)m->NO[l] * m->KPL[l], 0.f);
        if (!to_dev) MCHK(m, hipMemcpyAsync(tmp.data(), dw, tmp.size() * 4, hipMemcpyDeviceToHost, m->stream));
        if (!to_dev) MCHK(m, hipStreamSynchronize(m->stream));
        for (int n = 0; n < real_out[l]; ++n)
            for (int k = 0; k < real_in[l]; ++k) {
                float &d = tmp[(size_t)n * m->KPL[l] + k];
                if (to_dev) d = host[off + (size_t)n * real_in[l] + k];
                else host[off + (size_t)n * real_in[l] + k] = d;
            }
        off += (size_t)real_out[l] * real_in[l];
        std::vector<float> tb(m->NO[l], 0.f);
        if (to_dev) {
            for (int n = 0; n < real_out[l]; ++n) tb[n] = host[off + n];
            MCHK(m, hipMemcpyAsync(dw, tmp.data(), tmp.size() * 4, hipMemcpyHostToDevice, m->stream));
            MCHK(m, hipMemcpyAsync(db, tb.data(), tb.size() * 4, hipMemcpyHostToDevice, m->stream));
            MCHK(m, hipStreamSynchronize(m->stream));
        } else {
            MCHK(m, hipMemcpyAsync(tb.data(), db, tb.size() * 4, hipMemcpyDeviceToHost, m->stream));
            MCHK(m, hipStreamSynchronize(m->stream));
            for (int n = 0; n < real_out[l]; ++n) host[off + n] = tb[n];
        }
        off += real_out[l];
    }
    return 0;
}

int gpm_mlp_set_params(gpm_mlp *m, const float *params)
{
    if (!m || !params) return -1;
    int rc = xfer_params(m, const_cast<float *>(params), true, false);
    if (rc) return rc;
    for (int l = 0; l < 4; ++l) {
        const size_t nw = (size_t)m->NO[l] * m->KPL[l];
        MCHK(m, hipMemsetAsync(m->m[l], 0, nw * 4, m->stream));
        MCHK(m, hipMemsetAsync(m->v[l], 0, nw * 4, m->stream));
        MCHK(m, hipMemsetAsync(m->mb[l], 0, m->NO[l] * 4, m->stream));
        MCHK(m, hipMemsetAsync(m->vb[l], 0, m->NO[l] * 4, m->stream));
    }
    m->step = 0;
    rc = refresh_bf16(m);
    if (rc) return rc;
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

int gpm_mlp_get_params(gpm_mlp *m, float *params)
{
    if (!m || !params) return -1;
    return xfer_params(m, params, false, false);
}

int gpm_mlp_get_grads(gpm_mlp *m, float *grads)
{
    if (!m || !grads) return -1;
    return xfer_params(m, grads, false, true);
}

// Forward + loss + backward on a batch in device memory (x fp32 [batch][in],
// y int32 classes); leaves the gradients of the mean cross-entropy in place.
int gpm_mlp_forward_backward(gpm_mlp *m, const float *x, const int32_t *y, int32_t batch, double *loss, double *accuracy)
{
    if (!m || !x || !y || batch < 1 || batch > m->max_batch) return -1;
    const int B = batch, nb64 = (B + 63) / 64, nb256 = (B + 255) / 256;
    hipStream_t s = m->stream;
    if (m->kp == 128) hipLaunchKernelGGL(k_fwd_in<128>, dim3(nb64), dim3(256), 0, s, x, m->in, B, m->W[0], m->pb[0], m->H[0]);
    else hipLaunchKernelGGL(k_fwd_in<224>, dim3(nb64), dim3(256), 0, s, x, m->in, B, m->W[0], m->pb[0], m->H[0]);
    hipLaunchKernelGGL((k_fwd_hidden<128, 128, true, false>), dim3(nb64), dim3(256), 0, s, m->H[0], B, m->W[1], m->pb[1], (void *)m->H[1]);
    hipLaunchKernelGGL((k_fwd_hidden<128, 64, true, false>), dim3(nb64), dim3(256), 0, s, m->H[1], B, m->W[2], m->pb[2], (void *)m->H[2]);
    hipLaunchKernelGGL((k_fwd_hidden<64, 16, false, true>), dim3(nb64), dim3(256), 0, s, m->H[2], B, m->W[3], m->pb[3], (void *)m->logits);
    hipLaunchKernelGGL(k_softmax_ce, dim3(nb256), dim3(256), 0, s, m->logits, y, B, m->C, m->dH[3], m->lossp);
    // layer 3: dW3 = dlogits^T H2, dH2 = dlogits W3 * (H2 > 0)
    hipLaunchKernelGGL((k_bwd_weight<16, 64, false>), dim3(nb64), dim3(256), 0, s, m->dH[3], (const void *)m->H[2], 64, B, m->part, m->partb);
    hipLaunchKernelGGL(k_reduce_parts, dim3((16 * 64 + 63) / 64), dim3(256), 0, s, m->part, nb64, 16 * 64, m->g[3]);
    hipLaunchKernelGGL(k_reduce_parts, dim3((16 + 63) / 64), dim3(256), 0, s, m->partb, nb64, 16, m->gb[3]);
    hipLaunchKernelGGL((k_bwd_data<16, 64>), dim3(nb64), dim3(256), 0, s, m->dH[3], B, m->WT[3], m->H[2], m->dH[2]);
    hipLaunchKernelGGL((k_bwd_weight<64, 128, false>), dim3(nb64), dim3(256), 0, s, m->dH[2], (const void *)m->H[1], 128, B, m->part, m->partb);
    hipLaunchKernelGGL(k_reduce_parts, dim3((64 * 128 + 63) / 64), dim3(256), 0, s, m->part, nb64, 64 * 128, m->g[2]);
    hipLaunchKernelGGL(k_reduce_parts, dim3((64 + 63) / 64), dim3(256), 0, s, m->partb, nb64, 64, m->gb[2]);
    hipLaunchKernelGGL((k_bwd_data<64, 128>), dim3(nb64), dim3(256), 0, s, m->dH[2], B, m->WT[2], m->H[1], m->dH[1]);
    hipLaunchKernelGGL((k_bwd_weight<128, 128, false>), dim3(nb64), dim3(256), 0, s, m->dH[1], (const void *)m->H[0], 128, B, m->part, m->partb);
    hipLaunchKernelGGL(k_reduce_parts, dim3((128 * 128 + 63) / 64), dim3(256), 0, s, m->part, nb64, 128 * 128, m->g[1]);
    hipLaunchKernelGGL(k_reduce_parts, dim3((128 + 63) / 64), dim3(256), 0, s, m->partb, nb64, 128, m->gb[1]);
    hipLaunchKernelGGL((k_bwd_data<128, 128>), dim3(nb64), dim3(256), 0, s, m->dH[1], B, m->WT[1], m->H[0], m->dH[0]);
    if (m->kp == 128) hipLaunchKernelGGL((k_bwd_weight<128, 128, true>), dim3(nb64), dim3(256), 0, s, m->dH[0], (const void *)x, m->in, B, m->part, m->partb);
    else hipLaunchKernelGGL((k_bwd_weight<128, 224, true>), dim3(nb64), dim3(256), 0, s, m->dH[0], (const void *)x, m->in, B, m->part, m->partb);
    hipLaunchKernelGGL(k_reduce_parts, dim3((128 * m->kp + 63) / 64), dim3(256), 0, s, m->part, nb64, 128 * m->kp, m->g[0]);
    hipLaunchKernelGGL(k_reduce_parts, dim3((128 + 63) / 64), dim3(256), 0, s, m->partb, nb64, 128, m->gb[0]);
    MCHK(m, hipGetLastError());
    if (loss || accuracy) {
        std::vector<float> lp(2 * nb256);
        MCHK(m, hipMemcpyAsync(lp.data(), m->lossp, lp.size() * 4, hipMemcpyDeviceToHost, s));
        MCHK(m, hipStreamSynchronize(s));
        double ls = 0, ac = 0;
        for (int b = 0; b < nb256; ++b) { ls += lp[2 * b]; ac += lp[2 * b + 1]; }
        if (loss) *loss = ls / B;
        if (accuracy) *accuracy = ac / B;
    }
    return 0;
}

int gpm_mlp_adamw_step(gpm_mlp *m, double lr, double beta1, double beta2, double eps, double weight_decay)
{
    if (!m) return -1;
    m->step += 1;
    const float bc1 = (float)(1.0 - std::pow(beta1, (double)m->step));
    const float bc2s = (float)std::sqrt(1.0 - std::pow(beta2, (double)m->step));
    AdamSet a{};
    a.off[0] = 0;
    for (int l = 0; l < 4; ++l) {
        const int t = 2 * l, tb = 2 * l + 1;
        a.p[t] = m->p[l], a.g[t] = m->g[l], a.m[t] = m->m[l], a.v[t] = m->v[l];
        a.W[t] = m->W[l], a.WT[t] = m->WT[l], a.kp[t] = m->KPL[l], a.no[t] = m->NO[l];
        a.off[t + 1] = a.off[t] + m->NO[l] * m->KPL[l];
        a.p[tb] = m->pb[l], a.g[tb] = m->gb[l], a.m[tb] = m->mb[l], a.v[tb] = m->vb[l];
        a.W[tb] = nullptr, a.WT[tb] = nullptr, a.kp[tb] = 1, a.no[tb] = m->NO[l];
        a.off[tb + 1] = a.off[tb] + m->NO[l];
    }
    int most = 0;
    for (int t = 0; t < 8; ++t) most = std::max(most, a.off[t + 1] - a.off[t]);
    hipLaunchKernelGGL(k_adamw, dim3((most + 255) / 256, 8), dim3(256), 0, m->stream, a, (float)lr, (float)beta1,
                       (float)beta2, (float)eps, (float)weight_decay, bc1, bc2s);
    MCHK(m, hipGetLastError());
    return 0;
}

int gpm_mlp_train_step(gpm_mlp *m, const float *x, const int32_t *y, int32_t batch, double lr, double weight_decay,
                       double *loss, double *accuracy)
{
    int rc = gpm_mlp_forward_backward(m, x, y, batch, loss, accuracy);
    if (rc) return rc;
    return gpm_mlp_adamw_step(m, lr, 0.9, 0.999, 1e-8, weight_decay);
}

// Class logits for n feature rows (device memory), written to logits [n][C].
int gpm_mlp_logits(gpm_mlp *m, const float *x, int32_t n, float *logits)
{
    if (!m || !x || !logits || n < 1) return -1;
    for (int off = 0; off < n; off += m->max_batch) {
        const int B = n - off < m->max_batch ? n - off : m->max_batch;
        const int nb64 = (B + 63) / 64;
        hipStream_t s = m->stream;
        const float *xb = x + (size_t)off * m->in;
        if (m->kp == 128) hipLaunchKernelGGL(k_fwd_in<128>, dim3(nb64), dim3(256), 0, s, xb, m->in, B, m->W[0], m->pb[0], m->H[0]);
        else hipLaunchKernelGGL(k_fwd_in<224>, dim3(nb64), dim3(256), 0, s, xb, m->in, B, m->W[0], m->pb[0], m->H[0]);
        hipLaunchKernelGGL((k_fwd_hidden<128, 128, true, false>), dim3(nb64), dim3(256), 0, s, m->H[0], B, m->W[1], m->pb[1], (void *)m->H[1]);
        hipLaunchKernelGGL((k_fwd_hidden<128, 64, true, false>), dim3(nb64), dim3(256), 0, s, m->H[1], B, m->W[2], m->pb[2], (void *)m->H[2]);
        hipLaunchKernelGGL((k_fwd_hidden<64, 16, false, true>), dim3(nb64), dim3(256), 0, s, m->H[2], B, m->W[3], m->pb[3], (void *)m->logits);
        MCHK(m, hipGetLastError());
        MCHK(m, hipMemcpy2DAsync(logits + (size_t)off * m->C, m->C * sizeof(float), m->logits, 16 * sizeof(float),
                                 m->C * sizeof(float), B, hipMemcpyDeviceToDevice, s));
    }
    MCHK(m, hipStreamSynchronize(m->stream));
    return 0;
}

}  // extern "C"
