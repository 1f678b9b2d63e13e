// Adjoint of the SSP-RK3 step with the MOOD order mask and dt frozen.
//
// Frozen, one RK stage is W = V + L_m(V): GP reconstruction of both sides of
// every face at its Gauss points with the face order min(m(A), m(B)), HLLC at
// each point, Gauss quadrature and the four-face update (the same arithmetic
// as the forward kernels).  Its transpose applied to an output adjoint Wb runs
// as three deterministic kernels (no atomics):
//   k_adj_face    one thread per (face, Gauss point): Fb = dt/dx (Wb(B) - Wb(A))
//                 times the Gauss weight, the reconstructed states, and the
//                 hand-written reverse mode of the HLLC solver; stores the
//                 adjoint of each side's reconstructed state per point.
//   k_adj_gather  one thread per cell of the ghost-extended grid: the transpose
//                 of the reconstruction, sum over the stencil points s and the
//                 four sides of cell p - off(s) of zT(o, side, s, g) * ub.
//   k_adj_fold    one thread per interior cell: the transpose of the ghost
//                 fill (ghost adjoints summed into the cells they copy, with
//                 the normal momentum negated for reflecting walls), plus the
//                 identity part and the RK combination terms.
#include "gpm_adjoint.h"
#include "gpm_device.h"

namespace gpm {

// Reverse mode of hllc_adm<0> at an admissible pair of states: given
// Fb = dJ/dF, returns ulb = dJ/dul and urb = dJ/dur.  Branches and MIN/MAX
// selections are those the forward evaluation takes.
__device__ __forceinline__ void hllc_x_vjp(const double ul[4], const double ur[4], const double Fb[4], double ulb[4],
                                           double urb[4])
{
    constexpr double g1 = kGamma - 1.0;
    const double rL = ul[0], mxL = ul[1], myL = ul[2], EL = ul[3];
    const double rR = ur[0], mxR = ur[1], myR = ur[2], ER = ur[3];
    const double vxl = mxL / rL, vyl = myL / rL;
    const double pl = pressure(rL, mxL, myL, EL);
    const double cl = sqrt(kGamma * pl / rL);
    const double vxr = mxR / rR, vyr = myR / rR;
    const double pr = pressure(rR, mxR, myR, ER);
    const double cr = sqrt(kGamma * pr / rR);
    const bool sl_left = (vxl - cl) < (vxr - cr);
    const double SL = sl_left ? vxl - cl : vxr - cr;
    const bool sr_left = (vxl + cl) > (vxr + cr);
    const double SR = sr_left ? vxl + cl : vxr + cr;
    const double nL = SL - vxl, nR = SR - vxr;
    const double num = (pr - pl) + (rL * vxl * nL - rR * vxr * nR);
    const double den = rL * nL - rR * nR;
    const double uS = num / den;
    // Above: the forward quantities, in the forward's arithmetic, so that the
    // branches taken are the forward's.  Below: the reverse sweep, which needs
    // no bit-level agreement with anything and may contract to fused
    // multiply-adds.
  {
#pragma clang fp contract(fast)
    double aL[4] = {0.0, 0.0, 0.0, 0.0}, aR[4] = {0.0, 0.0, 0.0, 0.0};  // direct adjoints of ul, ur
    double Flb[4] = {0.0, 0.0, 0.0, 0.0}, Frb[4] = {0.0, 0.0, 0.0, 0.0};
    double vxlb = 0.0, vylb = 0.0, plb = 0.0, clb = 0.0;
    double vxrb = 0.0, vyrb = 0.0, prb = 0.0, crb = 0.0;
    double SLb = 0.0, SRb = 0.0, uSb = 0.0, nLb = 0.0, nRb = 0.0;
    bool star = false;
    if (SL >= 0.) {
        for (int k = 0; k < 4; ++k) Flb[k] = Fb[k];
    } else if (SL < 0. && uS >= 0.) {
        // F = Fl + SL*(Us*dL - ul), dL = rL nL/(SL - uS),
        // Us = (1, uS, vyl, EL/rL + (uS - vxl)(uS + pl/(rL nL)))
        star = true;
        const double D = SL - uS, dL = rL * nL / D, q = pl / (rL * nL), B = uS + q;
        const double Es = EL / rL + (uS - vxl) * B;
        const double Us[4] = {1.0, uS, vyl, Es};
        double Usb[4], dLb = 0.0;
        for (int k = 0; k < 4; ++k) {
            Flb[k] = Fb[k];
            SLb += Fb[k] * (Us[k] * dL - ul[k]);
            Usb[k] = Fb[k] * SL * dL;
            dLb += Fb[k] * SL * Us[k];
            aL[k] -= Fb[k] * SL;
        }
        uSb += Usb[1];
        vylb += Usb[2];
        const double Esb = Usb[3];
        aL[3] += Esb / rL;
        aL[0] -= Esb * EL / (rL * rL);
        uSb += Esb * B;
        vxlb -= Esb * B;
        const double Bb = Esb * (uS - vxl);
        uSb += Bb;
        plb += Bb / (rL * nL);
        aL[0] -= Bb * q / rL;
        nLb -= Bb * q / nL;
        aL[0] += dLb * nL / D;
        nLb += dLb * rL / D;
        const double Db = -dLb * dL / D;
        SLb += Db;
        uSb -= Db;
    } else if (uS < 0. && SR >= 0.) {
        star = true;
        const double D = SR - uS, dR = rR * nR / D, q = pr / (rR * nR), B = uS + q;
        const double Es = ER / rR + (uS - vxr) * B;
        const double Us[4] = {1.0, uS, vyr, Es};
        double Usb[4], dRb = 0.0;
        for (int k = 0; k < 4; ++k) {
            Frb[k] = Fb[k];
            SRb += Fb[k] * (Us[k] * dR - ur[k]);
            Usb[k] = Fb[k] * SR * dR;
            dRb += Fb[k] * SR * Us[k];
            aR[k] -= Fb[k] * SR;
        }
        uSb += Usb[1];
        vyrb += Usb[2];
        const double Esb = Usb[3];
        aR[3] += Esb / rR;
        aR[0] -= Esb * ER / (rR * rR);
        uSb += Esb * B;
        vxrb -= Esb * B;
        const double Bb = Esb * (uS - vxr);
        uSb += Bb;
        prb += Bb / (rR * nR);
        aR[0] -= Bb * q / rR;
        nRb -= Bb * q / nR;
        aR[0] += dRb * nR / D;
        nRb += dRb * rR / D;
        const double Db = -dRb * dR / D;
        SRb += Db;
        uSb -= Db;
    } else {
        for (int k = 0; k < 4; ++k) Frb[k] = Fb[k];
    }
    if (star) {
        // uS = num/den, num = rL vxl nL - rR vxr nR + (pr - pl), den = rL nL - rR nR
        const double numb = uSb / den, denb = -uSb * uS / den;
        aL[0] += numb * vxl * nL + denb * nL;
        vxlb += numb * rL * nL;
        nLb += numb * rL * vxl + denb * rL;
        aR[0] -= numb * vxr * nR + denb * nR;
        vxrb -= numb * rR * nR;
        nRb -= numb * rR * vxr + denb * rR;
        prb += numb;
        plb -= numb;
        // nL = SL - vxl, nR = SR - vxr
        SLb += nLb;
        vxlb -= nLb;
        SRb += nRb;
        vxrb -= nRb;
        if (sl_left) { vxlb += SLb; clb -= SLb; } else { vxrb += SLb; crb -= SLb; }
        if (sr_left) { vxlb += SRb; clb += SRb; } else { vxrb += SRb; crb += SRb; }
    }
    // physical fluxes Fl = (mx, r vx^2 + p, my vx, vx (E + p)), likewise Fr
    aL[1] += Flb[0];
    aL[0] += Flb[1] * (vxl * vxl);
    vxlb += Flb[1] * 2.0 * rL * vxl + Flb[2] * myL + Flb[3] * (EL + pl);
    plb += Flb[1] + Flb[3] * vxl;
    aL[2] += Flb[2] * vxl;
    aL[3] += Flb[3] * vxl;
    aR[1] += Frb[0];
    aR[0] += Frb[1] * (vxr * vxr);
    vxrb += Frb[1] * 2.0 * rR * vxr + Frb[2] * myR + Frb[3] * (ER + pr);
    prb += Frb[1] + Frb[3] * vxr;
    aR[2] += Frb[2] * vxr;
    aR[3] += Frb[3] * vxr;
    // c = sqrt(gamma p / r); p = (gamma-1)(E - 0.5 (mx^2 + my^2)/r); v = m/r
    plb += clb * kGamma / (2.0 * cl * rL);
    aL[0] -= clb * cl / (2.0 * rL);
    prb += crb * kGamma / (2.0 * cr * rR);
    aR[0] -= crb * cr / (2.0 * rR);
    aL[3] += plb * g1;
    aL[1] -= plb * g1 * vxl;
    aL[2] -= plb * g1 * vyl;
    aL[0] += plb * g1 * 0.5 * (vxl * vxl + vyl * vyl);
    aR[3] += prb * g1;
    aR[1] -= prb * g1 * vxr;
    aR[2] -= prb * g1 * vyr;
    aR[0] += prb * g1 * 0.5 * (vxr * vxr + vyr * vyr);
    aL[1] += vxlb / rL;
    aL[2] += vylb / rL;
    aL[0] -= (vxlb * vxl + vylb * vyl) / rL;
    aR[1] += vxrb / rR;
    aR[2] += vyrb / rR;
    aR[0] -= (vxrb * vxr + vyrb * vyr) / rR;
    for (int k = 0; k < 4; ++k) {
        ulb[k] = aL[k];
        urb[k] = aR[k];
    }
  }
}

// Reverse mode of riemann_dir<kHLLC>: a y-face solve is the x-solve on momentum-swapped
// states with the flux's momenta swapped back.
__device__ __forceinline__ void hllc_vjp(int dir, const double ul[4], const double ur[4], const double Fb[4],
                                         double ulb[4], double urb[4])
{
    if (dir == 0) {
        hllc_x_vjp(ul, ur, Fb, ulb, urb);
        return;
    }
    const double sl[4] = {ul[0], ul[2], ul[1], ul[3]}, sr[4] = {ur[0], ur[2], ur[1], ur[3]};
    const double Gb[4] = {Fb[0], Fb[2], Fb[1], Fb[3]};
    double a[4], b[4];
    hllc_x_vjp(sl, sr, Gb, a, b);
    ulb[0] = a[0]; ulb[1] = a[2]; ulb[2] = a[1]; ulb[3] = a[3];
    urb[0] = b[0]; urb[1] = b[2]; urb[2] = b[1]; urb[3] = b[3];
}

__device__ __forceinline__ int mask_at(const int8_t *__restrict__ m, int N, int i, int j, int mord)
{
    return (i < 0 || i >= N || j < 0 || j >= N) ? mord : (int)m[(size_t)j * N + i];
}

__device__ __forceinline__ double4 cell_or_zero(const double4 *__restrict__ W, int N, int i, int j)
{
    return (i < 0 || i >= N || j < 0 || j >= N) ? make_double4(0.0, 0.0, 0.0, 0.0) : W[(size_t)j * N + i];
}

constexpr int kAdjTX = 32, kAdjTY = 8;  // 256 threads, one item per thread and trip

// Reconstructions of both sides of a face at Gauss point g with the face order
// known at run time (dot_product order of mod_GP_MOOD.f90:46-65).
template <int MORD, class Acc>
__device__ __forceinline__ void recon_pair(const DevTables *__restrict__ T, const Acc &acc, int ord, int fa, int fb,
                                           int g, int ai, int aj, int bi, int bj, double ua[4], double ub[4])
{
    const int oi = (ord - 1) / 2, sz = stencil_sz(ord);
    const double *__restrict__ za_row = &T->zT[oi][fa][0][g];
    const double *__restrict__ zb_row = &T->zT[oi][fb][0][g];
    for (int k = 0; k < 4; ++k) { ua[k] = 0.0; ub[k] = 0.0; }
#pragma unroll 5
    for (int s = 0; s < stencil_sz(MORD); ++s) {
        if (s >= sz) break;
        const double4 qa = acc(ai + kSX[s], aj + kSY[s]);
        const double4 qb = acc(bi + kSX[s], bj + kSY[s]);
        const double za = za_row[s * 5], zb = zb_row[s * 5];
        ua[0] = ua[0] + za * qa.x; ua[1] = ua[1] + za * qa.y; ua[2] = ua[2] + za * qa.z; ua[3] = ua[3] + za * qa.w;
        ub[0] = ub[0] + zb * qb.x; ub[1] = ub[1] + zb * qb.y; ub[2] = ub[2] + zb * qb.z; ub[3] = ub[3] + zb * qb.w;
    }
}

struct AdjLds {
    const double4 *s;
    int pitch;
    __device__ __forceinline__ double4 operator()(int i, int j) const { return s[j * pitch + i]; }
};

// Faces are owned by the cell on their high side: the x-face left of cell
// (cx, cy), cx in [0, N], and the y-face below it, cy in [0, N].  A workgroup
// takes a 32x8 tile of owner cells, stages the stage input of the tile with a
// halo of R+1 cells and the output adjoint and orders of the tile and its
// low-side ring in LDS, then runs one (face, Gauss point) item per thread and
// trip.
template <int MORD, int NGP>
__global__ __launch_bounds__(256) void k_adj_face(AdjFaceArgs a)
{
    constexpr int R = (MORD - 1) / 2, TX = kAdjTX, TY = kAdjTY;
    constexpr int PW = TX + 2 * R + 1, PH = TY + 2 * R + 1;  // cells x0-1-R .. x0+TX-1+R
    constexpr int WW = TX + 1, WH = TY + 1;                  // cells x0-1 .. x0+TX-1
    __shared__ double4 sV[PW * PH];
    __shared__ double4 sW[WW * WH];
    __shared__ int8_t sM[WW * WH];
    const int N = a.N, tid = threadIdx.x;
    const int tiles_x = (N + TX) / TX;  // owner cells [0, N]
    const int x0 = (blockIdx.x % tiles_x) * TX, y0 = (blockIdx.x / tiles_x) * TY;
    for (int idx = tid; idx < PW * PH; idx += 256)
        sV[idx] = load_bc(a.V, N, a.bc, x0 - 1 - R + idx % PW, y0 - 1 - R + idx / PW);
    for (int idx = tid; idx < WW * WH; idx += 256) {
        const int ci = x0 - 1 + idx % WW, cj = y0 - 1 + idx / WW;
        sW[idx] = cell_or_zero(a.Wb, N, ci, cj);
        sM[idx] = (int8_t)mask_at(a.mask, N, ci, cj, MORD);
    }
    __syncthreads();
    const AdjLds acc{sV, PW};
    const int W = N + 2;
    const size_t plane = (size_t)W * W;
    for (int it = tid; it < 2 * NGP * TX * TY; it += 256) {
        const int lx = it % TX, ly = (it / TX) % TY, g = (it / (TX * TY)) % NGP, dir = it / (TX * TY * NGP);
        const int bx = x0 + lx, by = y0 + ly;  // owner cell B
        if (dir == 0 ? (bx > N || by > N - 1) : (bx > N - 1 || by > N)) continue;
        const int ax = dir == 0 ? bx - 1 : bx, ay = dir == 0 ? by : by - 1;
        const int wa = (ay - y0 + 1) * WW + (ax - x0 + 1), wb = (by - y0 + 1) * WW + (bx - x0 + 1);
        const int o = min((int)sM[wa], (int)sM[wb]);
        const int fa = dir == 0 ? kR : kT, fb = dir == 0 ? kL : kB;
        double ua[4], ub[4];
        recon_pair<MORD>(a.T, acc, o, fa, fb, g, ax - x0 + 1 + R, ay - y0 + 1 + R, bx - x0 + 1 + R,
                         by - y0 + 1 + R, ua, ub);
        const double4 wA = sW[wa], wB = sW[wb];
        const double c = a.alpha * (dir == 0 ? a.dtdx : a.dtdy) * a.T->gw[g];
        const double Fb[4] = {c * (wB.x - wA.x), c * (wB.y - wA.y), c * (wB.z - wA.z), c * (wB.w - wA.w)};
        double uab[4], ubb[4];
        hllc_vjp(dir, ua, ub, Fb, uab, ubb);
        a.ub[((size_t)fa * NGP + g) * plane + (size_t)(ay + 1) * W + (ax + 1)] = make_double4(uab[0], uab[1], uab[2], uab[3]);
        a.ub[((size_t)fb * NGP + g) * plane + (size_t)(by + 1) * W + (bx + 1)] = make_double4(ubb[0], ubb[1], ubb[2], ubb[3]);
        if (g == 0) {
            if (dir == 0) a.ordx[(size_t)ay * (N + 1) + (ax + 1)] = (int8_t)o;
            else a.ordy[(size_t)(ay + 1) * N + ax] = (int8_t)o;
        }
    }
}

// Order of the face on side `side` of cell (ci, cj), 0 if that face does not
// exist (faces lie between cells of [-1, N]^2 with at least one interior side).
__device__ __forceinline__ int side_order(const AdjGatherArgs &a, int side, int ci, int cj)
{
    const int N = a.N;
    if (side == kR) return (ci >= -1 && ci <= N - 1 && cj >= 0 && cj <= N - 1) ? a.ordx[(size_t)cj * (N + 1) + ci + 1] : 0;
    if (side == kL) return (ci >= 0 && ci <= N && cj >= 0 && cj <= N - 1) ? a.ordx[(size_t)cj * (N + 1) + ci] : 0;
    if (side == kT) return (cj >= -1 && cj <= N - 1 && ci >= 0 && ci <= N - 1) ? a.ordy[(size_t)(cj + 1) * N + ci] : 0;
    return (cj >= 0 && cj <= N && ci >= 0 && ci <= N - 1) ? a.ordy[(size_t)cj * N + ci] : 0;
}

// Transpose of the reconstruction over the ghost-extended grid [-(R+1), N+R]^2:
// G(p) = sum over sides, stencil points s and Gauss points g of
// zT(o, side, s, g) * ub[side][g][p - off(s)].  A workgroup takes a 32x8 tile
// of positions and, side by side, stages the adjoints and orders of the cells
// within R of the tile in LDS.
template <int MORD, int NGP>
__global__ __launch_bounds__(256) void k_adj_gather(AdjGatherArgs a)
{
    constexpr int R = (MORD - 1) / 2, TX = kAdjTX, TY = kAdjTY;
    constexpr int QW = TX + 2 * R, QH = TY + 2 * R;
    constexpr int NO = (MORD + 1) / 2, SZ = stencil_sz(MORD);
    __shared__ double4 sU[NGP][QW * QH];
    __shared__ int8_t sO[QW * QH];
    __shared__ double sZ[NO][SZ][NGP];  // zT(o, side, s, g) of the current side
    const int N = a.N, E = N + 2 * R + 2, tid = threadIdx.x;
    const int tiles_x = (E + TX - 1) / TX;
    const int px0 = -(R + 1) + (blockIdx.x % tiles_x) * TX, py0 = -(R + 1) + (blockIdx.x / tiles_x) * TY;
    const int lx = tid % TX, ly = tid / TX;
    const int px = px0 + lx, py = py0 + ly;
    const bool valid = px <= N + R && py <= N + R;
    const int W = N + 2;
    const size_t plane = (size_t)W * W;
    // the adjoint is checked against autograd and finite differences, not bit
    // for bit, so its sums use fused multiply-adds
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int side = 0; side < 4; ++side) {
        __syncthreads();
        for (int idx = tid; idx < QW * QH; idx += 256) {
            const int ci = px0 - R + idx % QW, cj = py0 - R + idx / QW;
            const int o = side_order(a, side, ci, cj);
            sO[idx] = (int8_t)o;
            const size_t base = (size_t)side * NGP * plane + (size_t)(cj + 1) * W + (ci + 1);
#pragma unroll
            for (int g = 0; g < NGP; ++g) sU[g][idx] = o ? a.ub[base + (size_t)g * plane] : make_double4(0.0, 0.0, 0.0, 0.0);
        }
        for (int idx = tid; idx < NO * SZ * NGP; idx += 256) {
            const int g = idx % NGP, sidx = (idx / NGP) % SZ, oi = idx / (NGP * SZ);
            sZ[oi][sidx][g] = a.T->zT[oi][side][sidx][g];
        }
        __syncthreads();
        if (!valid) continue;
        for (int s = 0; s < SZ; ++s) {
            const int q = (ly + R - kSY[s]) * QW + (lx + R - kSX[s]);
            const int o = sO[q];
            if (o == 0 || s >= stencil_sz(o)) continue;
            const int oi = (o - 1) / 2;
#pragma unroll
            for (int g = 0; g < NGP; ++g) {
                const double z = sZ[oi][s][g];
                const double4 v = sU[g][q];
                acc[0] = fma(z, v.x, acc[0]);
                acc[1] = fma(z, v.y, acc[1]);
                acc[2] = fma(z, v.z, acc[2]);
                acc[3] = fma(z, v.w, acc[3]);
            }
        }
    }
    if (valid) a.G[(size_t)(py + R + 1) * E + (px + R + 1)] = make_double4(acc[0], acc[1], acc[2], acc[3]);
}

// out = alpha*Wn + fold(G) + gamma*B;  A = aA*A + aU*Wn (if A is given).
// Cells further than R+1 from every edge are not copied by any ghost cell.
template <int R>
__global__ __launch_bounds__(256) void k_adj_fold(AdjFoldArgs a)
{
    const int N = a.N, E = N + 2 * R + 2;
    const size_t n = (size_t)N * N;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int cx = (int)(c % N), cy = (int)(c / N);
    const double4 w = a.Wn[c];
    double v[4] = {a.alpha * w.x, a.alpha * w.y, a.alpha * w.z, a.alpha * w.w};
    const bool near_x = cx <= R || cx >= N - 1 - R, near_y = cy <= R || cy >= N - 1 - R;
    if (!near_x && !near_y) {
        const double4 gq = a.G[(size_t)(cy + R + 1) * E + (cx + R + 1)];
        v[0] += gq.x; v[1] += gq.y; v[2] += gq.z; v[3] += gq.w;
    } else {
        // positions of [-(R+1), N+R] whose ghost rule copies cx (cy): the cell
        // itself (e = 0) and the ghost candidates -k, N-1+k, k = 1..R+1
        for (int ey = 0; ey <= 2 * (R + 1); ++ey) {
            const int qy = ey == 0 ? cy : (ey <= R + 1 ? -ey : N - 1 + ey - (R + 1));
            bool fy = false;
            if (ey > 0 && bc_map(qy, N, bc_y(a.bc), fy) != cy) continue;
            for (int ex = 0; ex <= 2 * (R + 1); ++ex) {
                const int qx = ex == 0 ? cx : (ex <= R + 1 ? -ex : N - 1 + ex - (R + 1));
                bool fx = false;
                if (ex > 0 && bc_map(qx, N, bc_x(a.bc), fx) != cx) continue;
                const double4 gq = a.G[(size_t)(qy + R + 1) * E + (qx + R + 1)];
                v[0] += gq.x;
                v[1] += fx ? -gq.y : gq.y;
                v[2] += fy ? -gq.z : gq.z;
                v[3] += gq.w;
            }
        }
    }
    if (a.B) {
        const double4 b = a.B[c];
        v[0] += a.gamma * b.x; v[1] += a.gamma * b.y; v[2] += a.gamma * b.z; v[3] += a.gamma * b.w;
    }
    a.out[c] = make_double4(v[0], v[1], v[2], v[3]);
    if (a.A) {
        const double4 p = a.aA != 0.0 ? a.A[c] : make_double4(0.0, 0.0, 0.0, 0.0);
        a.A[c] = make_double4(a.aA * p.x + a.aU * w.x, a.aA * p.y + a.aU * w.y, a.aA * p.z + a.aU * w.z,
                              a.aA * p.w + a.aU * w.w);
    }
}

// Flux reverse mode on arrays of state pairs (component test entry).
__global__ void k_hllc_vjp(int dir, int n, const double *__restrict__ ul, const double *__restrict__ ur,
                           const double *__restrict__ Fb, double *__restrict__ ulb, double *__restrict__ urb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a[4], b[4], f[4], x[4], y[4];
    for (int k = 0; k < 4; ++k) { a[k] = ul[4 * i + k]; b[k] = ur[4 * i + k]; f[k] = Fb[4 * i + k]; }
    hllc_vjp(dir, a, b, f, x, y);
    for (int k = 0; k < 4; ++k) { ulb[4 * i + k] = x[k]; urb[4 * i + k] = y[k]; }
}

template <int MORD, int NGP>
static hipError_t adj_stage_mg(const AdjFaceArgs &fa, const AdjGatherArgs &ga, const AdjFoldArgs &fo, hipStream_t s)
{
    constexpr int R = (MORD - 1) / 2;
    const int N = fa.N;
    const int ftiles = ((N + kAdjTX) / kAdjTX) * ((N + kAdjTY) / kAdjTY);
    k_adj_face<MORD, NGP><<<ftiles, 256, 0, s>>>(fa);
    const int E = N + 2 * R + 2;
    const int gtiles = ((E + kAdjTX - 1) / kAdjTX) * ((E + kAdjTY - 1) / kAdjTY);
    k_adj_gather<MORD, NGP><<<gtiles, 256, 0, s>>>(ga);
    k_adj_fold<R><<<(unsigned)(((long long)N * N + 255) / 256), 256, 0, s>>>(fo);
    return hipGetLastError();
}

hipError_t launch_adj_stage(int mord, int ngp, const AdjFaceArgs &fa, const AdjGatherArgs &ga, const AdjFoldArgs &fo,
                            hipStream_t s)
{
#define GPM_ADJ_CASE(M, G) \
    if (mord == M && ngp == G) return adj_stage_mg<M, G>(fa, ga, fo, s);
    GPM_ADJ_CASE(3, 2)
    GPM_ADJ_CASE(5, 3)
    GPM_ADJ_CASE(7, 4)
#undef GPM_ADJ_CASE
    return hipErrorInvalidValue;
}

bool adj_supported(int mord, int ngp) { return (mord == 3 && ngp == 2) || (mord == 5 && ngp == 3) || (mord == 7 && ngp == 4); }

hipError_t launch_hllc_vjp(int dir, int n, const double *ul, const double *ur, const double *Fb, double *ulb,
                           double *urb, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_hllc_vjp<<<(n + 255) / 256, 256, 0, s>>>(dir, n, ul, ur, Fb, ulb, urb);
    return hipGetLastError();
}

}  // namespace gpm
