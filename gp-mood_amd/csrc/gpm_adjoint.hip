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

// Reverse mode of hllc_dir: a y-face solve is the x-solve on momentum-swapped
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

// Reconstruction of cell (ci, cj) at face `face`, Gauss point g, order ord
// (dot_product order of mod_GP_MOOD.f90:46-65).
__device__ __forceinline__ void recon_point(const DevTables *__restrict__ T, const double4 *__restrict__ V, int N,
                                            int bc, int ord, int face, int g, int ci, int cj, double u[4])
{
    const int oi = (ord - 1) / 2, sz = stencil_sz(ord);
    u[0] = 0.0; u[1] = 0.0; u[2] = 0.0; u[3] = 0.0;
    for (int s = 0; s < sz; ++s) {
        const double4 q = load_bc(V, N, bc, ci + kSX[s], cj + kSY[s]);
        const double z = T->zT[oi][face][s][g];
        u[0] = u[0] + z * q.x;
        u[1] = u[1] + z * q.y;
        u[2] = u[2] + z * q.z;
        u[3] = u[3] + z * q.w;
    }
}

template <int MORD, int NGP>
__global__ __launch_bounds__(256) void k_adj_face(AdjFaceArgs a)
{
    const int N = a.N;
    const long long nfx = (long long)(N + 1) * N, nf = 2 * nfx;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * NGP) return;
    const int g = (int)(t / nf);
    const long long f = t % nf;
    const int dir = f < nfx ? 0 : 1;
    int ai, aj, bi, bj;
    if (dir == 0) {
        ai = (int)(f % (N + 1)) - 1; aj = (int)(f / (N + 1));
        bi = ai + 1; bj = aj;
    } else {
        const long long h = f - nfx;
        ai = (int)(h % N); aj = (int)(h / N) - 1;
        bi = ai; bj = aj + 1;
    }
    const int o = min(mask_at(a.mask, N, ai, aj, MORD), mask_at(a.mask, N, bi, bj, MORD));
    if (g == 0) (dir == 0 ? a.ordx : a.ordy)[f - (dir == 0 ? 0 : nfx)] = (int8_t)o;
    const int fa = dir == 0 ? kR : kT, fb = dir == 0 ? kL : kB;
    double ua[4], ub[4];
    recon_point(a.T, a.V, N, a.bc, o, fa, g, ai, aj, ua);
    recon_point(a.T, a.V, N, a.bc, o, fb, g, bi, bj, ub);
    const double4 wa = cell_or_zero(a.Wb, N, ai, aj), wb = cell_or_zero(a.Wb, N, bi, bj);
    const double c = a.alpha * (dir == 0 ? a.dtdx : a.dtdy) * a.T->gw[g];
    const double Fb[4] = {c * (wb.x - wa.x), c * (wb.y - wa.y), c * (wb.z - wa.z), c * (wb.w - wa.w)};
    double uab[4], ubb[4];
    hllc_vjp(dir, ua, ub, Fb, uab, ubb);
    const int W = N + 2;
    const size_t plane = (size_t)W * W;
    a.ub[((size_t)fa * NGP + g) * plane + (size_t)(aj + 1) * W + (ai + 1)] = make_double4(uab[0], uab[1], uab[2], uab[3]);
    a.ub[((size_t)fb * NGP + g) * plane + (size_t)(bj + 1) * W + (bi + 1)] = make_double4(ubb[0], ubb[1], ubb[2], ubb[3]);
}

template <int MORD, int NGP>
__global__ __launch_bounds__(256) void k_adj_gather(AdjGatherArgs a)
{
    constexpr int R = (MORD - 1) / 2;
    const int N = a.N, E = N + 2 * R + 2;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)E * E) return;
    const int px = (int)(t % E) - (R + 1), py = (int)(t / E) - (R + 1);
    const int W = N + 2;
    const size_t plane = (size_t)W * W;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < stencil_sz(MORD); ++s) {
        const int ci = px - kSX[s], cj = py - kSY[s];
        if (ci < -1 || ci > N || cj < -1 || cj > N) continue;
#pragma unroll
        for (int side = 0; side < 4; ++side) {
            int o;
            if (side == kR) {
                if (ci > N - 1 || cj < 0 || cj > N - 1) continue;
                o = a.ordx[(size_t)cj * (N + 1) + (ci + 1)];
            } else if (side == kL) {
                if (ci < 0 || cj < 0 || cj > N - 1) continue;
                o = a.ordx[(size_t)cj * (N + 1) + ci];
            } else if (side == kT) {
                if (cj > N - 1 || ci < 0 || ci > N - 1) continue;
                o = a.ordy[(size_t)(cj + 1) * N + ci];
            } else {
                if (cj < 0 || ci < 0 || ci > N - 1) continue;
                o = a.ordy[(size_t)cj * N + ci];
            }
            if (s >= stencil_sz(o)) continue;
            const int oi = (o - 1) / 2;
            const double4 *__restrict__ u = a.ub + (size_t)side * NGP * plane + (size_t)(cj + 1) * W + (ci + 1);
#pragma unroll
            for (int g = 0; g < NGP; ++g) {
                const double z = a.T->zT[oi][side][s][g];
                const double4 v = u[(size_t)g * plane];
                acc[0] += z * v.x;
                acc[1] += z * v.y;
                acc[2] += z * v.z;
                acc[3] += z * v.w;
            }
        }
    }
    a.G[t] = make_double4(acc[0], acc[1], acc[2], acc[3]);
}

// out = alpha*Wn + fold(G) + gamma*B;  A = aA*A + aU*Wn (if A is given)
template <int R>
__global__ __launch_bounds__(256) void k_adj_fold(AdjFoldArgs a)
{
    const int N = a.N, E = N + 2 * R + 2;
    const size_t n = (size_t)N * N;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int cx = (int)(c % N), cy = (int)(c / N);
    // preimages of cx (cy) under the ghost rule within [-(R+1), N+R]
    int px[2 * R + 3], py[2 * R + 3];
    bool fx[2 * R + 3], fy[2 * R + 3];
    int nx = 0, ny = 0;
    px[nx] = cx; fx[nx++] = false;
    py[ny] = cy; fy[ny++] = false;
    for (int k = 1; k <= R + 1; ++k) {
        const int cand[2] = {-k, N - 1 + k};
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            bool fl;
            if (bc_map(cand[e], N, a.bc, fl) == cx) { px[nx] = cand[e]; fx[nx++] = fl; }
            if (bc_map(cand[e], N, a.bc, fl) == cy) { py[ny] = cand[e]; fy[ny++] = fl; }
        }
    }
    const double4 w = a.Wn[c];
    double v[4] = {a.alpha * w.x, a.alpha * w.y, a.alpha * w.z, a.alpha * w.w};
    for (int jy = 0; jy < ny; ++jy)
        for (int jx = 0; jx < nx; ++jx) {
            const double4 gq = a.G[(size_t)(py[jy] + R + 1) * E + (px[jx] + R + 1)];
            v[0] += gq.x;
            v[1] += fx[jx] ? -gq.y : gq.y;
            v[2] += fy[jy] ? -gq.z : gq.z;
            v[3] += gq.w;
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
    const long long N = fa.N;
    const long long nface = 2 * (N + 1) * N * NGP;
    k_adj_face<MORD, NGP><<<(unsigned)((nface + 255) / 256), 256, 0, s>>>(fa);
    const long long E = N + 2 * R + 2;
    k_adj_gather<MORD, NGP><<<(unsigned)((E * E + 255) / 256), 256, 0, s>>>(ga);
    k_adj_fold<R><<<(unsigned)((N * N + 255) / 256), 256, 0, s>>>(fo);
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
