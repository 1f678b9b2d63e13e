// Host precompute of the GP-SE reconstruction tables in quad precision.
//
// Follows GP_presim_sphere (GP_init.f90:26-357), its SE kernels
// (Tpred_d2x_1D_fun :431-443, Tpred_1D_fun :446-455, intg_kernel_SE :459-511)
// and the Cholesky inverse (linalg.f90:5-95), evaluated in __float128 with
// libquadmath.  Mixed precision is kept where the reference has it: Gauss
// weights are FP64 expressions; Gauss points for 2 and 3 points are quad, for
// 4 and 5 points FP64 expressions widened to quad (GP_init.f90:45-83).
//
// The output tables use the reference's Fortran layouts:
//   zT_sp(Mord, sz_sphere, 4, ngp), GP_d2x_sp(Mord, sz), GP_d2y_sp(Mord, sz),
//   gauss_weight(5, 5), ixiy_sp(7, 25, 2).
#include <quadmath.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gpm_internal.h"

namespace gpm {

typedef __float128 qd;

static int stcl_sz(int ord) { return ord == 1 ? 1 : ord == 3 ? 5 : ord == 5 ? 13 : 25; }

namespace {

struct SE {
    qd l16, dx16, dy16;
    qd pi16, ospi;

    // Tpred_1D_fun, GP_init.f90:446-455
    qd tpred(qd delta, qd h) const {
        qd loh = l16 / h;
        qd expp = (delta + (qd)1 / 2) / (sqrtq((qd)2) * loh);
        qd expm = (delta - (qd)1 / 2) / (sqrtq((qd)2) * loh);
        return sqrtq(pi16 / 2) * loh * (erfq(expp) - erfq(expm));
    }
    // Tpred_d2x_1D_fun, GP_init.f90:431-443
    qd tpred_d2(qd delta, qd h) const {
        qd fm = delta - (qd)0.5, fp = delta + (qd)0.5;
        qd loh = l16 / h;
        qd expm = -(fm * fm) / (2 * (loh * loh));
        qd expp = -(fp * fp) / (2 * (loh * loh));
        return ((qd)1 / (l16 * l16)) * (fm * expq(expm) - fp * expq(expp));
    }
    // one direction of intg_kernel_SE, GP_init.f90:474-489: returns the
    // prefactor sqrt(pi)*lod**2 in *pre and the bracket as the result, so that
    // the caller multiplies in the reference's order r = r * pre * bracket.
    qd kfac(qd d, qd h, qd *pre) const {
        qd lod = l16 / h;
        qd s2l = sqrtq((qd)2) * lod;
        qd a11 = (d + 1) / s2l;
        qd a12 = (d - 1) / s2l;
        qd a21 = -(((d + 1) * (d + 1)) / (2 * (lod * lod)));
        qd a22 = -(((d - 1) * (d - 1)) / (2 * (lod * lod)));
        qd a31 = d / s2l;
        qd a32 = -(d * d) / (2 * (lod * lod));
        *pre = sqrtq(pi16) * (lod * lod);
        return a11 * erfq(a11) + a12 * erfq(a12) + ospi * (expq(a21) + expq(a22)) -
               2 * (a31 * erfq(a31) + ospi * expq(a32));
    }
    qd cov(qd ddx, qd ddy) const {
        qd pre, br;
        br = kfac(ddx, dx16, &pre);
        qd r = pre * br;
        br = kfac(ddy, dy16, &pre);
        return r * pre * br;
    }
};

// chol + invchol, linalg.f90:5-44,78-95 (reschol :46-76)
static void inv_chol(const std::vector<qd> &A, int n, std::vector<qd> &inv)
{
    std::vector<qd> L(n * n, 0);
#define AA(i, j) A[(i) + n * (j)]
#define LL(i, j) L[(i) + n * (j)]
    for (int i = 0; i < n; ++i) {
        qd S = 0;
        for (int k = 0; k < i; ++k) S = S + LL(i, k) * LL(i, k);
        LL(i, i) = sqrtq(AA(i, i) - S);
        for (int j = i + 1; j < n; ++j) {
            S = 0;
            for (int k = 0; k < i; ++k) S = S + LL(i, k) * LL(j, k);
            LL(j, i) = (AA(j, i) - S) / LL(i, i);
        }
    }
    inv.assign(n * n, 0);
    std::vector<qd> y(n), x(n);
    for (int kk = 0; kk < n; ++kk) {
        for (int i = 0; i < n; ++i) {
            qd S = 0;
            for (int j = 0; j < i; ++j) S = S + LL(i, j) * y[j];
            qd b = (i == kk) ? (qd)1 : (qd)0;
            y[i] = (b - S) / LL(i, i);
        }
        for (int i = n - 1; i >= 0; --i) {
            qd S = 0;
            for (int j = i + 1; j < n; ++j) S = S + LL(j, i) * x[j];
            x[i] = (y[i] - S) / LL(i, i);
        }
        for (int h = 0; h < n; ++h) inv[kk + n * h] = x[h];  // AM1(k,:) = x
    }
#undef AA
#undef LL
}

}  // namespace

// Sphere stencil offsets, GP_init.f90:97-176 (order of the entries matters).
const int kStencilX[25] = {0, -1, 0, 1, 0, -1, -1, 1, 1, 0, -2, 0, 2,
                           -3, -2, -1, 0, 1, 2, 3, 2, 1, 0, -1, -2};
const int kStencilY[25] = {0, 0, 1, 0, -1, -1, 1, 1, -1, -2, 0, 2, 0,
                           0, 1, 2, 3, 2, 1, 0, -1, -2, -3, -2, -1};

double quad_ratio(double L, int n) { return (double)((qd)L / n); }

int gp_tables(int mord, int ngp, double ell_o_dx, double lx, double ly, int lf, int nf,
              double *zT, double *d2x, double *d2y, double *gw_out, int32_t *ixiy)
{
    if (!(mord == 1 || mord == 3 || mord == 5 || mord == 7) || ngp < 1 || ngp > 5 || lf < 1 || nf < 1)
        return -1;
    const int szm = stcl_sz(mord);
    SE se;
    se.pi16 = 4 * atanq((qd)1);
    se.ospi = (qd)1 / sqrtq(se.pi16);
    se.dx16 = (qd)lx / lf;
    se.dy16 = (qd)ly / nf;
    se.l16 = (qd)ell_o_dx * (se.dx16 < se.dy16 ? se.dx16 : se.dy16);

    // Gauss weights (FP64) and points, GP_init.f90:45-83
    double gw[5][5] = {{0}};
    qd gpt[5][5];
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 5; ++b) gpt[a][b] = 0;
    const qd sq3 = sqrtq((qd)3), sq35 = sqrtq((qd)3 / 5);
    const double sq65 = std::sqrt(6. / 5), sq30 = std::sqrt(30.), third = 1. / 3, sq107 = std::sqrt(10. / 7),
                 sq70 = std::sqrt(70.);
    gw[0][0] = 1.0;
    gw[1][0] = 0.5; gw[1][1] = 0.5;
    gpt[1][0] = -(qd)0.5 / sq3; gpt[1][1] = (qd)0.5 / sq3;
    gw[2][0] = 0.5 * 5. / 9; gw[2][1] = 0.5 * 8. / 9; gw[2][2] = 0.5 * 5. / 9;
    gpt[2][0] = -(qd)0.5 * sq35; gpt[2][1] = 0; gpt[2][2] = (qd)0.5 * sq35;
    gw[3][0] = 0.5 * (1. / 36) * (18. - sq30);
    gw[3][1] = 0.5 * (1. / 36) * (18. + sq30);
    gw[3][2] = 0.5 * (1. / 36) * (18. + sq30);
    gw[3][3] = 0.5 * (1. / 36) * (18. - sq30);
    gpt[3][0] = -(0.5 * std::sqrt(3. / 7 + (2. / 7) * sq65));
    gpt[3][1] = -(0.5 * std::sqrt(3. / 7 - (2. / 7) * sq65));
    gpt[3][2] = 0.5 * std::sqrt(3. / 7 - (2. / 7) * sq65);
    gpt[3][3] = 0.5 * std::sqrt(3. / 7 + (2. / 7) * sq65);
    gw[4][0] = 0.5 * (322. - 13 * sq70) / 900;
    gw[4][1] = 0.5 * (322. + 13 * sq70) / 900;
    gw[4][2] = 0.5 * 128. / 225;
    gw[4][3] = 0.5 * (322. + 13 * sq70) / 900;
    gw[4][4] = 0.5 * (322. - 13 * sq70) / 900;
    gpt[4][0] = -(0.5 * third * std::sqrt(5. + 2 * sq107));
    gpt[4][1] = -(0.5 * third * std::sqrt(5. - 2 * sq107));
    gpt[4][2] = 0;
    gpt[4][3] = 0.5 * third * std::sqrt(5. - 2 * sq107);
    gpt[4][4] = 0.5 * third * std::sqrt(5. + 2 * sq107);
    for (int a = 0; a < 5; ++a)
        for (int b = 0; b < 5; ++b) gw_out[a + 5 * b] = gw[a][b];

    std::memset(zT, 0, sizeof(double) * mord * szm * 4 * ngp);
    std::memset(d2x, 0, sizeof(double) * mord * szm);
    std::memset(d2y, 0, sizeof(double) * mord * szm);
    std::memset(ixiy, 0, sizeof(int32_t) * 7 * 25 * 2);

    const int G = ngp - 1;
    for (int ord = 1; ord <= mord; ord += 2) {
        const int npt = stcl_sz(ord);
        for (int s = 0; s < npt; ++s) {
            ixiy[(ord - 1) + 7 * (s + 25 * 0)] = kStencilX[s];
            ixiy[(ord - 1) + 7 * (s + 25 * 1)] = kStencilY[s];
        }
        std::vector<qd> cov(npt * npt), inv;
        for (int k = 0; k < npt; ++k)
            for (int h = 0; h < npt; ++h)
                cov[k + npt * h] = se.cov((qd)(kStencilX[k] - kStencilX[h]), (qd)(kStencilY[k] - kStencilY[h]));
        // T_pred(k, face, gp) and the second-derivative vectors, GP_init.f90:194-244
        std::vector<qd> tp(npt * 4 * ngp), g2x(npt), g2y(npt);
        for (int k = 0; k < npt; ++k) {
            const qd ix = kStencilX[k], iy = kStencilY[k];
            for (int h = 0; h < ngp; ++h) {
                const qd g = gpt[G][h];
                qd dxs[4] = {ix - (-(qd)1 / 2), ix - g, ix - (qd)1 / 2, ix - g};
                qd dys[4] = {iy - g, iy - (qd)1 / 2, iy - g, iy - (-(qd)1 / 2)};
                for (int i = 0; i < 4; ++i) {
                    qd t = se.tpred(dxs[i], se.dx16);
                    t = t * se.tpred(dys[i], se.dy16);
                    tp[k + npt * (i + 4 * h)] = t;
                }
            }
            qd a = se.tpred_d2(ix, se.dx16);
            g2x[k] = a * se.tpred(iy, se.dy16);
            qd b = se.tpred_d2(iy, se.dy16);
            g2y[k] = b * se.tpred(ix, se.dx16);
        }
        inv_chol(cov, npt, inv);
        // zT = real(T_pred . C^-1, 8), then normalised by its FP64 sum (:254-265)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < ngp; ++j) {
                double z[25];
                for (int h = 0; h < npt; ++h) {
                    qd acc = 0;
                    for (int k = 0; k < npt; ++k) acc = acc + tp[k + npt * (i + 4 * j)] * inv[k + npt * h];
                    z[h] = (double)acc;
                }
                double sum = 0.0;
                for (int h = 0; h < npt; ++h) sum = sum + z[h];
                for (int h = 0; h < npt; ++h)
                    zT[(ord - 1) + mord * (h + szm * (i + 4 * j))] = z[h] / sum;
            }
        // GP_d2 = real(T_d2 . C^-1, 8) (:267-268)
        for (int h = 0; h < npt; ++h) {
            qd ax = 0, ay = 0;
            for (int k = 0; k < npt; ++k) {
                ax = ax + g2x[k] * inv[k + npt * h];
                ay = ay + g2y[k] * inv[k + npt * h];
            }
            d2x[(ord - 1) + mord * h] = (double)ax;
            d2y[(ord - 1) + mord * h] = (double)ay;
        }
    }
    return 0;
}

}  // namespace gpm
