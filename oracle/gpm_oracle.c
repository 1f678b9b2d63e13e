/*
 * CPU restatement of the reference GP-MOOD 2D Euler step -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * this library, and only as the checker (or the timed CPU baseline), never as the
 * product path.  The product is the HIP library in gp-mood_amd/csrc.
 *
 * It restates, operation for operation, the FP64 arithmetic of the reference
 * Fortran (slug-cfd/gp-mood, gp-mood-code/) for the configuration
 *   method = GP_MOOD, integrator = SSP_RK3,
 *   numFlux in {HLLC, HLL, unsplit},
 *   BC_type in {Neumann, Periodic, Reflective, KH_BC}
 * and keeps the reference's MOOD bookkeeping (CellGPO, DetCell, DetFace_x/y,
 * stored face fluxes L_Flux_x/y) literally, so that every decision is taken on
 * the same data as in the reference.  Build with -ffp-contract=off (no FMA) so
 * that every product and sum is rounded as in the reference build.
 *
 * Layout: a state array holds (N+8)^2 cells of 4 doubles, cell (l,n) with
 * l,n in [-3, N+4] at index ((n+3)*(N+8) + (l+3))*4 + k, k = rho, mx, my, E:
 * byte-identical to the reference's U(4, lb:le, nb:ne) (global_variables.f90:21-26).
 *
 * Pinned against: the reference binary built by oracle/ref/build_ref.py (golden
 * fixtures under tests/golden/), bit for bit.  See the test suite for the list.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NGC 4

typedef struct {
    int N, mord, ngp, bc;      /* bc: 0 Neumann, 1 Periodic, 2 Reflective, 3 KH_BC */
    int bcx, bcy;              /* rule per axis: 0 Neumann, 1 Periodic, 2 Reflective */
    int flux;                  /* numFlux: 0 HLLC, 1 HLL, 2 unsplit */
    int dmp, u2, u2tol;        /* MOOD switches (problems/<name>.f90: DMP, U2, U2_tol) */
    int cascade;               /* 0: reference (>3 -> 3, 3 -> 1); 1: full (7->5->3->1) */
    double cfl, dx, dy;
    /* Tables in the reference's Fortran layouts (global_variables.f90:37-38,53,61):
       zT_sp(Mord, sz_sphere, 4, ngp), GP_d2x_sp(Mord, sz_sphere),
       GP_d2y_sp(Mord, sz_sphere), gauss_weight(5,5), ixiy_sp(7,25,2). */
    double *zT, *d2x, *d2y, gw[25];
    int ixiy[7 * 25 * 2];
    int sz;                    /* sz_sphere = stencil size of Mord */
    /* work arrays */
    int W;                     /* N + 8 */
    double *Uh;                /* Uh(4, 4, ngp, 0:N+1, 0:N+1) */
    double *Fx, *Fy;           /* L_Flux_x(4,-1:N+1,0:N+1), L_Flux_y(4,0:N+1,-1:N+1) */
    int8_t *gpo;               /* CellGPO(lb:le, nb:ne) */
    uint8_t *detcell;          /* DetCell(lb:le, nb:ne) */
    uint8_t *detfx, *detfy;    /* DetFace_x(-1:N+1, 0:N+1), DetFace_y(0:N+1, -1:N+1) */
    uint8_t *decrease;         /* decrease(1:N, 1:N) */
    double *U1, *U2;           /* RK stage states */
    long count_detected;       /* count_detected_cell (global_variables.f90:63) */
    int8_t *init_orders;       /* a priori CellGPO of the interior [N][N] (NULL: Mord) */
    int passes[3];             /* MOOD passes per stage of the last step */
} orc_t;

static const double GAMMA = 1.4;

/* Fortran MIN/MAX as compiled by the reference build: the second operand is
   returned unless the first compares strictly smaller (larger). */
static inline double fmin2(double a, double b) { return (a < b) ? a : b; }
static inline double fmax2(double a, double b) { return (a > b) ? a : b; }

static inline int stcl_sz(int ord) /* GP_init.f90:14-23 */
{
    return ord == 1 ? 1 : ord == 3 ? 5 : ord == 5 ? 13 : 25;
}

#define CELL(o, l, n) ((size_t)(((n) + 3) * (o)->W + ((l) + 3)))
#define ZT(o, ord, s, i, j) ((o)->zT[((ord) - 1) + (o)->mord * ((s) + (o)->sz * ((i) + 4 * (j)))])
#define D2X(o, ord, s) ((o)->d2x[((ord) - 1) + (o)->mord * (s)])
#define D2Y(o, ord, s) ((o)->d2y[((ord) - 1) + (o)->mord * (s)])
#define IXIY(o, ord, s, d) ((o)->ixiy[((ord) - 1) + 7 * ((s) + 25 * (d))])
#define GW(o, ngp, j) ((o)->gw[((ngp) - 1) + 5 * (j)])
#define FXI(o, l, n) ((size_t)(((n) * ((o)->N + 3) + ((l) + 1))))
#define FYI(o, l, n) ((size_t)((((n) + 1) * ((o)->N + 2) + (l))))
#define DFX(o, l, n) ((o)->detfx[FXI(o, l, n)])
#define DFY(o, l, n) ((o)->detfy[FYI(o, l, n)])
#define UHI(o, k, i, j, l, n) ((size_t)(((((size_t)(n) * ((o)->N + 2) + (l)) * 4 + (i)) * (o)->ngp + (j)) * 4 + (k)))

enum { iL = 0, iT = 1, iR = 2, iB = 3 };

orc_t *orc_create(int N, int mord, int ngp, int bc, int dmp, int u2, int u2tol, int cascade,
                  double cfl, double dx, double dy, const double *zT, const double *d2x,
                  const double *d2y, const double *gw, const int *ixiy)
{
    orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
    o->N = N; o->mord = mord; o->ngp = ngp; o->bc = bc;
    /* KH_BC (BC.f90:383-406): periodic in x, Neumann in y */
    o->bcx = bc == 3 ? 1 : bc;
    o->bcy = bc == 3 ? 0 : bc;
    o->flux = 0;
    o->dmp = dmp; o->u2 = u2; o->u2tol = u2tol; o->cascade = cascade;
    o->cfl = cfl; o->dx = dx; o->dy = dy;
    o->sz = stcl_sz(mord);
    o->zT = (double *)malloc(sizeof(double) * mord * o->sz * 4 * ngp);
    memcpy(o->zT, zT, sizeof(double) * mord * o->sz * 4 * ngp);
    o->d2x = (double *)malloc(sizeof(double) * mord * o->sz);
    o->d2y = (double *)malloc(sizeof(double) * mord * o->sz);
    memcpy(o->d2x, d2x, sizeof(double) * mord * o->sz);
    memcpy(o->d2y, d2y, sizeof(double) * mord * o->sz);
    memcpy(o->gw, gw, sizeof(double) * 25);
    memcpy(o->ixiy, ixiy, sizeof(int) * 7 * 25 * 2);
    o->W = N + 2 * NGC;
    size_t ncell = (size_t)o->W * o->W;
    o->Uh = (double *)calloc((size_t)(N + 2) * (N + 2) * 16 * ngp, sizeof(double));
    o->Fx = (double *)calloc((size_t)(N + 3) * (N + 2) * 4, sizeof(double));
    o->Fy = (double *)calloc((size_t)(N + 2) * (N + 3) * 4, sizeof(double));
    o->gpo = (int8_t *)malloc(ncell);
    o->detcell = (uint8_t *)malloc(ncell);
    o->detfx = (uint8_t *)malloc((size_t)(N + 3) * (N + 2));
    o->detfy = (uint8_t *)malloc((size_t)(N + 2) * (N + 3));
    o->decrease = (uint8_t *)malloc((size_t)N * N);
    o->U1 = (double *)calloc(ncell * 4, sizeof(double));
    o->U2 = (double *)calloc(ncell * 4, sizeof(double));
    return o;
}

void orc_destroy(orc_t *o)
{
    if (!o) return;
    free(o->zT); free(o->d2x); free(o->d2y); free(o->Uh); free(o->Fx); free(o->Fy);
    free(o->gpo); free(o->detcell); free(o->detfx); free(o->detfy); free(o->decrease);
    free(o->U1); free(o->U2); free(o->init_orders); free(o);
}

/* Boundary_C, BC.f90:18-59 (Neumann, Periodic), :89-135 (Reflective) and
   :383-406 (KH_BC): x-ghosts for every row first, then y-ghosts for every
   column. */
void orc_bc(const orc_t *o, double *U)
{
    const int N = o->N;
    for (int n = 1 - NGC; n <= N + NGC; ++n) {
        for (int g = 1; g <= NGC; ++g) {
            int lr = N + g, ll = 1 - g, sr, sl;
            double sgn = 1.0;
            if (o->bcx == 0) { sr = N; sl = 1; }
            else if (o->bcx == 1) { sr = g; sl = N - g + 1; }
            else { sr = N - g + 1; sl = g; sgn = -1.0; }
            for (int k = 0; k < 4; ++k) {
                U[CELL(o, lr, n) * 4 + k] = U[CELL(o, sr, n) * 4 + k];
                U[CELL(o, ll, n) * 4 + k] = U[CELL(o, sl, n) * 4 + k];
            }
            if (sgn < 0) {
                U[CELL(o, lr, n) * 4 + 1] = -U[CELL(o, lr, n) * 4 + 1];
                U[CELL(o, ll, n) * 4 + 1] = -U[CELL(o, ll, n) * 4 + 1];
            }
        }
    }
    for (int l = 1 - NGC; l <= N + NGC; ++l) {
        for (int g = 1; g <= NGC; ++g) {
            int nt = N + g, nb = 1 - g, st, sb;
            double sgn = 1.0;
            if (o->bcy == 0) { st = N; sb = 1; }
            else if (o->bcy == 1) { st = g; sb = N - g + 1; }
            else { st = N - g + 1; sb = g; sgn = -1.0; }
            for (int k = 0; k < 4; ++k) {
                U[CELL(o, l, nt) * 4 + k] = U[CELL(o, l, st) * 4 + k];
                U[CELL(o, l, nb) * 4 + k] = U[CELL(o, l, sb) * 4 + k];
            }
            if (sgn < 0) {
                U[CELL(o, l, nt) * 4 + 2] = -U[CELL(o, l, nt) * 4 + 2];
                U[CELL(o, l, nb) * 4 + 2] = -U[CELL(o, l, nb) * 4 + 2];
            }
        }
    }
}

/* pressure, physics.f90:132-143 */
static inline double pressure(const double *u)
{
    return (GAMMA - 1.0) * (u[3] - (0.5 * (u[1] * u[1] + u[2] * u[2])) / u[0]);
}

/* Setdt, physics.f90:78-129 (dt_reduction is .false.): min over every cell
   including ghosts, then *CFL, then clipped to dtfinal. */
double orc_setdt(const orc_t *o, const double *U, double dtfinal)
{
    const int N = o->N;
    double dt = 1.0e30;
#pragma omp parallel for reduction(min : dt) schedule(static)
    for (int n = 1 - NGC; n <= N + NGC; ++n) {
        double dtl = 1.0e30;
        for (int l = 1 - NGC; l <= N + NGC; ++l) {
            const double *u = U + CELL(o, l, n) * 4;
            double vx = u[1] / u[0];
            double vy = u[2] / u[0];
            double pres = pressure(u);
            double c = sqrt(GAMMA * pres / u[0]);
            double v = 1.0 / (1.0 / (o->dx / (fabs(vx) + c)) + 1.0 / (o->dy / (fabs(vy) + c)));
            dtl = fmin2(dtl, v);
        }
        if (dtl < dt) dt = dtl;
    }
    dt = dt * o->cfl;
    if (dt > dtfinal) dt = dtfinal;
    return dt;
}

/* c2p_local (physics.f90:51-75) + Flux (physics.f90:9-48) + HLLC_Flux
   (mod_HLLC.f90:11-169).  dir 0 = x, 1 = y. */
static void hllc(const double *ul, const double *ur, int dir, double *F)
{
    int adm = 1;
    const int md = dir == 0 ? 1 : 2, om = dir == 0 ? 2 : 1;
    double vxl = ul[1] / ul[0], vyl = ul[2] / ul[0];
    double pl = pressure(ul);
    double cl = sqrt(GAMMA * pl / ul[0]);
    if (ul[0] <= 0. || pl <= 0. || isnan(ul[0]) || isnan(pl)) adm = 0;
    double vxr = ur[1] / ur[0], vyr = ur[2] / ur[0];
    double pr = pressure(ur);
    double cr = sqrt(GAMMA * pr / ur[0]);
    if (ur[0] <= 0. || pr <= 0. || isnan(ur[0]) || isnan(pr)) adm = 0;
    if (!adm) {
        /* mod_HLLC.f90:166: F = 1/(vx-vx) with vx of the right state's last c2p */
        double vx = dir == 0 ? vxr : vxr; /* c2p returns vx in both directions */
        double z = vx - vx;
        for (int k = 0; k < 4; ++k) F[k] = 1.0 / z;
        return;
    }
    double vnl = dir == 0 ? vxl : vyl, vtl = dir == 0 ? vyl : vxl;
    double vnr = dir == 0 ? vxr : vyr, vtr = dir == 0 ? vyr : vxr;
    double Fl[4], Fr[4];
    if (dir == 0) {
        Fl[0] = ul[1]; Fl[1] = ul[0] * (vxl * vxl) + pl; Fl[2] = ul[2] * vxl; Fl[3] = vxl * (ul[3] + pl);
        Fr[0] = ur[1]; Fr[1] = ur[0] * (vxr * vxr) + pr; Fr[2] = ur[2] * vxr; Fr[3] = vxr * (ur[3] + pr);
    } else {
        Fl[0] = ul[2]; Fl[1] = ul[1] * vyl; Fl[2] = ul[0] * (vyl * vyl) + pl; Fl[3] = vyl * (ul[3] + pl);
        Fr[0] = ur[2]; Fr[1] = ur[1] * vyr; Fr[2] = ur[0] * (vyr * vyr) + pr; Fr[3] = vyr * (ur[3] + pr);
    }
    double SL = fmin2(vnl - cl, vnr - cr);
    double SR = fmax2(vnl + cl, vnr + cr);
    double numerL = SL - vnl, numerR = SR - vnr;
    double uStar = ul[0] * vnl * numerL - ur[0] * vnr * numerR;
    uStar = (pr - pl) + uStar;
    uStar = uStar / (ul[0] * numerL - ur[0] * numerR);
    double denomL = SL - uStar, denomR = SR - uStar;
    double dStarL = ul[0] * numerL / denomL;
    double dStarR = ur[0] * numerR / denomR;
    double UsL[4], UsR[4];
    UsL[0] = 1.0; UsL[md] = uStar; UsL[om] = vtl;
    UsL[3] = ul[3] / ul[0] + (uStar - vnl) * (uStar + pl / (ul[0] * numerL));
    UsR[0] = 1.0; UsR[md] = uStar; UsR[om] = vtr;
    UsR[3] = ur[3] / ur[0] + (uStar - vnr) * (uStar + pr / (ur[0] * numerR));
    for (int k = 0; k < 4; ++k) { UsL[k] = UsL[k] * dStarL; UsR[k] = UsR[k] * dStarR; }
    if (SL >= 0.) {
        for (int k = 0; k < 4; ++k) F[k] = Fl[k];
    } else if (SL < 0. && uStar >= 0.) {
        for (int k = 0; k < 4; ++k) F[k] = Fl[k] + SL * (UsL[k] - ul[k]);
    } else if (uStar < 0. && SR >= 0.) {
        for (int k = 0; k < 4; ++k) F[k] = Fr[k] + SR * (UsR[k] - ur[k]);
    } else {
        for (int k = 0; k < 4; ++k) F[k] = Fr[k];
    }
}

/* HLL_Flux, mod_HLL.f90:11-166, with c2p_local and Flux (physics.f90:9-75). */
static void hll(const double *ul, const double *ur, int dir, double *F)
{
    int adm = 1;
    double vxl = ul[1] / ul[0], vyl = ul[2] / ul[0];
    double pl = pressure(ul);
    double cl = sqrt(GAMMA * pl / ul[0]);
    if (ul[0] <= 0. || pl <= 0. || isnan(ul[0]) || isnan(pl)) adm = 0;
    double vxr = ur[1] / ur[0], vyr = ur[2] / ur[0];
    double pr = pressure(ur);
    double cr = sqrt(GAMMA * pr / ur[0]);
    if (ur[0] <= 0. || pr <= 0. || isnan(ur[0]) || isnan(pr)) adm = 0;
    if (!adm) {
        /* mod_HLL.f90:162: F = 1/(vx-vx), vx of the right state's c2p */
        double z = vxr - vxr;
        for (int k = 0; k < 4; ++k) F[k] = 1.0 / z;
        return;
    }
    const double vnl = dir == 0 ? vxl : vyl, vnr = dir == 0 ? vxr : vyr;
    double Fl[4], Fr[4];
    if (dir == 0) {
        Fl[0] = ul[1]; Fl[1] = ul[0] * (vxl * vxl) + pl; Fl[2] = ul[2] * vxl; Fl[3] = vxl * (ul[3] + pl);
        Fr[0] = ur[1]; Fr[1] = ur[0] * (vxr * vxr) + pr; Fr[2] = ur[2] * vxr; Fr[3] = vxr * (ur[3] + pr);
    } else {
        Fl[0] = ul[2]; Fl[1] = ul[1] * vyl; Fl[2] = ul[0] * (vyl * vyl) + pl; Fl[3] = vyl * (ul[3] + pl);
        Fr[0] = ur[2]; Fr[1] = ur[1] * vyr; Fr[2] = ur[0] * (vyr * vyr) + pr; Fr[3] = vyr * (ur[3] + pr);
    }
    const double SL = fmin2(vnl - cl, vnr - cr);
    const double SR = fmax2(vnl + cl, vnr + cr);
    if (SL > 0.) {
        for (int k = 0; k < 4; ++k) F[k] = Fl[k];
    } else if (SR < 0.) {
        for (int k = 0; k < 4; ++k) F[k] = Fr[k];
    } else {
        for (int k = 0; k < 4; ++k) F[k] = (SR * Fl[k] - SL * Fr[k] + SR * SL * (ur[k] - ul[k])) / (SR - SL);
    }
}

/* unsplit_flux, mod_unsplit.f90:11-71, with c2p_local (physics.f90:51-75).
   The reference passes one variable for both vx2 and vy2 of c2p_local; it
   ends up holding the right state's vy**2, which the inadmissible branch uses. */
static void unsplit(const double *ul, const double *ur, int dir, double *F)
{
    int adm = 1;
    double vxl = ul[1] / ul[0], vyl = ul[2] / ul[0];
    double pl = pressure(ul);
    double cl = sqrt(GAMMA * pl / ul[0]);
    if (ul[0] <= 0. || pl <= 0. || isnan(ul[0]) || isnan(pl)) adm = 0;
    double vxr = ur[1] / ur[0], vyr = ur[2] / ur[0];
    double no = vyr * vyr;
    double pr = pressure(ur);
    double cr = sqrt(GAMMA * pr / ur[0]);
    if (ur[0] <= 0. || pr <= 0. || isnan(ur[0]) || isnan(pr)) adm = 0;
    const double a = 1.1 * fmax2(ul[0] * cl, ur[0] * cr);
    const double vnl = dir == 0 ? vxl : vyl, vnr = dir == 0 ? vxr : vyr;
    const double vtl = dir == 0 ? vyl : vxl, vtr = dir == 0 ? vyr : vxr;
    const double pistar = 0.5 * (pl + pr) - 0.5 * a * (vnr - vnl);
    const double ustar = 0.5 * (vnl + vnr) - 0.5 * (pr - pl) / a;
    const double trans = -(0.5 * a * (vtr - vtl));
    if (!adm) {
        double z = no - no;
        for (int k = 0; k < 4; ++k) F[k] = 1.0 / z;
        return;
    }
    const double *up = ustar > 0 ? ul : ur;
    for (int k = 0; k < 4; ++k) F[k] = up[k] * ustar;
    const int md = dir == 0 ? 1 : 2, om = dir == 0 ? 2 : 1;
    F[md] = F[md] + pistar;
    F[om] = F[om] + trans;
    F[3] = F[3] + pistar * ustar;
}

static void riemann(int flux, const double *ul, const double *ur, int dir, double *F)
{
    if (flux == 1) hll(ul, ur, dir, F);
    else if (flux == 2) unsplit(ul, ur, dir, F);
    else hllc(ul, ur, dir, F);
}

/* GP_MOOD_, mod_GP_MOOD.f90:11-76: face states at Gauss points for cells
   l,n in [0, N+1] whose face flag is set. */
static void gp_mood_recons(orc_t *o, const double *Uin)
{
    const int N = o->N, mord = o->mord, ngp = o->ngp, sz = o->sz;
#pragma omp parallel for schedule(static)
    for (int n = 0; n <= N + 1; ++n) {
        double q[4][25];
        for (int l = 0; l <= N + 1; ++l) {
            int done = 0;
            for (int i = iL; i <= iB; ++i) {
                int rec = 0, ord = 0;
                int me = o->gpo[CELL(o, l, n)];
                if (i == iL && DFX(o, l - 1, n)) { ord = (int)fmin2(o->gpo[CELL(o, l - 1, n)], me); rec = 1; }
                else if (i == iR && DFX(o, l, n)) { ord = (int)fmin2(o->gpo[CELL(o, l + 1, n)], me); rec = 1; }
                else if (i == iB && DFY(o, l, n - 1)) { ord = (int)fmin2(o->gpo[CELL(o, l, n - 1)], me); rec = 1; }
                else if (i == iT && DFY(o, l, n)) { ord = (int)fmin2(o->gpo[CELL(o, l, n + 1)], me); rec = 1; }
                if (!rec) continue;
                if (!done) {
                    for (int s = 0; s < sz; ++s) {
                        const double *u = Uin + CELL(o, l + IXIY(o, mord, s, 0), n + IXIY(o, mord, s, 1)) * 4;
                        for (int k = 0; k < 4; ++k) q[k][s] = u[k];
                    }
                    done = 1;
                }
                int ssz = stcl_sz(ord);
                for (int j = 0; j < ngp; ++j) {
                    for (int k = 0; k < 4; ++k) {
                        double acc = 0.0;
                        for (int s = 0; s < ssz; ++s) acc = acc + ZT(o, ord, s, i, j) * q[k][s];
                        o->Uh[UHI(o, k, i, j, l, n)] = acc;
                    }
                }
            }
        }
    }
}

/* shock_detector, mod_subroutine_MOOD.f90:191-238 */
static int shock_detector(const orc_t *o, const double *Uin, int l, int n)
{
    const double dx = o->dx, dy = o->dy;
#define UI(k, a, b) Uin[CELL(o, a, b) * 4 + (k)]
    double divV = (UI(1, l + 1, n) - UI(1, l - 1, n)) / dx + (UI(2, l, n + 1) - UI(2, l, n - 1)) / dy;
    divV = 0.5 * divV / UI(0, l, n);
    divV = divV - 0.5 * (UI(1, l, n) * (UI(0, l + 1, n) - UI(0, l - 1, n)) / dx +
                         UI(2, l, n) * (UI(0, l, n + 1) - UI(0, l, n - 1)) / dy) /
                      (UI(0, l, n) * UI(0, l, n));
    double Ca = 1.4 * pressure(Uin + CELL(o, l, n) * 4) / UI(0, l, n);
    double Mach = (UI(1, l, n) * UI(1, l, n) + UI(2, l, n) * UI(2, l, n)) / (UI(0, l, n) * UI(0, l, n));
    Mach = sqrt(Mach / Ca);
    double pL = pressure(Uin + CELL(o, l - 1, n) * 4);
    double pR = pressure(Uin + CELL(o, l + 1, n) * 4);
    double pB = pressure(Uin + CELL(o, l, n - 1) * 4);
    double pT = pressure(Uin + CELL(o, l, n + 1) * 4);
    double gradP = 0.5 * (fabs(pR - pL) / fmin2(pL, pR) / dx + fabs(pT - pB) / fmin2(pB, pT) / dy);
    double m = fmax2(dx, dy);
    return (divV < -(m * m) && Mach > 0.2 && gradP > 5.0);
#undef UI
}

/* DETECTION loop-1 decision for one cell, mod_subroutine_MOOD.f90:41-131 */
static int decide(const orc_t *o, const double *Uin, const double *Uout, int l, int n)
{
    const double *uo = Uout + CELL(o, l, n) * 4;
    if (uo[0] <= 0. || isnan(uo[0])) return 1;
    double p = pressure(uo);
    if (p <= 0. || isnan(p)) return 1;
    int shock = shock_detector(o, Uin, l, n);
    if (!(o->dmp && shock)) return 0;
#define R(a, b) Uin[CELL(o, a, b) * 4]
    double UB = R(l, n - 1), UL = R(l - 1, n), UM = R(l, n), UR = R(l + 1, n), UT = R(l, n + 1);
    double mn = fmin2(fmin2(fmin2(fmin2(UM, UL), UB), UT), UR);
    double mx = fmax2(fmax2(fmax2(fmax2(UM, UL), UB), UT), UR);
    double Mm = mx - mn;
    int dec = 0;
    if (Mm > o->dx * o->dx * o->dx) {
        if (uo[0] > mx || uo[0] < mn) {
            dec = 1;
            if (o->u2) {
                double D2x[3][3], D2y[3][3];
                for (int j = -1; j <= 1; ++j)
                    for (int i = -1; i <= 1; ++i) {
                        if (!(i == 0 || j == 0)) continue;
                        double q[5];
                        for (int k = 0; k < 5; ++k) q[k] = R(l + IXIY(o, 3, k, 0) + i, n + IXIY(o, 3, k, 1) + j);
                        double ax = 0.0, ay = 0.0;
                        for (int k = 0; k < 5; ++k) ax = ax + D2X(o, 3, k) * q[k];
                        for (int k = 0; k < 5; ++k) ay = ay + D2Y(o, 3, k) * q[k];
                        D2x[i + 1][j + 1] = ax;
                        D2y[i + 1][j + 1] = ay;
                    }
                double Xmin = fmin2(fmin2(fmin2(fmin2(D2x[0][1], D2x[1][1]), D2x[2][1]), D2x[1][0]), D2x[1][2]);
                double Xmax = fmax2(fmax2(fmax2(fmax2(D2x[0][1], D2x[1][1]), D2x[2][1]), D2x[1][0]), D2x[1][2]);
                double Ymin = fmin2(fmin2(fmin2(fmin2(D2y[0][1], D2y[1][1]), D2y[2][1]), D2y[1][0]), D2y[1][2]);
                double Ymax = fmax2(fmax2(fmax2(fmax2(D2y[0][1], D2y[1][1]), D2y[2][1]), D2y[1][0]), D2y[1][2]);
                double delta = 0.0;
                if (o->u2tol) delta = fmax2(o->dx, o->dy);
                if ((Xmin * Xmax > -delta) && ((fmax2(fabs(Xmin), fabs(Xmax)) < delta) || (fabs(Xmin / Xmax) >= 0.5)))
                    dec = 0;
                else if ((Ymin * Ymax > -delta) && ((fmax2(fabs(Ymin), fabs(Ymax)) < delta) || (fabs(Ymin / Ymax) >= 0.5)))
                    dec = 0;
            }
        }
    }
#undef R
    return dec;
}

static inline int lower_order(int ord, int cascade) /* mod_subroutine_MOOD.f90:159-163 */
{
    if (cascade == 0) return ord == 3 ? 1 : (ord > 3 ? 3 : ord);
    return ord > 1 ? ord - 2 : ord;
}

/* DETECTION, mod_subroutine_MOOD.f90:13-189.  Returns 1 when MOOD is finished. */
static int detection(orc_t *o, const double *Uin, double *Uout)
{
    const int N = o->N;
    orc_bc(o, Uout);
    memset(o->decrease, 0, (size_t)N * N);
#pragma omp parallel for schedule(dynamic, 4)
    for (int n = 1; n <= N; ++n)
        for (int l = 1; l <= N; ++l)
            if (o->detcell[CELL(o, l, n)] && o->gpo[CELL(o, l, n)] > 1)
                o->decrease[(size_t)(n - 1) * N + (l - 1)] = (uint8_t)decide(o, Uin, Uout, l, n);
    /* clearing of passing cells' flags (:133-141); idempotent, order free */
    for (int n = 1; n <= N; ++n)
        for (int l = 1; l <= N; ++l)
            if (!o->decrease[(size_t)(n - 1) * N + (l - 1)]) {
                o->detcell[CELL(o, l, n)] = 0;
                DFY(o, l, n - 1) = 0; DFY(o, l, n) = 0;
                DFX(o, l - 1, n) = 0; DFX(o, l, n) = 0;
            }
    int finished = 1;
    for (int n = 1; n <= N; ++n)
        for (int l = 1; l <= N; ++l) {
            if (!o->decrease[(size_t)(n - 1) * N + (l - 1)]) continue;
            o->count_detected++;
            finished = 0;
            int8_t *g = &o->gpo[CELL(o, l, n)];
            *g = (int8_t)lower_order(*g, o->cascade);
            const int nb[5][2] = {{l, n - 1}, {l - 1, n}, {l, n}, {l + 1, n}, {l, n + 1}};
            for (int t = 0; t < 5; ++t) o->detcell[CELL(o, nb[t][0], nb[t][1])] = 1;
            DFY(o, l, n - 1) = 1; DFY(o, l, n) = 1; DFX(o, l, n) = 1; DFX(o, l - 1, n) = 1;
            for (int t = 0; t < 5; ++t)
                for (int k = 0; k < 4; ++k)
                    Uout[CELL(o, nb[t][0], nb[t][1]) * 4 + k] = Uin[CELL(o, nb[t][0], nb[t][1]) * 4 + k];
        }
    return finished;
}

/* One flux at a face from the stored Gauss-point face states, quadrature
   L_Flux = sum_j F_j * gauss_weight(ngp, j) (mod_FE.f90:80-98). */
static void face_flux(const orc_t *o, int dir, int l, int n, double *Lf)
{
    const int ngp = o->ngp;
    double Fg[5][4];
    for (int j = 0; j < ngp; ++j) {
        double ul[4], ur[4];
        for (int k = 0; k < 4; ++k) {
            if (dir == 0) { ul[k] = o->Uh[UHI(o, k, iR, j, l, n)]; ur[k] = o->Uh[UHI(o, k, iL, j, l + 1, n)]; }
            else { ul[k] = o->Uh[UHI(o, k, iT, j, l, n)]; ur[k] = o->Uh[UHI(o, k, iB, j, l, n + 1)]; }
        }
        riemann(o->flux, ul, ur, dir, Fg[j]);
    }
    for (int k = 0; k < 4; ++k) {
        double acc = 0.0;
        for (int j = 0; j < ngp; ++j) acc = acc + Fg[j][k] * GW(o, ngp, j);
        Lf[k] = acc;
    }
}

/* Forward_Euler, mod_FE.f90:20-190 (GP_MOOD, HLLC, no NN).  The face loop
   (:75-145) is evaluated as: every flagged face flux first, then for every
   interior cell with DetCell set the four updates in the order in which the
   reference's n-outer/l-inner loop applies them:
     y-face (l,n-1) [+], x-face (l-1,n) [+], x-face (l,n) [-], y-face (l,n) [-].
   Interior cells without DetCell are not touched by a re-pass (the reference
   updates them only through flagged faces, and a flagged face always has both
   neighbours flagged).  Ghost cells of Uout are overwritten by Boundary_C at the
   start of DETECTION, so their face-loop updates are not formed. */
void orc_fe(orc_t *o, double *Uin, double *Uout, double dt, int *npass)
{
    const int N = o->N;
    const size_t ncell = (size_t)o->W * o->W;
    for (size_t c = 0; c < ncell; ++c) { o->gpo[c] = (int8_t)o->mord; o->detcell[c] = 1; }
    /* a priori orders (compute_CellGPO_with_NN, mod_FE.f90:51-56): interior only */
    if (o->init_orders)
        for (int n = 1; n <= N; ++n)
            for (int l = 1; l <= N; ++l) o->gpo[CELL(o, l, n)] = o->init_orders[(size_t)(n - 1) * N + (l - 1)];
    memset(o->detfx, 1, (size_t)(N + 3) * (N + 2));
    memset(o->detfy, 1, (size_t)(N + 2) * (N + 3));
    orc_bc(o, Uin);
    memcpy(Uout, Uin, ncell * 4 * sizeof(double));
    const double dtdx = dt / o->dx, dtdy = dt / o->dy;
    int passes = 0;
    for (;;) {
        ++passes;
        gp_mood_recons(o, Uin);
#pragma omp parallel for schedule(static)
        for (int n = 0; n <= N; ++n)
            for (int l = 0; l <= N; ++l) {
                if (n >= 1 && DFX(o, l, n)) face_flux(o, 0, l, n, o->Fx + FXI(o, l, n) * 4);
                if (l >= 1 && DFY(o, l, n)) face_flux(o, 1, l, n, o->Fy + FYI(o, l, n) * 4);
            }
#pragma omp parallel for schedule(static)
        for (int n = 1; n <= N; ++n)
            for (int l = 1; l <= N; ++l) {
                if (!o->detcell[CELL(o, l, n)]) continue;
                double *u = Uout + CELL(o, l, n) * 4;
                const double *fyb = o->Fy + FYI(o, l, n - 1) * 4, *fxl = o->Fx + FXI(o, l - 1, n) * 4;
                const double *fxr = o->Fx + FXI(o, l, n) * 4, *fyt = o->Fy + FYI(o, l, n) * 4;
                for (int k = 0; k < 4; ++k) {
                    double v = u[k];
                    v = v + dtdy * fyb[k];
                    v = v + dtdx * fxl[k];
                    v = v - dtdx * fxr[k];
                    v = v - dtdy * fyt[k];
                    u[k] = v;
                }
            }
        if (detection(o, Uin, Uout)) break;
    }
    if (npass) *npass = passes;
}

/* SSP_RK3_, mod_SSP_RK3.f90:12-39 */
void orc_ssprk3(orc_t *o, double *U, double *Ur, double dt)
{
    const size_t n4 = (size_t)o->W * o->W * 4;
    o->count_detected = 0;
    orc_fe(o, U, o->U1, dt, &o->passes[0]);
    orc_fe(o, o->U1, o->U2, dt, &o->passes[1]);
    for (size_t i = 0; i < n4; ++i) o->U2[i] = 0.75 * U[i] + 0.25 * o->U2[i];
    orc_fe(o, o->U2, Ur, dt, &o->passes[2]);
    const double a = 1.0 / 3, b = 2.0 / 3;
    for (size_t i = 0; i < n4; ++i) Ur[i] = a * U[i] + b * Ur[i];
}

/* The step loop of main.f90:145-162 for nsteps steps on the full array U
   (ghosts refreshed by Boundary_C before the first step, as main.f90:125).
   Optional outputs: dts[nsteps], counts[nsteps], masks[nsteps][N][N] (CellGPO
   after the last stage of each step), passes[nsteps][3]. */
int orc_run(orc_t *o, double *U, int nsteps, double *t_io, double tmax,
            double *dts, long *counts, int8_t *masks, int *passes)
{
    const int N = o->N;
    const size_t n4 = (size_t)o->W * o->W * 4;
    double *Ur = (double *)calloc(n4, sizeof(double));
    double t = t_io ? *t_io : 0.0;
    orc_bc(o, U);
    for (int s = 0; s < nsteps; ++s) {
        double dt = orc_setdt(o, U, tmax - t);
        orc_ssprk3(o, U, Ur, dt);
        orc_bc(o, U);
        t = t + dt;
        memcpy(U, Ur, n4 * sizeof(double));
        if (dts) dts[s] = dt;
        if (counts) counts[s] = o->count_detected;
        if (passes) for (int k = 0; k < 3; ++k) passes[s * 3 + k] = o->passes[k];
        if (masks)
            for (int n = 1; n <= N; ++n)
                for (int l = 1; l <= N; ++l)
                    masks[(size_t)s * N * N + (size_t)(n - 1) * N + (l - 1)] = o->gpo[CELL(o, l, n)];
    }
    if (t_io) *t_io = t;
    free(Ur);
    return 0;
}

/* Per-stage variant for parity on intermediate data: runs one step and returns
   the three stage outputs (after RK combination) and the three stage masks. */
int orc_step_stages(orc_t *o, double *U, double dt, double *stage_out /* 3*(N+8)^2*4 */,
                    int8_t *stage_masks /* 3*N*N */)
{
    const int N = o->N;
    const size_t n4 = (size_t)o->W * o->W * 4;
    double *Ur = (double *)calloc(n4, sizeof(double));
    orc_bc(o, U);
    o->count_detected = 0;
    double *outs[3] = {o->U1, o->U2, Ur};
    double *ins[3] = {U, o->U1, o->U2};
    for (int st = 0; st < 3; ++st) {
        orc_fe(o, ins[st], outs[st], dt, &o->passes[st]);
        if (st == 1) for (size_t i = 0; i < n4; ++i) o->U2[i] = 0.75 * U[i] + 0.25 * o->U2[i];
        if (st == 2) {
            const double a = 1.0 / 3, b = 2.0 / 3;
            for (size_t i = 0; i < n4; ++i) Ur[i] = a * U[i] + b * Ur[i];
        }
        if (stage_out) memcpy(stage_out + st * n4, outs[st], n4 * sizeof(double));
        if (stage_masks)
            for (int n = 1; n <= N; ++n)
                for (int l = 1; l <= N; ++l)
                    stage_masks[(size_t)st * N * N + (size_t)(n - 1) * N + (l - 1)] = o->gpo[CELL(o, l, n)];
    }
    orc_bc(o, U);
    memcpy(U, Ur, n4 * sizeof(double));
    free(Ur);
    return 0;
}

long orc_count_detected(const orc_t *o) { return o->count_detected; }
void orc_passes(const orc_t *o, int *p) { for (int k = 0; k < 3; ++k) p[k] = o->passes[k]; }

/* The Riemann solvers alone, for unit checks (dir 0 = x, 1 = y; flux 0 HLLC,
   1 HLL, 2 unsplit). */
void orc_hllc(const double *ul, const double *ur, int dir, double *F) { hllc(ul, ur, dir, F); }
void orc_riemann(int flux, const double *ul, const double *ur, int dir, double *F) { riemann(flux, ul, ur, dir, F); }

/* numFlux of the step (0 HLLC, 1 HLL, 2 unsplit). */
void orc_set_flux(orc_t *o, int flux) { o->flux = flux; }

/* Fixed a priori orders for every stage (NULL restores CellGPO = Mord). */
void orc_set_initial_orders(orc_t *o, const int8_t *orders)
{
    free(o->init_orders);
    o->init_orders = NULL;
    if (!orders) return;
    o->init_orders = (int8_t *)malloc((size_t)o->N * o->N);
    memcpy(o->init_orders, orders, (size_t)o->N * o->N);
}
