// Device-side numerics of the GP-MOOD step for gfx950.
//
// Every FP64 expression keeps the operation order of the reference Fortran
// (built with -fdefault-real-8, no FMA) so that results are bit-identical:
// the library is compiled with -ffp-contract=off, IEEE division and sqrt.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpm {

constexpr double kGamma = 1.4;  // constants.f90:9 (y = 1.4)

// Boundary-condition ids (constants.f90:24-25, 29; DMR is the problem id the
// reference also uses as BC_type).
enum { kNeumann = 17, kPeriodic = 18, kReflective = 20, kKH = 403 };
enum { kDirichlet = 19, kMach800 = 201, kDoubleMach800 = 301, kRMI = 401, kDMR = 35, kRT = 402 };
// Flux ids (constants.f90:21) and the template tags of the Riemann solvers.
enum { kHLLC = 100, kLLF = 101, kHLL = 102, kUnsplit = 103 };
// Integrator ids (constants.f90:19) and the problem id with a source term.
enum { kFE = 8, kSSPRK2 = 9, kSSPRK3 = 10, kSSPRK4 = 12 };
enum { kProblemRT = 51 };

// g of global_variables.f90:129 is real(4); it enters FP64 expressions widened.
constexpr double kGravity = (double)0.1f;
// Face ids (constants.f90:37): iL, iT, iR, iB.
enum { kL = 0, kT = 1, kR = 2, kB = 3 };

// Sphere stencil offsets in the reference's order (GP_init.f90:97-176).
__device__ constexpr int kSX[25] = {0, -1, 0, 1, 0, -1, -1, 1, 1, 0, -2, 0, 2,
                                    -3, -2, -1, 0, 1, 2, 3, 2, 1, 0, -1, -2};
__device__ constexpr int kSY[25] = {0, 0, 1, 0, -1, -1, 1, 1, -1, -2, 0, 2, 0,
                                    0, 1, 2, 3, 2, 1, 0, -1, -2, -3, -2, -1};

__host__ __device__ constexpr int stencil_sz(int ord) { return ord == 1 ? 1 : ord == 3 ? 5 : ord == 5 ? 13 : 25; }

// Reconstruction of a method: GP_MOOD_ (and NN_GP_MOOD, whose reconstruction
// is GP_MOOD_'s), POL_MOOD_ (mod_pol_MOOD.f90: order 3 from the polynomial
// weights, which the host puts in zT's order-3 block, and the cell values at
// any other order), or NN_GP_MOOD_CC_ (mod_NN_GP_MOOD_CC.f90: the cell values
// at order 1, the blend of values_NN at order 3).
enum { kReconGP = 0, kReconPol = 1, kReconCC = 2 };

// GP tables on the device, laid out for scalar loads in the unrolled loops:
// zT[(ord-1)/2][face][s][gp].
struct DevTables {
    double zT[4][4][25][5];
    double d2x[5], d2y[5];  // GP_d2x_sp(3, 1:5), GP_d2y_sp(3, 1:5)
    double gw[5];           // gauss_weight(ngp, 1:ngp)
};

// a / b correctly rounded (IEEE round to nearest even).  The compiler's FP64
// division (v_div_scale, v_rcp, two Newton steps, v_div_fmas, v_div_fixup) is
// faithful but not always correctly rounded: -0x1.6666666666663p+0 /
// -0x1.ffffffffffffbp-1 gives 0x1.6666666666666p+0 instead of
// 0x1.6666666666667p+0 (the exact quotient lies 0.49999999999999994 ulp from
// the latter).  Here a third Newton step makes the reciprocal y of the scaled
// divisor its correctly rounded value (Markstein: a faithful y and e = 1 - d y
// computed exactly give RN(1/d) in y + e y, unless d's significand is all
// ones), so the final fused step q + (n - d q) y is the correctly rounded
// quotient (Markstein: y = RN(1/d) and a faithful q).  The scale and fix-up
// instructions keep the compiler sequence's handling of extreme exponents,
// zeros, infinities and NaN.  A divisor with an all-ones significand takes an
// exact comparison of the remainders of q and its neighbour instead.
__device__ __forceinline__ double ddiv(double a, double b)
{
    bool flag_d, flag_n;
    const double d = __builtin_amdgcn_div_scale(a, b, false, &flag_d);  // scaled divisor
    const double n = __builtin_amdgcn_div_scale(a, b, true, &flag_n);   // scaled dividend
    double y = __builtin_amdgcn_rcp(d);
#pragma unroll
    for (int it = 0; it < 3; ++it) y = __builtin_fma(y, __builtin_fma(-d, y, 1.0), y);
    const double m = n * y;
    const double r = __builtin_fma(-d, m, n);
    double q = __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(r, y, m, flag_n), b, a);
    const unsigned long long bb = (unsigned long long)__double_as_longlong(b);
    if ((bb & 0xFFFFFFFFFFFFFull) == 0xFFFFFFFFFFFFFull) {
        const double aq = fabs(q), ab = fabs(b);
        if (aq >= 0x1p-900 && aq <= 0x1p900 && ab >= 0x1p-900 && ab <= 0x1p900) {
            const double e = __builtin_fma(-q, b, a);  // exact remainder of the faithful q
            if (e != 0.0) {
                // neighbour of q towards the exact quotient q + e / b
                const bool up = (e > 0.0) == (b > 0.0);
                const long long qb = __double_as_longlong(q);
                const double q2 = __longlong_as_double(qb + (((q > 0.0) == up) ? 1 : -1));
                const double e2 = __builtin_fma(-q2, b, a);
                if (fabs(e2) < fabs(e) || (fabs(e2) == fabs(e) && (__double_as_longlong(q2) & 1) == 0)) q = q2;
            }
        }
    }
    return q;
}

// A divisor shared by several quotients (the density of one Riemann state
// divides its momenta, kinetic energy, pressure and total energy).  When
// div_scale leaves both operands alone (no scaling, flag clear), ddiv's
// reciprocal chain -- v_rcp and the three Newton steps -- depends on the
// divisor only, so it is computed once here and qdiv replays the rest of
// ddiv's sequence (m = n y, r = n - d m, fmas, fixup) with the same operands:
// the same instructions on the same values, hence the same bits.
//
// div_scale scales only near the ends of the range: an exponent difference
// of 768 or more, a denormal divisor, reciprocal or quotient, a numerator
// exponent of 53 or less (biased).  With the divisor in [2^-300, 2^300) and
// the numerator's frexp exponent in [-600, 400] (|a| in [2^-601, 2^400)) none
// of these can happen.  Numerators that are zero, infinite or NaN (frexp
// exponent 0) take the shared form too: div_fixup returns their quotient
// (signed zero, signed infinity, quiet NaN) whatever the approximation.
// Every other numerator, and a divisor outside the range or with an all-ones
// significand (ddiv's exact remainder comparison), goes through ddiv itself.
struct Recip {
    double b, y;
    bool fast;
};

// biased frexp exponent of a numerator: at most 1000 when a may share
__device__ __forceinline__ unsigned qdiv_span(double a)
{
    return (unsigned)(__builtin_amdgcn_frexp_exp(a) + 600);
}

__host__ __device__ __forceinline__ Recip recip(double b)
{
#ifdef __HIP_DEVICE_COMPILE__
    double y = __builtin_amdgcn_rcp(b);
#pragma unroll
    for (int it = 0; it < 3; ++it) y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
    const unsigned long long bb = (unsigned long long)__double_as_longlong(b);
    const unsigned e = (unsigned)(bb >> 52) & 0x7FFu;  // biased exponent: normal b in [2^-300, 2^300)
    return Recip{b, y, e - 723u < 600u && (bb & 0xFFFFFFFFFFFFFull) != 0xFFFFFFFFFFFFFull};
#else
    return Recip{b, 0.0, false};
#endif
}

// Whether a / R.b may take the shared form.
__device__ __forceinline__ bool qdiv_ok(double a, const Recip &R) { return R.fast && qdiv_span(a) <= 1000u; }

// a / R.b in the shared form, for a with qdiv_ok(a, R): ddiv's sequence after
// the reciprocal, unscaled.
__device__ __forceinline__ double qdiv_shared(double a, const Recip &R)
{
    const double m = a * R.y;
    const double r = __builtin_fma(-R.b, m, a);
    return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(r, R.y, m, false), R.b, a);
}

// a / R.b correctly rounded, bit-identical to ddiv(a, R.b).
__host__ __device__ __forceinline__ double qdiv(double a, const Recip &R)
{
#ifdef __HIP_DEVICE_COMPILE__
    return qdiv_ok(a, R) ? qdiv_shared(a, R) : ddiv(a, R.b);
#else
    return a / R.b;
#endif
}

// The reference's FP64 a / b in code compiled for both sides: ddiv on the
// device, the host's IEEE division on the host.
__host__ __device__ __forceinline__ double dv(double a, double b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return ddiv(a, b);
#else
    return a / b;
#endif
}

// Fortran MIN/MAX of the reference build: the second operand unless the first
// compares strictly smaller (larger).
__device__ __forceinline__ double fmin2(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double fmax2(double a, double b) { return (a > b) ? a : b; }

// pressure, physics.f90:132-143
__host__ __device__ __forceinline__ double pressure(double r, double mx, double my, double E)
{
    return (kGamma - 1.0) * (E - dv(0.5 * (mx * mx + my * my), r));
}

// Conserved state of a cell as read from memory (rho, rho*u, rho*v, E).
struct Cell {
    double r, mx, my, E;
};

__device__ __forceinline__ Cell to_cell(const double4 &v) { return Cell{v.x, v.y, v.z, v.w}; }

// Ghost-cell rule of Boundary_C (BC.f90:18-135) as a map from a cell index in
// [-4, N+3] to the interior cell it copies and whether the normal momentum is
// negated.  Indices further out (only reached by partial tiles, whose values are
// never used) are clamped into the grid.
__device__ __forceinline__ int bc_map(int i, int N, int bc, bool &flip)
{
    flip = false;
    if (i < 0) {
        if (bc == kNeumann) i = 0;
        else if (bc == kPeriodic) i = i + N;
        else { i = -1 - i; flip = true; }
    } else if (i >= N) {
        if (bc == kNeumann) i = N - 1;
        else if (bc == kPeriodic) i = i - N;
        else { i = 2 * N - 1 - i; flip = true; }
    }
    i = i < 0 ? 0 : (i > N - 1 ? N - 1 : i);
    return i;
}

// The rule of each axis: KH_BC (BC.f90:383-406) is periodic in x and Neumann
// in y; the other boundary types use one rule on both axes.
__device__ __forceinline__ int bc_x(int bc) { return bc == kKH ? kPeriodic : bc; }
__device__ __forceinline__ int bc_y(int bc) { return bc == kKH ? kNeumann : bc; }

// A stage input's cell (i, j) of an NX x NY grid under a copy boundary
// (Neumann, Periodic, Reflective, KH_BC): branch-free index arithmetic.
__device__ __forceinline__ double4 load_copy(const double4 *__restrict__ U, int NX, int NY, int bc, int i, int j)
{
    bool fx, fy;
    int si = bc_map(i, NX, bc_x(bc), fx);
    int sj = bc_map(j, NY, bc_y(bc), fy);
    double4 v = U[(size_t)sj * NX + si];
    if (fx) v.y = -v.y;
    if (fy) v.z = -v.z;
    return v;
}

__device__ __forceinline__ double4 load_bc(const double4 *__restrict__ U, int N, int bc, int i, int j)
{
    return load_copy(U, N, N, bc, i, j);
}

// primitive_to_conservative, physics.f90:146-160
__host__ __device__ __forceinline__ double4 p2c(double r, double vx, double vy, double p)
{
    return make_double4(r, r * vx, r * vy, dv(p, kGamma - 1.) + 0.5 * r * (vx * vx + vy * vy));
}

// Cell centre of 0-based cell i: mesh_x(l) = l*dx - dx/2 with l = i+1 (mesh_init.f90:14-20).
__host__ __device__ __forceinline__ double mesh_c(int i, double d) { return (double)(i + 1) * d - d / 2; }

// The boundary of a stage input on an NX x NY grid (lf x nf): the BC_type and
// a pointer (device memory in kernels, host memory on the host; kept out of the
// kernel argument so that indexing it does not copy the argument to scratch)
// to the fixed ghost states of the formula rules, computed on the host with
// primitive_to_conservative (BC.f90:60-381):
//   Mach800_BC        st[0] jet (1.4, 0, 800, 1)
//   DoubleMach800_BC  st[0] bottom jet (1.4, 0, 800, 1), st[1] top jet (1.4, 0, -800, 1)
//   DMR               st[0] post-shock (8, 7.1447096, -4.125, 116.5), st[1] pre-shock (1.4, 0, 0, 1)
//   RMI_BC            st[0] inflow (2.666666666666667, 2 sqrt(1.4), 0, 4.5)
//   Dirichlet         st[0] left (3.857143, 2.629369, 0, 10.33333), st[g] right ghost column g = 1..4
struct Bnd {
    int NX, NY, bc;
    double dx, dy, sqrt3;
    const double4 *st;  // [5]
};
constexpr int kBndStates = 5;

__host__ __device__ __forceinline__ bool index_map_bc(int bc)
{
    return bc == kNeumann || bc == kPeriodic || bc == kReflective || bc == kKH;
}

// Abscissa where the DMR top boundary switches from the post- to the
// pre-shock state at time t_rk (BC.f90:301).
__host__ __device__ __forceinline__ double dmr_xmin_top(double t_rk, double sqrt3)
{
    return 1. / 6 + dv(10. * t_rk, 0.5 * sqrt3) + dv(1., sqrt3);
}

// Ghost column i (i < 0 or i >= NX) of interior row j: the x part of
// Boundary_C, which the reference applies to every row before the y part.
// src(i, j) returns interior cell (i, j).
template <class Src>
__host__ __device__ __forceinline__ double4 ghost_col(const Bnd &B, const Src &src, int i, int j)
{
    const int NX = B.NX;
    const bool right = i >= NX;
    const int g = right ? i - NX + 1 : -i;  // ghost layer 1..4
    double4 v;
    switch (B.bc) {
    case kPeriodic: case kKH:
        return src(right ? g - 1 : NX - g, j);
    case kReflective: case kRT:
        v = src(right ? NX - g : g - 1, j);
        v.y = -v.y;
        return v;
    case kDirichlet:
        return right ? B.st[g] : B.st[0];
    case kDMR: {  // f_DMR, IC.f90:336-355
        const double x = mesh_c(i, B.dx), y = mesh_c(j, B.dy);
        return y > (x - 1. / 6) * B.sqrt3 ? B.st[0] : B.st[1];
    }
    case kRMI:
        return right ? src(NX - 1, j) : B.st[0];
    default:  // Neumann, Mach800_BC, DoubleMach800_BC: outflow
        return src(right ? NX - 1 : 0, j);
    }
}

// Value of cell (i, j), i in [-4, NX+3], j in [-4, NY+3], of a stage input
// after Boundary_C (BC.f90:18-413): interior cells as they are; ghost columns
// by ghost_col; ghost rows (corners included) by the y rule applied to the
// column's values, as the reference's y loop runs over the ghost columns too.
// Indices further out (reached only by partial tiles, whose values are never
// used) are clamped first.  xmin_top: dmr_xmin_top at the stage's t_rk.
template <class Src>
__host__ __device__ __forceinline__ double4 bnd_value(const Bnd &B, const Src &src, int i, int j, double xmin_top)
{
    const int NX = B.NX, NY = B.NY;
    i = i < -4 ? -4 : (i > NX + 3 ? NX + 3 : i);
    j = j < -4 ? -4 : (j > NY + 3 ? NY + 3 : j);
    const bool inx = i >= 0 && i < NX;
    if (j >= 0 && j < NY) return inx ? src(i, j) : ghost_col(B, src, i, j);
    const bool top = j >= NY;
    const int g = top ? j - NY + 1 : -j;  // ghost layer 1..4
    auto col = [&](int jj) { return inx ? src(i, jj) : ghost_col(B, src, i, jj); };
    double4 v;
    switch (B.bc) {
    case kPeriodic:
        return col(top ? g - 1 : NY - g);
    case kReflective: case kRMI:
        v = col(top ? NY - g : g - 1);
        v.z = -v.z;
        return v;
    case kMach800:
        if (!top && fabs(mesh_c(i, B.dx) - 0.75) <= 0.05) return B.st[0];
        return col(top ? NY - 1 : 0);
    case kDoubleMach800:
        if (fabs(mesh_c(i, B.dx) - 0.75) <= 0.05) return top ? B.st[1] : B.st[0];
        return col(top ? NY - 1 : 0);
    case kDMR: {
        const double x = mesh_c(i, B.dx);
        if (top) return x < xmin_top ? B.st[0] : B.st[1];
        if (x > 1. / 6) {
            v = col(g - 1);
            v.z = -v.z;
            return v;
        }
        return B.st[0];
    }
    case kRT: {  // hydrostatic reflection in primitive variables
        const int k = g - 1;
        const double4 u = col(top ? NY - 1 - k : k);
        const double vx = dv(u.y, u.x), vy = dv(u.z, u.x), p = pressure(u.x, u.y, u.z, u.w);
        const double h = (double)(2 * k) * B.dy * kGravity * u.x;
        return p2c(u.x, vx, -vy, top ? p - h : p + h);
    }
    default:  // Neumann, Dirichlet, KH_BC: outflow
        return col(top ? NY - 1 : 0);
    }
}

// Interior cells of a device state (row-major, NX per row).
struct DevSrc {
    const double4 *__restrict__ U;
    int NX;
    __device__ __forceinline__ double4 operator()(int i, int j) const { return U[(size_t)j * NX + i]; }
};

// A stage input's cell (i, j) with the boundary rule applied.
__device__ __forceinline__ double4 load_bnd(const double4 *__restrict__ U, const Bnd &B, int i, int j,
                                           double xmin_top)
{
    if (i >= 0 && i < B.NX && j >= 0 && j < B.NY) return U[(size_t)j * B.NX + i];
    return bnd_value(B, DevSrc{U, B.NX}, i, j, xmin_top);
}

// c2p_local (physics.f90:51-75) of one Riemann state: velocities, pressure
// (physics.f90:132-143) and sound speed, the four divisions by the density
// sharing its reciprocal.
struct Prim {
    double vx, vy, p, c;
    Recip r;
};

__device__ __forceinline__ Prim c2p_shared(const double u[4])
{
    Prim s;
    s.r = recip(u[0]);
    const double ke = 0.5 * (u[1] * u[1] + u[2] * u[2]);
    // the shared form unconditionally, then one check for all four quotients
    // and one (rarely taken) branch to the full divisions
    s.vx = qdiv_shared(u[1], s.r);
    s.vy = qdiv_shared(u[2], s.r);
    s.p = (kGamma - 1.0) * (u[3] - qdiv_shared(ke, s.r));
    double cq = qdiv_shared(kGamma * s.p, s.r);
    const unsigned span = max(max(qdiv_span(u[1]), qdiv_span(u[2])), max(qdiv_span(ke), qdiv_span(kGamma * s.p)));
    if (!s.r.fast || span > 1000u) {
        s.vx = ddiv(u[1], u[0]);
        s.vy = ddiv(u[2], u[0]);
        s.p = (kGamma - 1.0) * (u[3] - ddiv(ke, u[0]));
        cq = ddiv(kGamma * s.p, u[0]);
    }
    s.c = sqrt(cq);
    return s;
}

// HLLC flux (mod_HLLC.f90:11-169) with c2p_local and Flux (physics.f90:9-75).
// DIR 0: x-face (normal momentum mx), DIR 1: y-face.
// hllc_adm returns false, leaving F untouched, when either state is
// inadmissible (rho <= 0, p <= 0 or NaN: c2p_local, physics.f90:72).
template <int DIR>
__device__ __forceinline__ bool hllc_adm(const double ul[4], const double ur[4], double F[4])
{
    const Prim L = c2p_shared(ul), R = c2p_shared(ur);
    const double vxl = L.vx, vyl = L.vy, pl = L.p, cl = L.c;
    const double vxr = R.vx, vyr = R.vy, pr = R.p, cr = R.c;
    const bool adm = !(ul[0] <= 0. || pl <= 0. || isnan(ul[0]) || isnan(pl) ||
                       ur[0] <= 0. || pr <= 0. || isnan(ur[0]) || isnan(pr));
    if (!adm) return false;
    constexpr int md = DIR == 0 ? 1 : 2, om = DIR == 0 ? 2 : 1;
    const double vnl = DIR == 0 ? vxl : vyl, vtl = DIR == 0 ? vyl : vxl;
    const double vnr = DIR == 0 ? vxr : vyr, vtr = DIR == 0 ? vyr : vxr;
    double Fl[4], Fr[4];
    if (DIR == 0) {
        Fl[0] = ul[1]; Fl[1] = ul[0] * (vxl * vxl) + pl; Fl[2] = ul[2] * vxl; Fl[3] = vxl * (ul[3] + pl);
        Fr[0] = ur[1]; Fr[1] = ur[0] * (vxr * vxr) + pr; Fr[2] = ur[2] * vxr; Fr[3] = vxr * (ur[3] + pr);
    } else {
        Fl[0] = ul[2]; Fl[1] = ul[1] * vyl; Fl[2] = ul[0] * (vyl * vyl) + pl; Fl[3] = vyl * (ul[3] + pl);
        Fr[0] = ur[2]; Fr[1] = ur[1] * vyr; Fr[2] = ur[0] * (vyr * vyr) + pr; Fr[3] = vyr * (ur[3] + pr);
    }
    const double SL = fmin2(vnl - cl, vnr - cr);
    const double SR = fmax2(vnl + cl, vnr + cr);
    const double numerL = SL - vnl, numerR = SR - vnr;
    double uStar = ul[0] * vnl * numerL - ur[0] * vnr * numerR;
    uStar = (pr - pl) + uStar;
    uStar = dv(uStar, ul[0] * numerL - ur[0] * numerR);
    if (SL >= 0.) {
        F[0] = Fl[0]; F[1] = Fl[1]; F[2] = Fl[2]; F[3] = Fl[3];
    } else if (SL < 0. && uStar >= 0.) {
        const double dStarL = dv(ul[0] * numerL, SL - uStar);
        double Us[4];
        Us[0] = 1.0; Us[md] = uStar; Us[om] = vtl;
        Us[3] = qdiv(ul[3], L.r) + (uStar - vnl) * (uStar + dv(pl, ul[0] * numerL));
#pragma unroll
        for (int k = 0; k < 4; ++k) F[k] = Fl[k] + SL * (Us[k] * dStarL - ul[k]);
    } else if (uStar < 0. && SR >= 0.) {
        const double dStarR = dv(ur[0] * numerR, SR - uStar);
        double Us[4];
        Us[0] = 1.0; Us[md] = uStar; Us[om] = vtr;
        Us[3] = qdiv(ur[3], R.r) + (uStar - vnr) * (uStar + dv(pr, ur[0] * numerR));
#pragma unroll
        for (int k = 0; k < 4; ++k) F[k] = Fr[k] + SR * (Us[k] * dStarR - ur[k]);
    } else {
        F[0] = Fr[0]; F[1] = Fr[1]; F[2] = Fr[2]; F[3] = Fr[3];
    }
    return true;
}

// Flux of an inadmissible face: mod_HLLC.f90:166, F = 1/(vx-vx) with vx the
// x-velocity of the right state's c2p (+Inf for finite vx, NaN otherwise).
__device__ __forceinline__ void hllc_inadmissible(const double ur[4], double F[4])
{
    const double vx = dv(ur[1], ur[0]);
    const double z = vx - vx;
    const double f = dv(1.0, z);
    F[0] = f; F[1] = f; F[2] = f; F[3] = f;
}

template <int DIR>
__device__ __forceinline__ void hllc(const double ul[4], const double ur[4], double F[4])
{
    if (!hllc_adm<DIR>(ul, ur, F)) hllc_inadmissible(ur, F);
}

// HLL flux (mod_HLL.f90:11-166); false, leaving F untouched, when either
// state is inadmissible (the flux is then that of hllc_inadmissible).
template <int DIR>
__device__ __forceinline__ bool hll_adm(const double ul[4], const double ur[4], double F[4])
{
    const Prim L = c2p_shared(ul), R = c2p_shared(ur);
    const double vxl = L.vx, vyl = L.vy, pl = L.p, cl = L.c;
    const double vxr = R.vx, vyr = R.vy, pr = R.p, cr = R.c;
    const bool adm = !(ul[0] <= 0. || pl <= 0. || isnan(ul[0]) || isnan(pl) ||
                       ur[0] <= 0. || pr <= 0. || isnan(ur[0]) || isnan(pr));
    if (!adm) return false;
    const double vnl = DIR == 0 ? vxl : vyl, vnr = DIR == 0 ? vxr : vyr;
    double Fl[4], Fr[4];
    if (DIR == 0) {
        Fl[0] = ul[1]; Fl[1] = ul[0] * (vxl * vxl) + pl; Fl[2] = ul[2] * vxl; Fl[3] = vxl * (ul[3] + pl);
        Fr[0] = ur[1]; Fr[1] = ur[0] * (vxr * vxr) + pr; Fr[2] = ur[2] * vxr; Fr[3] = vxr * (ur[3] + pr);
    } else {
        Fl[0] = ul[2]; Fl[1] = ul[1] * vyl; Fl[2] = ul[0] * (vyl * vyl) + pl; Fl[3] = vyl * (ul[3] + pl);
        Fr[0] = ur[2]; Fr[1] = ur[1] * vyr; Fr[2] = ur[0] * (vyr * vyr) + pr; Fr[3] = vyr * (ur[3] + pr);
    }
    const double SL = fmin2(vnl - cl, vnr - cr);
    const double SR = fmax2(vnl + cl, vnr + cr);
    if (SL > 0.) {
#pragma unroll
        for (int k = 0; k < 4; ++k) F[k] = Fl[k];
    } else if (SR < 0.) {
#pragma unroll
        for (int k = 0; k < 4; ++k) F[k] = Fr[k];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) F[k] = dv(SR * Fl[k] - SL * Fr[k] + SR * SL * (ur[k] - ul[k]), SR - SL);
    }
    return true;
}

// Local Lax-Friedrichs flux (mod_LLF.f90:11-93); false when either state is
// inadmissible (the flux is then that of hllc_inadmissible, mod_LLF.f90:89).
template <int DIR>
__device__ __forceinline__ bool llf_adm(const double ul[4], const double ur[4], double F[4])
{
    const Prim L = c2p_shared(ul), R = c2p_shared(ur);
    const double vxl = L.vx, vyl = L.vy, pl = L.p, cl = L.c;
    const double vxr = R.vx, vyr = R.vy, pr = R.p, cr = R.c;
    const bool adm = !(ul[0] <= 0. || pl <= 0. || isnan(ul[0]) || isnan(pl) ||
                       ur[0] <= 0. || pr <= 0. || isnan(ur[0]) || isnan(pr));
    if (!adm) return false;
    const double vnl = DIR == 0 ? vxl : vyl, vnr = DIR == 0 ? vxr : vyr;
    double Fl[4], Fr[4];
    if (DIR == 0) {
        Fl[0] = ul[1]; Fl[1] = ul[0] * (vxl * vxl) + pl; Fl[2] = ul[2] * vxl; Fl[3] = vxl * (ul[3] + pl);
        Fr[0] = ur[1]; Fr[1] = ur[0] * (vxr * vxr) + pr; Fr[2] = ur[2] * vxr; Fr[3] = vxr * (ur[3] + pr);
    } else {
        Fl[0] = ul[2]; Fl[1] = ul[1] * vyl; Fl[2] = ul[0] * (vyl * vyl) + pl; Fl[3] = vyl * (ul[3] + pl);
        Fr[0] = ur[2]; Fr[1] = ur[1] * vyr; Fr[2] = ur[0] * (vyr * vyr) + pr; Fr[3] = vyr * (ur[3] + pr);
    }
    const double vmcl = vnl - cl, vpcl = vnl + cl, vmcr = vnr - cr, vpcr = vnr + cr;
    const double smax = fmax2(fmax2(fmax2(fabs(vmcl), fabs(vmcr)), fabs(vpcl)), fabs(vpcr));
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] = 0.5 * (Fr[k] + Fl[k]) - 0.5 * smax * (ur[k] - ul[k]);
    return true;
}

// Pressure-velocity ("unsplit") flux (mod_unsplit.f90:11-71); false when
// either state is inadmissible.
template <int DIR>
__device__ __forceinline__ bool unsplit_adm(const double ul[4], const double ur[4], double F[4])
{
    const Prim L = c2p_shared(ul), R = c2p_shared(ur);
    const double vxl = L.vx, vyl = L.vy, pl = L.p, cl = L.c;
    const double vxr = R.vx, vyr = R.vy, pr = R.p, cr = R.c;
    const bool adm = !(ul[0] <= 0. || pl <= 0. || isnan(ul[0]) || isnan(pl) ||
                       ur[0] <= 0. || pr <= 0. || isnan(ur[0]) || isnan(pr));
    if (!adm) return false;
    const double a = 1.1 * fmax2(ul[0] * cl, ur[0] * cr);
    const double vnl = DIR == 0 ? vxl : vyl, vnr = DIR == 0 ? vxr : vyr;
    const double vtl = DIR == 0 ? vyl : vxl, vtr = DIR == 0 ? vyr : vxr;
    const double pistar = 0.5 * (pl + pr) - 0.5 * a * (vnr - vnl);
    const double ustar = 0.5 * (vnl + vnr) - dv(0.5 * (pr - pl), a);
    const double trans = -(0.5 * a * (vtr - vtl));
    const bool left = ustar > 0.;
    constexpr int md = DIR == 0 ? 1 : 2, om = DIR == 0 ? 2 : 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] = (left ? ul[k] : ur[k]) * ustar;
    F[md] = F[md] + pistar;
    F[om] = F[om] + trans;
    F[3] = F[3] + pistar * ustar;
    return true;
}

// Flux of an inadmissible face in mod_unsplit.f90:68: F = 1/(no-no), where the
// reference's c2p_local calls write vx2 and vy2 into the one variable `no`,
// which ends up holding the right state's vy**2.
__device__ __forceinline__ void unsplit_inadmissible(const double ur[4], double F[4])
{
    const double vy = dv(ur[2], ur[0]);
    const double no = vy * vy;
    const double z = no - no;
    const double f = dv(1.0, z);
    F[0] = f; F[1] = f; F[2] = f; F[3] = f;
}

// The Riemann solver of numFlux FLUX on a DIR face.
template <int FLUX, int DIR>
__device__ __forceinline__ void riemann(const double ul[4], const double ur[4], double F[4])
{
    if (FLUX == kHLL) {
        if (!hll_adm<DIR>(ul, ur, F)) hllc_inadmissible(ur, F);
    } else if (FLUX == kLLF) {
        if (!llf_adm<DIR>(ul, ur, F)) hllc_inadmissible(ur, F);
    } else if (FLUX == kUnsplit) {
        if (!unsplit_adm<DIR>(ul, ur, F)) unsplit_inadmissible(ur, F);
    } else {
        if (!hllc_adm<DIR>(ul, ur, F)) hllc_inadmissible(ur, F);
    }
}

// Run-time direction: a y-face solve is the x-solve on momentum-swapped states
// with the flux's momenta swapped back (exact for all four solvers: their
// y-direction formulas are the x-direction ones with the roles of the
// momenta exchanged, operation for operation).  One x-solve serves both
// directions, with the swaps as selects: the lanes of a MOOD wavefront hold
// x- and y-faces together, and a branch per direction made every wavefront
// run the solver twice.  For an x-face the selects pass the states through,
// so the operations are those of riemann<FLUX, 0>.
template <int FLUX>
__device__ __forceinline__ void riemann_dir(int dir, const double ul[4], const double ur[4], double F[4])
{
    const bool y = dir != 0;
    const double sl[4] = {ul[0], y ? ul[2] : ul[1], y ? ul[1] : ul[2], ul[3]};
    const double sr[4] = {ur[0], y ? ur[2] : ur[1], y ? ur[1] : ur[2], ur[3]};
    double G[4];
    const bool ok = FLUX == kHLL       ? hll_adm<0>(sl, sr, G)
                    : FLUX == kLLF     ? llf_adm<0>(sl, sr, G)
                    : FLUX == kUnsplit ? unsplit_adm<0>(sl, sr, G)
                                       : hllc_adm<0>(sl, sr, G);
    if (ok) {
        F[0] = G[0]; F[1] = y ? G[2] : G[1]; F[2] = y ? G[1] : G[2]; F[3] = G[3];
    } else if (FLUX == kUnsplit) {
        unsplit_inadmissible(ur, F);
    } else {
        hllc_inadmissible(ur, F);
    }
}

// A stencil accessor anchored at cell (i, j): q(dx, dy) = acc(i + dx, j + dy).
// Accessors over an LDS tile overload anchor() with a shifted pointer, so a
// stencil point costs one address add per side (the uniform offset comes
// from the scalar unit) instead of a row product per point.
template <class Acc>
struct Anchored {
    Acc acc;
    int i, j;
    __device__ __forceinline__ double4 operator()(int dx, int dy) const { return acc(i + dx, j + dy); }
};
template <class Acc>
__device__ __forceinline__ Anchored<Acc> anchor(const Acc &acc, int i, int j)
{
    return Anchored<Acc>{acc, i, j};
}

// GP reconstruction of both sides of a face at all NGP Gauss points
// (mod_GP_MOOD.f90:46-65): ua[j][k] = sum_s zT(ORD, s, FA, j) * q_A(k, s) and
// likewise ub for cell B, each summed from s = 1 in stencil order starting at 0
// (dot_product).  The stencil loop stays rolled (UNR points per trip) so the
// weights are fetched with scalar loads per trip instead of being held in
// registers across the whole stencil.
template <int ORD, int NGP, int FA, int FB, int UNR, class Acc>
__device__ __forceinline__ void recon2(const DevTables *__restrict__ T, const Acc &acc, int ai, int aj, int bi,
                                       int bj, double (&ua)[NGP][4], double (&ub)[NGP][4])
{
    constexpr int SZ = stencil_sz(ORD);
    constexpr int OI = (ORD - 1) / 2;
#pragma unroll
    for (int j = 0; j < NGP; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) { ua[j][k] = 0.0; ub[j][k] = 0.0; }
    const auto A = anchor(acc, ai, aj);
    const auto B = anchor(acc, bi, bj);
#pragma unroll UNR
    for (int s = 0; s < SZ; ++s) {
        const int dx = kSX[s], dy = kSY[s];
        const double4 qa = A(dx, dy);
        const double4 qb = B(dx, dy);
#pragma unroll
        for (int j = 0; j < NGP; ++j) {
            const double za = T->zT[OI][FA][s][j];
            const double zb = T->zT[OI][FB][s][j];
            ua[j][0] = ua[j][0] + za * qa.x;
            ua[j][1] = ua[j][1] + za * qa.y;
            ua[j][2] = ua[j][2] + za * qa.z;
            ua[j][3] = ua[j][3] + za * qa.w;
            ub[j][0] = ub[j][0] + zb * qb.x;
            ub[j][1] = ub[j][1] + zb * qb.y;
            ub[j][2] = ub[j][2] + zb * qb.z;
            ub[j][3] = ub[j][3] + zb * qb.w;
        }
    }
}

#ifndef GPM_RECON_UNROLL
#define GPM_RECON_UNROLL 1
#endif

// Face flux: Riemann solve at each Gauss point of the face between cell A
// (left/bottom) and B (right/top), then L_Flux = sum_j F_j * gauss_weight(ngp,j)
// from 0 (mod_FE.f90:80-98 and 112-130).
template <int ORD, int NGP, int DIR, int FLUX, class Acc>
__device__ __forceinline__ void face_flux(const DevTables *__restrict__ T, const Acc &acc, int ai, int aj,
                                          int bi, int bj, double L[4])
{
    double ua[NGP][4], ub[NGP][4];
    recon2<ORD, NGP, DIR == 0 ? kR : kT, DIR == 0 ? kL : kB, GPM_RECON_UNROLL>(T, acc, ai, aj, bi, bj, ua, ub);
    L[0] = 0.0; L[1] = 0.0; L[2] = 0.0; L[3] = 0.0;
    if (ORD == 1) {
        // First order: every Gauss point carries the same two states, so one
        // Riemann solve serves all of them; the quadrature sum is unchanged.
        double F[4];
        riemann<FLUX, DIR>(ua[0], ub[0], F);
#pragma unroll
        for (int j = 0; j < NGP; ++j) {
            const double w = T->gw[j];
#pragma unroll
            for (int k = 0; k < 4; ++k) L[k] = L[k] + F[k] * w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < NGP; ++j) {
            double F[4];
            riemann<FLUX, DIR>(ua[j], ub[j], F);
            const double w = T->gw[j];
#pragma unroll
            for (int k = 0; k < 4; ++k) L[k] = L[k] + F[k] * w;
        }
    }
}

// Face flux for a face whose order is only known at run time (ord <= MORD).
template <int MORD, int NGP, int DIR, int FLUX, class Acc>
__device__ __forceinline__ void face_flux_dyn(const DevTables *__restrict__ T, const Acc acc, int ord, int ai,
                                              int aj, int bi, int bj, double L[4])
{
    if (MORD >= 7 && ord == 7) face_flux<(MORD >= 7 ? 7 : 1), NGP, DIR, FLUX>(T, acc, ai, aj, bi, bj, L);
    else if (MORD >= 5 && ord == 5) face_flux<(MORD >= 5 ? 5 : 1), NGP, DIR, FLUX>(T, acc, ai, aj, bi, bj, L);
    else if (MORD >= 3 && ord == 3) face_flux<(MORD >= 3 ? 3 : 1), NGP, DIR, FLUX>(T, acc, ai, aj, bi, bj, L);
    else face_flux<1, NGP, DIR, FLUX>(T, acc, ai, aj, bi, bj, L);
}

// Flux of one face at one Gauss point with the face order known only at run
// time (ord <= MORD): one reconstruction loop whose trip count follows the
// order (the same operations, in the same order, as recon2 for that order) and one
// Riemann solve, so lanes of a wavefront that hold different orders run a
// single code path instead of one inlined solver per order.
//
// With recon = kReconPol or kReconCC (Mord 3) the face states at order 1 are
// the cell values themselves; with kReconCC those at order 3 are the blend
// v1 * U(cell) + v2 * GP value (face_flux_cc).
template <int MORD, int FLUX, class Acc>
__device__ __forceinline__ void face_gp_flux_dyn(const DevTables *__restrict__ T, const Acc &acc, int ord, int fa,
                                                 int fb, int dir, int j, int ai, int aj, int bi, int bj, double F[4],
                                                 int recon = kReconGP, double va1 = 0.0, double va2 = 1.0,
                                                 double vb1 = 0.0, double vb2 = 1.0)
{
    constexpr int SZM = stencil_sz(MORD);
    const int oi = (ord - 1) / 2, sz = stencil_sz(ord);
    const double *__restrict__ za_row = &T->zT[oi][fa][0][j];
    const double *__restrict__ zb_row = &T->zT[oi][fb][0][j];
    double ua[4] = {0.0, 0.0, 0.0, 0.0}, ub[4] = {0.0, 0.0, 0.0, 0.0};
    const auto A = anchor(acc, ai, aj);
    const auto B = anchor(acc, bi, bj);
#pragma unroll 5
    for (int s = 0; s < SZM; ++s) {
        if (s >= sz) break;
        const int dx = kSX[s], dy = kSY[s];
        const double4 qa = A(dx, dy);
        const double4 qb = B(dx, dy);
        const double za = za_row[s * 5], zb = zb_row[s * 5];
        ua[0] = ua[0] + za * qa.x; ua[1] = ua[1] + za * qa.y; ua[2] = ua[2] + za * qa.z; ua[3] = ua[3] + za * qa.w;
        ub[0] = ub[0] + zb * qb.x; ub[1] = ub[1] + zb * qb.y; ub[2] = ub[2] + zb * qb.z; ub[3] = ub[3] + zb * qb.w;
    }
    if (MORD == 3 && recon != kReconGP) {
        const double4 ca = acc(ai, aj), cb = acc(bi, bj);
        const double qa[4] = {ca.x, ca.y, ca.z, ca.w}, qb[4] = {cb.x, cb.y, cb.z, cb.w};
        if (ord == 1) {
#pragma unroll
            for (int k = 0; k < 4; ++k) { ua[k] = qa[k]; ub[k] = qb[k]; }
        } else if (recon == kReconCC) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ua[k] = va1 * qa[k] + va2 * ua[k];
                ub[k] = vb1 * qb[k] + vb2 * ub[k];
            }
        }
    }
    riemann_dir<FLUX>(dir, ua, ub, F);
}

// values_NN(l, n, 1:2) of cell (i, j) (0-based interior indices, l = i+1,
// n = j+1) from its Fortran storage, dimension(-1:lf+2, -1:nf+2, 2); cells
// outside the reconstructed range [-1, NX] x [-1, NY] read (0, 1).
__device__ __forceinline__ void vnn_at(const float *__restrict__ vnn, int NX, int NY, int i, int j, double &v1,
                                       double &v2)
{
    if (i < -1 || i > NX || j < -1 || j > NY) {
        v1 = 0.0;
        v2 = 1.0;
        return;
    }
    const size_t W = (size_t)NX + 4, c = (size_t)(i + 2) + (size_t)(j + 2) * W;
    v1 = (double)vnn[c];
    v2 = (double)vnn[c + W * ((size_t)NY + 4)];
}

// Face flux of NN_GP_MOOD_CC (NN_GP_MOOD_CC_, mod_NN_GP_MOOD_CC.f90:57-68;
// Mord 3): order 1 takes each cell's own value, order 3 the GP value blended
// with it, v1 * U(cell) + v2 * dot_product(zT, q), with the cell's values_NN.
template <int NGP, int DIR, int FLUX, class Acc>
__device__ __forceinline__ void face_flux_cc(const DevTables *__restrict__ T, const Acc &acc, int ord, int ai, int aj,
                                             int bi, int bj, double va1, double va2, double vb1, double vb2,
                                             double L[4])
{
    const double4 ca = acc(ai, aj), cb = acc(bi, bj);
    L[0] = 0.0; L[1] = 0.0; L[2] = 0.0; L[3] = 0.0;
    if (ord == 3) {
        double ua[NGP][4], ub[NGP][4];
        recon2<3, NGP, DIR == 0 ? kR : kT, DIR == 0 ? kL : kB, GPM_RECON_UNROLL>(T, acc, ai, aj, bi, bj, ua, ub);
        const double qa[4] = {ca.x, ca.y, ca.z, ca.w}, qb[4] = {cb.x, cb.y, cb.z, cb.w};
#pragma unroll
        for (int j = 0; j < NGP; ++j) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ua[j][k] = va1 * qa[k] + va2 * ua[j][k];
                ub[j][k] = vb1 * qb[k] + vb2 * ub[j][k];
            }
            double F[4];
            riemann<FLUX, DIR>(ua[j], ub[j], F);
            const double w = T->gw[j];
#pragma unroll
            for (int k = 0; k < 4; ++k) L[k] = L[k] + F[k] * w;
        }
    } else {
        // every Gauss point carries the two cell values: one Riemann solve
        double sa[4] = {ca.x, ca.y, ca.z, ca.w}, sb[4] = {cb.x, cb.y, cb.z, cb.w}, F[4];
        riemann<FLUX, DIR>(sa, sb, F);
#pragma unroll
        for (int j = 0; j < NGP; ++j) {
            const double w = T->gw[j];
#pragma unroll
            for (int k = 0; k < 4; ++k) L[k] = L[k] + F[k] * w;
        }
    }
}

struct MoodParams {
    int dmp, u2, u2tol;
    double dx, dy;
    int recon;  // kRecon*: the U2 second derivatives of POL_MOOD are its own
};

// shock_detector, mod_subroutine_MOOD.f90:191-238: compression (divV <
// -max(dx,dy)^2), Mach > 0.2 and a pressure jump (gradP > 5).  The three tests
// are evaluated one after the other and the first that fails ends the
// detector (same result as evaluating all three; most cells fail the Mach or
// compression test, so the divisions of the others are skipped).
__device__ __forceinline__ bool shock_mach(const Cell &c, double pc)
{
    const double Ca = dv(1.4 * pc, c.r);
    double Mach = dv(c.mx * c.mx + c.my * c.my, c.r * c.r);
    Mach = sqrt(dv(Mach, Ca));
    return Mach > 0.2;
}

__device__ __forceinline__ bool shock_compression(const Cell &c, const Cell &e, const Cell &w, const Cell &nn,
                                                  const Cell &s, double dx, double dy)
{
    double divV = dv(e.mx - w.mx, dx) + dv(nn.my - s.my, dy);
    divV = dv(0.5 * divV, c.r);
    divV = divV - dv(0.5 * (dv(c.mx * (e.r - w.r), dx) + dv(c.my * (nn.r - s.r), dy)), c.r * c.r);
    const double m = fmax2(dx, dy);
    return divV < -(m * m);
}

__device__ __forceinline__ bool shock_pressure_jump(double pL, double pR, double pB, double pT, double dx, double dy)
{
    const double gradP = 0.5 * (dv(dv(fabs(pR - pL), fmin2(pL, pR)), dx) + dv(dv(fabs(pT - pB), fmin2(pB, pT)), dy));
    return gradP > 5.0;
}

template <class Acc>
__device__ __forceinline__ bool shock_detector(const MoodParams &P, const Acc &acc, int i, int j)
{
    const Cell c = to_cell(acc(i, j));
    const Cell e = to_cell(acc(i + 1, j)), w = to_cell(acc(i - 1, j));
    const Cell nn = to_cell(acc(i, j + 1)), s = to_cell(acc(i, j - 1));
    if (!shock_mach(c, pressure(c.r, c.mx, c.my, c.E))) return false;
    if (!shock_compression(c, e, w, nn, s, P.dx, P.dy)) return false;
    return shock_pressure_jump(pressure(w.r, w.mx, w.my, w.E), pressure(e.r, e.mx, e.my, e.E),
                               pressure(s.r, s.mx, s.my, s.E), pressure(nn.r, nn.mx, nn.my, nn.E), P.dx, P.dy);
}

// DETECTION loop-1 decision of one cell (mod_subroutine_MOOD.f90:41-131) for a
// cell whose order is above 1: PAD, then the shock-gated DMP with plateau test
// and U2 smooth-extrema relaxation.  `cand` is the cell's candidate Uout.
// Kept out of line (it runs only for DMP failures); its arguments are passed by
// value so that the call does not make every thread build a stack frame.
//
// pol: POL_MOOD's second derivatives, dot_product(D2O2, rho(l+i-1:l+i+1, n+j))
// / dx**2 and the y form with D2O2 = (1, -2, 1) (mod_subroutine_MOOD.f90:86-87,
// constants.f90:43), instead of the GP_d2x/y_sp weights.
template <class Acc>
__device__ __noinline__ bool u2_relax(int u2tol, double dx, double dy, const DevTables *__restrict__ T,
                                      const Acc acc, int i, int j, bool pol)
{
    double D2x[5], D2y[5];  // (-1,0), (0,0), (1,0), (0,-1), (0,1)
    const int oi[5] = {-1, 0, 1, 0, 0}, oj[5] = {0, 0, 0, -1, 1};
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const int ci = i + oi[t], cj = j + oj[t];
        double ax = 0.0, ay = 0.0;
        if (pol) {
            ax = ax + 1.0 * acc(ci - 1, cj).x;
            ax = ax + -2.0 * acc(ci, cj).x;
            ax = ax + 1.0 * acc(ci + 1, cj).x;
            ay = ay + 1.0 * acc(ci, cj - 1).x;
            ay = ay + -2.0 * acc(ci, cj).x;
            ay = ay + 1.0 * acc(ci, cj + 1).x;
            ax = dv(ax, dx * dx);
            ay = dv(ay, dy * dy);
        } else {
            double q[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) q[k] = acc(ci + kSX[k], cj + kSY[k]).x;
#pragma unroll
            for (int k = 0; k < 5; ++k) ax = ax + T->d2x[k] * q[k];
#pragma unroll
            for (int k = 0; k < 5; ++k) ay = ay + T->d2y[k] * q[k];
        }
        D2x[t] = ax;
        D2y[t] = ay;
    }
    const double Xmin = fmin2(fmin2(fmin2(fmin2(D2x[0], D2x[1]), D2x[2]), D2x[3]), D2x[4]);
    const double Xmax = fmax2(fmax2(fmax2(fmax2(D2x[0], D2x[1]), D2x[2]), D2x[3]), D2x[4]);
    const double Ymin = fmin2(fmin2(fmin2(fmin2(D2y[0], D2y[1]), D2y[2]), D2y[3]), D2y[4]);
    const double Ymax = fmax2(fmax2(fmax2(fmax2(D2y[0], D2y[1]), D2y[2]), D2y[3]), D2y[4]);
    double delta = 0.0;
    if (u2tol) delta = fmax2(dx, dy);
    if ((Xmin * Xmax > -delta) && ((fmax2(fabs(Xmin), fabs(Xmax)) < delta) || (fabs(dv(Xmin, Xmax)) >= 0.5)))
        return true;
    if ((Ymin * Ymax > -delta) && ((fmax2(fabs(Ymin), fabs(Ymax)) < delta) || (fabs(dv(Ymin, Ymax)) >= 0.5)))
        return true;
    return false;
}

template <class Acc>
__device__ __forceinline__ bool decide(const MoodParams &P, const DevTables *__restrict__ T, const Acc &acc, int i,
                                       int j, const double cand[4])
{
    if (cand[0] <= 0. || isnan(cand[0])) return true;
    const double p = pressure(cand[0], cand[1], cand[2], cand[3]);
    if (p <= 0. || isnan(p)) return true;
    if (!P.dmp) return false;
    if (!shock_detector(P, acc, i, j)) return false;
    const double UB = acc(i, j - 1).x, UL = acc(i - 1, j).x, UM = acc(i, j).x;
    const double UR = acc(i + 1, j).x, UT = acc(i, j + 1).x;
    const double mn = fmin2(fmin2(fmin2(fmin2(UM, UL), UB), UT), UR);
    const double mx = fmax2(fmax2(fmax2(fmax2(UM, UL), UB), UT), UR);
    const double Mm = mx - mn;
    if (!(Mm > P.dx * P.dx * P.dx)) return false;
    if (!(cand[0] > mx || cand[0] < mn)) return false;
    if (P.u2 && u2_relax(P.u2tol, P.dx, P.dy, T, acc, i, j, P.recon == kReconPol)) return false;
    return true;
}

// The parts of the decision that depend on the stage input alone (the shock
// detector, and the U2 relaxation of a shocked cell whose 5-point density
// range exceeds the plateau gap) do not change during a stage's MOOD passes.
// The first pass records them per cell (decide_record), and the correction
// passes finish the decision from the record with the cell's new candidate
// (decide_with_flags); both return decide's boolean.
enum { kFlagShock = 1, kFlagU2 = 2 };

template <class Acc>
__device__ __forceinline__ bool decide_record(const MoodParams &P, const DevTables *__restrict__ T, const Acc &acc,
                                              int i, int j, const double cand[4], uint8_t &flags)
{
    bool pad = cand[0] <= 0. || isnan(cand[0]);
    if (!pad) {
        const double p = pressure(cand[0], cand[1], cand[2], cand[3]);
        pad = p <= 0. || isnan(p);
    }
    flags = 0;
    if (!P.dmp) return pad;
    if (!shock_detector(P, acc, i, j)) return pad;
    flags = kFlagShock;
    const double UB = acc(i, j - 1).x, UL = acc(i - 1, j).x, UM = acc(i, j).x;
    const double UR = acc(i + 1, j).x, UT = acc(i, j + 1).x;
    const double mn = fmin2(fmin2(fmin2(fmin2(UM, UL), UB), UT), UR);
    const double mx = fmax2(fmax2(fmax2(fmax2(UM, UL), UB), UT), UR);
    if (!(mx - mn > P.dx * P.dx * P.dx)) return pad;
    const bool relax = P.u2 && u2_relax(P.u2tol, P.dx, P.dy, T, acc, i, j, P.recon == kReconPol);
    if (relax) flags |= kFlagU2;
    if (pad) return true;
    if (!(cand[0] > mx || cand[0] < mn)) return false;
    return !relax;
}

template <class Acc>
__device__ __forceinline__ bool decide_with_flags(const MoodParams &P, const Acc &acc, int i, int j,
                                                  const double cand[4], uint8_t flags)
{
    if (cand[0] <= 0. || isnan(cand[0])) return true;
    const double p = pressure(cand[0], cand[1], cand[2], cand[3]);
    if (p <= 0. || isnan(p)) return true;
    if (!P.dmp || !(flags & kFlagShock)) return false;
    const double UB = acc(i, j - 1).x, UL = acc(i - 1, j).x, UM = acc(i, j).x;
    const double UR = acc(i + 1, j).x, UT = acc(i, j + 1).x;
    const double mn = fmin2(fmin2(fmin2(fmin2(UM, UL), UB), UT), UR);
    const double mx = fmax2(fmax2(fmax2(fmax2(UM, UL), UB), UT), UR);
    if (!(mx - mn > P.dx * P.dx * P.dx)) return false;
    if (!(cand[0] > mx || cand[0] < mn)) return false;
    return !(flags & kFlagU2);
}

// Cascade of a failing cell (mod_subroutine_MOOD.f90:159-163 or the full one).
__device__ __forceinline__ int lower_order(int ord, int cascade)
{
    if (cascade == 0) return ord == 3 ? 1 : (ord > 3 ? 3 : ord);
    return ord > 1 ? ord - 2 : ord;
}

// Per-cell order word (32 bits): bits 0-1 the order index (CellGPO = 2*idx+1),
// bits 2-3 the index of the order the cell had before its last lowering, bits
// 4-31 the MOOD pass in which the cascade last lowered it (0: the first pass,
// or the a priori order).  A correction pass reads the orders its cells had
// when the pass began: a cell lowered during the same pass p (version p) still
// counts with its previous order, so decisions and lowering share one grid
// barrier per pass.
typedef uint32_t order_t;
__host__ __device__ __forceinline__ order_t pack_order(int ord, int prev, int ver)
{
    return (order_t)(((unsigned)ver << 4) | (((prev - 1) >> 1) << 2) | ((ord - 1) >> 1));
}
__host__ __device__ __forceinline__ int order_of_word(order_t v) { return 2 * (v & 3) + 1; }
__device__ __forceinline__ int order_at_pass(order_t v, int p)
{
    return ((int)(v >> 4) == p) ? 2 * ((v >> 2) & 3) + 1 : 2 * (v & 3) + 1;
}

}  // namespace gpm
