// Internal declarations shared by the host parts of libgpmood.
#pragma once
#include <cstdint>

namespace gpm {

// Sphere stencil offsets in the reference's order (GP_init.f90:97-176).
extern const int kStencilX[25];
extern const int kStencilY[25];

// Quad-precision GP table precompute (GP_init.f90:26-357); tables in the
// reference's Fortran layouts.  Returns 0 on success.
int gp_tables(int mord, int ngp, double ell_o_dx, double lx, double ly, int lf, int nf,
              double *zT, double *d2x, double *d2y, double *gw, int32_t *ixiy);

// real(L_16/n, 8): a length divided by a cell count in quad precision, rounded
// to FP64 (global_variables.f90:19-21).
double quad_ratio(double L, int n);

inline int stencil_size(int ord) { return ord == 1 ? 1 : ord == 3 ? 5 : ord == 5 ? 13 : 25; }

}  // namespace gpm
