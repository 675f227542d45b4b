// Host entry points over the library's solver and adjoint device code, so
// that tests can check them on the CPU (TEST INFRASTRUCTURE ONLY).
#include "swe_adjoint.hpp"

extern "C" {

void shim_edge_flux(int solver, double nx, double ny, const double *QL, const double *QR,
                    double drytol, double g, double *fL, double *fR) {
    swe::Flux a, b;
    const double inv16g = 1.0 / (16.0 * g);
    if (solver == 1)
        swe::edge_flux<1>(nx, ny, QL[0], QL[1], QL[2], QL[3], QR[0], QR[1], QR[2], QR[3], drytol,
                          g, inv16g, a, b);
    else
        swe::edge_flux<0>(nx, ny, QL[0], QL[1], QL[2], QL[3], QR[0], QR[1], QR[2], QR[3], drytol,
                          g, inv16g, a, b);
    fL[0] = a.h; fL[1] = a.hu; fL[2] = a.hv; fL[3] = a.s;
    fR[0] = b.h; fR[1] = b.hu; fR[2] = b.hv; fR[3] = b.s;
}

void shim_edge_flux_adj(int solver, double nx, double ny, const double *QL, const double *QR,
                        double drytol, double g, const double *gfL, const double *gfR,
                        double *gQL, double *gQR) {
    const double inv16g = 1.0 / (16.0 * g);
    for (int i = 0; i < 4; ++i) gQL[i] = gQR[i] = 0.0;
    if (solver == 1)
        swe::edge_flux_adj<1>(nx, ny, QL[0], QL[1], QL[2], QL[3], QR[0], QR[1], QR[2], QR[3],
                              drytol, g, inv16g, gfL, gfR, gQL, gQR);
    else
        swe::edge_flux_adj<0>(nx, ny, QL[0], QL[1], QL[2], QL[3], QR[0], QR[1], QR[2], QR[3],
                              drytol, g, inv16g, gfL, gfR, gQL, gQR);
}
}
