/*
 * swe_oracle.c -- CPU restatement of the sam(oa)^2 SWE Euler step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP implementation.
 * Build with -ffp-contract=off (oracle/Makefile does) so that every
 * expression rounds exactly as written, in the reference's order.
 *
 * Parity pins for this file (see tests/test_oracle_*.py):
 *   - swo_geoclaw / swo_riemanntype, augmented and f-wave solvers:
 *     known-answer vectors produced by the reference's own riemannsolvers.f +
 *     c_bind_riemannsolvers.f90 compiled with amdflang (oracle/_ref/ref_kat),
 *     plus invariants (dry-dry, identical states, lake at rest, mirror
 *     symmetry).
 *   - lf_flux (global and local Lax-Friedrichs, flat-bed and well-balanced):
 *     VTU dumps of the reference SWE binary built with each of _LF_FLUX,
 *     _LLF_FLUX, _LF_BATH_FLUX and _LLF_BATH_FLUX (oracle/_ref/samoa_swe_*lf*),
 *     on wet dam breaks, a partly dry sloped bed and the half-dry dam break.
 *   - swo_leg_sizes: table dumped by oracle/_ref/ref_tables (the reference's
 *     parameter expressions compiled by amdflang).
 *   - swo_mesh_uniform / swo_step / swo_initial_dt: VTU dumps of the full
 *     reference SWE binary (oracle/_ref/samoa_swe) and closed forms.
 */
#include "swe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* SFC_data_types.f90:557 r_leg_sizes = [(sqrt(0.5)**i, i=-1,MAX_DEPTH)].
 * amdflang folds x**i (integer i) by binary exponentiation. */
void swo_leg_sizes(double out[64]) {
    const double s = sqrt(0.5);
    for (int k = -1; k <= 62; ++k) {
        int e = k < 0 ? -k : k;
        double r = 1.0, base = s;
        while (e) {
            if (e & 1) r = r * base;
            base = base * base;
            e >>= 1;
        }
        if (k < 0) r = 1.0 / r;
        out[k + 1] = r;
    }
}

static double leg(int k) {
    static double tab[64];
    static int init = 0;
    if (!init) {
        swo_leg_sizes(tab);
        init = 1;
    }
    return tab[k + 1];
}

/* SFC_data_types.f90:524 r_volumes = 0.5**(i+1) (exact powers of two). */
static double volume_of(int depth) { return ldexp(1.0, -(depth + 1)); }

static double dmin2(double a, double b) { return a < b ? a : b; }
static double dmax2(double a, double b) { return a > b ? a : b; }

/* riemannsolvers.f:507-618, maxiter = 1. */
void swo_riemanntype(double hL, double hR, double uL, double uR, double drytol,
                     double g, double *hm_out, double *s1m_out, double *s2m_out,
                     int *rare1, int *rare2) {
    double h_min = dmin2(hR, hL);
    double h_max = dmax2(hR, hL);
    double delu = uR - uL;
    double hm, s1m, s2m;

    if (h_min <= drytol) { /* :536-548 */
        hm = 0.0;
        s1m = ((uR + uL) - 2.0 * sqrt(g * hR)) + 2.0 * sqrt(g * hL);
        s2m = ((uR + uL) - 2.0 * sqrt(g * hR)) + 2.0 * sqrt(g * hL);
        if (hL <= 0.0) {
            *rare2 = 1;
            *rare1 = 0;
        } else {
            *rare1 = 1;
            *rare2 = 0;
        }
    } else {
        double F_min = delu + 2.0 * (sqrt(g * h_min) - sqrt(g * h_max));
        double F_max = delu + (h_max - h_min) *
                                  sqrt(((0.5 * g) * (h_max + h_min)) / (h_max * h_min));
        if (F_min > 0.0) { /* 2-rarefactions, :554-564 */
            double t = dmax2(0.0, -delu + 2.0 * (sqrt(g * hL) + sqrt(g * hR)));
            hm = (1.0 / (16.0 * g)) * (t * t);
            s1m = (uL + 2.0 * sqrt(g * hL)) - 3.0 * sqrt(g * hm);
            s2m = (uR - 2.0 * sqrt(g * hR)) + 3.0 * sqrt(g * hm);
            *rare1 = 1;
            *rare2 = 1;
        } else if (F_max <= 0.0) { /* 2 shocks, Newton on sqrt(h), :566-588 */
            double h0 = h_max;
            double gL = sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hL));
            double gR = sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hR));
            double F0 = (delu + (h0 - hL) * gL) + (h0 - hR) * gR;
            double dfdh = ((gL - (g * (h0 - hL)) / ((4.0 * (h0 * h0)) * gL)) + gR) -
                          (g * (h0 - hR)) / ((4.0 * (h0 * h0)) * gR);
            double slope = (2.0 * sqrt(h0)) * dfdh;
            double q = sqrt(h0) - F0 / slope;
            h0 = q * q;
            hm = h0;
            double u1m = uL - (hm - hL) * sqrt((0.5 * g) * (1.0 / hm + 1.0 / hL));
            double u2m = uR + (hm - hR) * sqrt((0.5 * g) * (1.0 / hm + 1.0 / hR));
            s1m = u1m - sqrt(g * hm);
            s2m = u2m + sqrt(g * hm);
            *rare1 = 0;
            *rare2 = 0;
        } else { /* one shock, one rarefaction, :589-612 */
            double h0 = h_min;
            double F0 = (delu + 2.0 * (sqrt(g * h0) - sqrt(g * h_max))) +
                        (h0 - h_min) * sqrt((0.5 * g) * (1.0 / h0 + 1.0 / h_min));
            double slope = (F_max - F0) / (h_max - h_min);
            h0 = h0 - F0 / slope;
            hm = h0;
            if (hL > hR) {
                s1m = (uL + 2.0 * sqrt(g * hL)) - 3.0 * sqrt(g * hm);
                s2m = (uL + 2.0 * sqrt(g * hL)) - sqrt(g * hm);
                *rare1 = 1;
                *rare2 = 0;
            } else {
                s2m = (uR - 2.0 * sqrt(g * hR)) + 3.0 * sqrt(g * hm);
                s1m = (uR - 2.0 * sqrt(g * hR)) + sqrt(g * hm);
                *rare2 = 1;
                *rare1 = 0;
            }
        }
    }
    *hm_out = hm;
    *s1m_out = s1m;
    *s2m_out = s2m;
}

/* 3x3 determinant as written at riemannsolvers.f:188-191 / :201-204. */
static double det3x3(const double A[3][3]) {
    double d1 = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
    double d2 = A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]);
    double d3 = A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    return (d1 - d2) + d3;
}

/* riemannsolvers.f:2-258 with maxiter = 1, meqn = mwaves = 3. sw[3], fw[3][3]
 * indexed fw[equation][wave]. sE1/sE2 are in/out as in the reference. */
static void riemann_aug_JCP(double hL, double hR, double huL, double huR,
                            double hvL, double hvR, double bL, double bR,
                            double uL, double uR, double vL, double vR,
                            double delphi, double *sE1, double *sE2,
                            double drytol, double g, double sw[3], double fw[3][3]) {
    (void)hvL;
    (void)hvR;
    double delh = hR - hL;
    double delhu = huR - huL;
    double delb = bR - bL;
    double hm, s1m, s2m;
    int rare1, rare2;
    double lambda[3], r[3][3], del[3], beta[3];

    swo_riemanntype(hL, hR, uL, uR, drytol, g, &hm, &s1m, &s2m, &rare1, &rare2);

    lambda[0] = dmin2(*sE1, s2m); /* :57-60 */
    lambda[2] = dmax2(*sE2, s1m);
    *sE1 = lambda[0];
    *sE2 = lambda[2];
    double hstarHLL = dmax2((((huL - huR) + *sE2 * hR) - *sE1 * hL) / (*sE2 - *sE1), 0.0);

    /* :92-104: rarefaction corrector disabled; lambda(2) is the mean. */
    for (int mw = 0; mw < 3; mw += 2) {
        r[0][mw] = 1.0;
        r[1][mw] = lambda[mw];
        r[2][mw] = lambda[mw] * lambda[mw];
    }
    lambda[1] = 0.5 * (lambda[0] + lambda[2]);
    if (fabs(lambda[1]) < 1.0e-20) lambda[1] = 0.0;
    r[0][1] = 0.0;
    r[1][1] = 0.0;
    r[2][1] = 1.0;

    /* steady-state wave, :108-185 (single iteration) */
    const double criticaltol = 1.0e-6;
    double hLstar = hL, hRstar = hR, uLstar = uL, uRstar = uR;
    double hbar = dmax2(0.5 * (hLstar + hRstar), 0.0);
    double s1s2bar = 0.25 * ((uLstar + uRstar) * (uLstar + uRstar)) - g * hbar;
    double s1s2tilde = dmax2(0.0, uLstar * uRstar) - g * hbar;

    int sonic = 0;
    if (fabs(s1s2bar) <= criticaltol) sonic = 1;
    if (s1s2bar * s1s2tilde <= criticaltol) sonic = 1;
    if ((s1s2bar * *sE1) * *sE2 <= criticaltol) sonic = 1;
    if (dmin2(fabs(*sE1), fabs(*sE2)) < criticaltol) sonic = 1;
    if (*sE1 < 0.0 && s1m > 0.0) sonic = 1;
    if (*sE2 > 0.0 && s2m < 0.0) sonic = 1;
    if ((uL + sqrt(g * hL)) * (uR + sqrt(g * hR)) < 0.0) sonic = 1;
    if ((uL - sqrt(g * hL)) * (uR - sqrt(g * hR)) < 0.0) sonic = 1;

    double deldelh;
    if (sonic)
        deldelh = -delb;
    else
        deldelh = ((delb * g) * hbar) / s1s2bar;
    if (*sE1 < -criticaltol && *sE2 > criticaltol) {
        deldelh = dmin2(deldelh, (hstarHLL * (*sE2 - *sE1)) / *sE2);
        deldelh = dmax2(deldelh, (hstarHLL * (*sE2 - *sE1)) / *sE1);
    } else if (*sE1 >= criticaltol) {
        deldelh = dmin2(deldelh, (hstarHLL * (*sE2 - *sE1)) / *sE1);
        deldelh = dmax2(deldelh, -hL);
    } else if (*sE2 <= -criticaltol) {
        deldelh = dmin2(deldelh, hR);
        deldelh = dmax2(deldelh, (hstarHLL * (*sE2 - *sE1)) / *sE2);
    }

    double deldelphi;
    if (sonic)
        deldelphi = -((g * hbar) * delb);
    else
        deldelphi = -((((delb * g) * hbar) * s1s2tilde) / s1s2bar);
    deldelphi = dmin2(deldelphi, g * dmax2(-(hLstar * delb), -(hRstar * delb)));
    deldelphi = dmax2(deldelphi, g * dmin2(-(hLstar * delb), -(hRstar * delb)));

    del[0] = delh - deldelh;
    del[1] = delhu;
    del[2] = delphi - deldelphi;

    /* Cramer's rule, :188-205 */
    double determinant = det3x3((const double(*)[3])r);
    for (int k = 0; k < 3; ++k) {
        double A[3][3];
        for (int mw = 0; mw < 3; ++mw)
            for (int m = 0; m < 3; ++m) A[m][mw] = (mw == k) ? del[m] : r[m][mw];
        beta[k] = det3x3((const double(*)[3])A) / determinant;
    }

    /* :245-254 */
    for (int mw = 0; mw < 3; ++mw) {
        sw[mw] = lambda[mw];
        fw[0][mw] = beta[mw] * r[1][mw];
        fw[1][mw] = beta[mw] * r[2][mw];
        fw[2][mw] = beta[mw] * r[1][mw];
    }
    fw[2][0] = fw[2][0] * vL;
    fw[2][2] = fw[2][2] * vR;
    fw[2][1] = (((hR * uR) * vR - (hL * uL) * vL) - fw[2][0]) - fw[2][2];
}

/* riemannsolvers.f:446-500 */
static void riemann_fwave(double hL, double hR, double huL, double huR,
                          double hvL, double hvR, double bL, double bR,
                          double uL, double uR, double vL, double vR,
                          double delphi, double s1, double s2, double g,
                          double sw[3], double fw[3][3]) {
    (void)hvL;
    (void)hvR;
    (void)uL;
    (void)uR;
    double delhu = huR - huL;
    double delb = bR - bL;
    double deldelphi = -(((g * 0.5) * (hR + hL)) * delb);
    double delphidecomp = delphi - deldelphi;
    double beta1 = (s2 * delhu - delphidecomp) / (s2 - s1);
    double beta2 = (delphidecomp - s1 * delhu) / (s2 - s1);
    sw[0] = s1;
    sw[1] = 0.5 * (s1 + s2);
    sw[2] = s2;
    fw[0][0] = beta1;
    fw[1][0] = beta1 * s1;
    fw[2][0] = beta1 * vL;
    fw[0][2] = beta2;
    fw[1][2] = beta2 * s2;
    fw[2][2] = beta2 * vR;
    fw[0][1] = 0.0;
    fw[1][1] = 0.0;
    fw[2][1] = ((huR * vR - huL * vL) - fw[2][0]) - fw[2][2];
}

/* c_bind_riemannsolvers.f90:16-176 (double precision). */
void swo_geoclaw(int solver, double hL, double hR, double huL, double huR,
                 double hvL, double hvR, double bL, double bR, double drytol,
                 double g, double netL[3], double netR[3], double *maxs) {
    double sw[3] = {0.0, 0.0, 0.0};
    double fw[3][3] = {{0.0}};
    double wall[3] = {1.0, 1.0, 1.0};
    double uL, uR, vL, vR;

    *maxs = 0.0;
    for (int i = 0; i < 3; ++i) netL[i] = netR[i] = 0.0;

    if (hL <= drytol && hR <= drytol) return; /* :55-57 */

    if (hR > drytol) { /* :60-69 */
        uR = huR / hR;
        vR = hvR / hR;
    } else {
        hR = 0.0;
        huR = 0.0;
        hvR = 0.0;
        uR = 0.0;
        vR = 0.0;
    }
    if (hL > drytol) { /* :71-80 */
        uL = huL / hL;
        vL = hvL / hL;
    } else {
        hL = 0.0;
        huL = 0.0;
        hvL = 0.0;
        uL = 0.0;
        vL = 0.0;
    }

    double hstar, s1m, s2m;
    int rare1, rare2;
    if (hR <= drytol) { /* :87-100 */
        swo_riemanntype(hL, hL, uL, -uL, drytol, g, &hstar, &s1m, &s2m, &rare1, &rare2);
        double hstartest = dmax2(hL, hstar);
        if (hstartest + bL < bR) {
            wall[1] = 0.0;
            wall[2] = 0.0;
            hR = hL;
            huR = -huL;
            bR = bL;
            uR = -uL;
            vR = vL;
        } else if (hL + bL < bR) {
            bR = hL + bL;
        }
    } else if (hL <= drytol) { /* :101-114 */
        swo_riemanntype(hR, hR, -uR, uR, drytol, g, &hstar, &s1m, &s2m, &rare1, &rare2);
        double hstartest = dmax2(hR, hstar);
        if (hstartest + bR < bL) {
            wall[0] = 0.0;
            wall[1] = 0.0;
            hL = hR;
            huL = -huR;
            bL = bR;
            uL = -uR;
            vL = vR;
        } else if (hR + bR < bL) {
            bL = hR + bR;
        }
    }

    /* :120 */
    double delphi = (((huR - huL) * (uL + uR) - (uL * uR) * (hR - hL)) +
                     ((0.5 * g) * (((bR + hR) - bL) - hL)) * (hR + hL)) -
                    ((0.5 * g) * (hR + hL)) * (bR - bL);

    /* :123-132 */
    double sL = uL - sqrt(g * hL);
    double sR = uR + sqrt(g * hR);
    double uhat = (sqrt(g * hL) * uL + sqrt(g * hR) * uR) / (sqrt(g * hR) + sqrt(g * hL));
    double chat = sqrt((g * 0.5) * (hR + hL));
    double sRoe1 = uhat - chat;
    double sRoe2 = uhat + chat;
    double sE1 = dmin2(sL, sRoe1);
    double sE2 = dmax2(sR, sRoe2);

    if (solver == SWO_SOLVER_FWAVE)
        riemann_fwave(hL, hR, huL, huR, hvL, hvR, bL, bR, uL, uR, vL, vR, delphi, sE1,
                      sE2, g, sw, fw);
    else
        riemann_aug_JCP(hL, hR, huL, huR, hvL, hvR, bL, bR, uL, uR, vL, vR, delphi,
                        &sE1, &sE2, drytol, g, sw, fw);

    for (int mw = 0; mw < 3; ++mw) { /* :148-153 */
        sw[mw] = sw[mw] * wall[mw];
        for (int m = 0; m < 3; ++m) fw[m][mw] = fw[m][mw] * wall[mw];
    }
    for (int m = 0; m < 3; ++m) { /* :156-167 */
        for (int mw = 0; mw < 3; ++mw) {
            if (sw[mw] < 0.0)
                netL[m] = netL[m] + fw[m][mw];
            else if (sw[mw] > 0.0)
                netR[m] = netR[m] + fw[m][mw];
            else {
                netL[m] = netL[m] + 0.5 * fw[m][mw];
                netR[m] = netR[m] + 0.5 * fw[m][mw];
            }
        }
    }
    /* :174-175 */
    double m0 = fabs(sw[0]), m1 = fabs(sw[1]), m2 = fabs(sw[2]);
    double mx = m0;
    if (m1 > mx) mx = m1;
    if (m2 > mx) mx = m2;
    *maxs = mx;
}

/* compute_lf_flux, SWE_euler_timestep.f90:359-446, in world coordinates:
 * Q = (eta, hu, hv, b), n the left cell's outward normal, f = (h, hu, hv,
 * max wave speed) per side.  Fortran's vector expressions are written out per
 * component in the order the source gives them (left to right; a unary minus
 * binds looser than a product, which does not change a value). */
static void lf_flux(int solver, const double n[2], const double QL[4], const double QR[4],
                    double drytol, double g, double fL[4], double fR[4]) {
    const int llf = solver == SWO_SOLVER_LLF || solver == SWO_SOLVER_LLF_BATH;
    double hL, hR, vL, vR, alpha;
    if (solver == SWO_SOLVER_LF_BATH || solver == SWO_SOLVER_LLF_BATH) {
        const int dryL = QL[0] - QL[3] < drytol, dryR = QR[0] - QR[3] < drytol;
        if (dryL || dryR) { /* :367-392, a wall against the dry side */
            for (int i = 0; i < 4; ++i) fL[i] = fR[i] = 0.0;
            if (dryL && dryR) {
            } else if (dryL) {
                hR = fmax(QR[0] - QR[3], 0.0);
                vR = n[0] * (QR[1] / (QR[0] - QR[3])) + n[1] * (QR[2] / (QR[0] - QR[3]));
                fR[3] = sqrt(g * hR) + fabs(vR);
                for (int i = 0; i < 2; ++i)
                    fR[1 + i] = (-((0.5 * vR) * QR[1 + i]) - (((0.5 * g) * hR) * hR) * n[i]) +
                                (0.5 * fR[3]) * QR[1 + i];
            } else {
                hL = fmax(QL[0] - QL[3], 0.0);
                vL = n[0] * (QL[1] / (QL[0] - QL[3])) + n[1] * (QL[2] / (QL[0] - QL[3]));
                fL[3] = sqrt(g * hL) + fabs(vL);
                for (int i = 0; i < 2; ++i)
                    fL[1 + i] = (((0.5 * vL) * QL[1 + i]) + (((0.5 * g) * hL) * hL) * n[i]) +
                                (0.5 * fL[3]) * QL[1 + i];
            }
            return;
        }
        hL = fmax(QL[0] - QL[3], 0.0); /* :394-418 */
        hR = fmax(QR[0] - QR[3], 0.0);
        vL = n[0] * (QL[1] / (QL[0] - QL[3])) + n[1] * (QL[2] / (QL[0] - QL[3]));
        vR = n[0] * (QR[1] / (QR[0] - QR[3])) + n[1] * (QR[2] / (QR[0] - QR[3]));
        fL[3] = sqrt(g * hL) + fabs(vL);
        fR[3] = sqrt(g * hR) + fabs(vR);
        alpha = llf ? (fL[3] > fR[3] ? fL[3] : fR[3]) : 100.0;
        fL[0] = 0.5 * ((hL * vL + hR * vR) + alpha * (QL[0] - QR[0]));
        fR[0] = -fL[0];
        for (int i = 0; i < 2; ++i) {
            fL[1 + i] = 0.5 * (((QL[1 + i] * vL + QR[1 + i] * vR) +
                                ((0.5 * g) * (hL * hL + hR * hR)) * n[i]) +
                               alpha * (QL[1 + i] - QR[1 + i]));
            fR[1 + i] = -fL[1 + i];
            const double src = (((0.25 * g) * (hL + hR)) * (QR[3] - QL[3])) * n[i];
            fL[1 + i] = fL[1 + i] + src;
            fR[1 + i] = fR[1 + i] + src;
        }
        return;
    }
    /* :422-445, the constant bed b = 0 */
    hL = fmax(QL[0] - 0.0, 0.0);
    hR = fmax(QR[0] - 0.0, 0.0);
    vL = n[0] * (QL[1] / (QL[0] - 0.0)) + n[1] * (QL[2] / (QL[0] - 0.0));
    vR = n[0] * (QR[1] / (QR[0] - 0.0)) + n[1] * (QR[2] / (QR[0] - 0.0));
    fL[3] = sqrt(g * (QL[0] - 0.0)) + fabs(vL);
    fR[3] = sqrt(g * (QR[0] - 0.0)) + fabs(vR);
    alpha = llf ? (fL[3] > fR[3] ? fL[3] : fR[3]) : 100.0;
    fL[0] = 0.5 * ((vL * hL + vR * hR) + alpha * (QL[0] - QR[0]));
    fR[0] = -fL[0];
    for (int i = 0; i < 2; ++i) {
        /* vL vL hL + vR vR hR is a scalar, added to both components */
        fL[1 + i] = 0.5 * ((((vL * vL) * hL + (vR * vR) * hR) +
                            ((0.5 * g) * (hL * hL + hR * hR)) * n[i]) +
                           alpha * (QL[1 + i] - QR[1 + i]));
        fR[1 + i] = -fL[1 + i];
    }
}

/* SWE_euler_timestep.f90:449-481 */
void swo_flux(int solver, const double n[2], const double QL[4], const double QR[4],
              double drytol, double g, double fL[4], double fR[4]) {
    if (solver >= SWO_SOLVER_LF) {
        lf_flux(solver, n, QL, QR, drytol, g, fL, fR);
        return;
    }
    /* transform_matrix(1,:) = n, (2,:) = (-n2, n1); p' = T p */
    double pL1 = n[0] * QL[1] + n[1] * QL[2];
    double pL2 = (-n[1]) * QL[1] + n[0] * QL[2];
    double pR1 = n[0] * QR[1] + n[1] * QR[2];
    double pR2 = (-n[1]) * QR[1] + n[0] * QR[2];
    double hL = QL[0] - QL[3];
    double hR = QR[0] - QR[3];
    double netL[3], netR[3], s;
    swo_geoclaw(solver, hL, hR, pL1, pR1, pL2, pR2, QL[3], QR[3], drytol, g, netL, netR,
                &s);
    /* flux%p = net(2:3) T */
    fL[0] = netL[0];
    fL[1] = netL[1] * n[0] + netL[2] * (-n[1]);
    fL[2] = netL[1] * n[1] + netL[2] * n[0];
    fL[3] = s;
    fR[0] = netR[0];
    fR[1] = netR[1] * n[0] + netR[2] * (-n[1]);
    fR[2] = netR[1] * n[1] + netR[2] * n[0];
    fR[3] = s;
}

/* SFC_data_types.f90:627-686: Jacobian of plotter type p (rounded to
 * integers, :645-647), normals n_i = normalize(e_i J^-1) with
 * e = (-1,0), (1,1), (0,-1). */
static void jacobian_of(int p, double J[2][2]) {
    const double PI = 4.0 * atan(1.0);
    double a = PI / 4.0 * (double)abs(p);
    /* J[row][col] */
    if (p > 0) {
        J[0][0] = cos(a);
        J[1][0] = sin(a);
        J[0][1] = -sin(a);
        J[1][1] = cos(a);
    } else {
        J[0][0] = -sin(a);
        J[1][0] = cos(a);
        J[0][1] = cos(a);
        J[1][1] = sin(a);
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) J[i][j] = trunc(1.5 * J[i][j]);
}

void swo_normals(int p, double n[3][2]) {
    double J[2][2], Ji[2][2];
    jacobian_of(p, J);
    double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    double inv = 1.0 / det;
    Ji[0][0] = inv * J[1][1];
    Ji[1][0] = inv * (-J[1][0]);
    Ji[0][1] = inv * (-J[0][1]);
    Ji[1][1] = inv * J[0][0];
    const double e[3][2] = {{-1.0, 0.0}, {1.0, 1.0}, {0.0, -1.0}};
    for (int i = 0; i < 3; ++i) {
        /* row vector times matrix: (e J^-1)_j = e_1 Ji[0][j] + e_2 Ji[1][j] */
        double x = e[i][0] * Ji[0][0] + e[i][1] * Ji[1][0];
        double y = e[i][0] * Ji[0][1] + e[i][1] * Ji[1][1];
        double f = 1.0 / sqrt(x * x + y * y);
        n[i][0] = f * x;
        n[i][1] = f * y;
    }
}

/* ---------------------------------------------------------------------- */
/* Uniform mesh: SFC_node_traversal.f90:90-145 (roots, recursion) and
 * :252-292 (child plotter types). */

static const int8_t plotter_child[17] = {3, 2, 1, 8, 7, 6, 5, 4, 0,
                                         -6, -7, -8, -1, -2, -3, -4, -5};

typedef struct {
    int8_t *plotter;
    double *coords;
    int64_t next;
} mesh_ctx;

static void recurse(mesh_ctx *m, int p, const double c[6], int levels) {
    if (levels == 0) {
        int64_t i = m->next++;
        m->plotter[i] = (int8_t)p;
        memcpy(m->coords + 6 * i, c, 6 * sizeof(double));
        return;
    }
    double mid[2] = {0.5 * (c[0] + c[4]), 0.5 * (c[1] + c[5])};
    double first[6] = {c[0], c[1], mid[0], mid[1], c[2], c[3]};
    double second[6] = {c[2], c[3], mid[0], mid[1], c[4], c[5]};
    int p1 = plotter_child[p + 8];
    int p2 = -plotter_child[-p + 8];
    recurse(m, p1, first, levels - 1);
    recurse(m, p2, second, levels - 1);
}

typedef struct {
    int64_t kx, ky;
    int64_t slot;
} edge_key;

static int cmp_key(const void *a, const void *b) {
    const edge_key *x = (const edge_key *)a, *y = (const edge_key *)b;
    if (x->kx != y->kx) return x->kx < y->kx ? -1 : 1;
    if (x->ky != y->ky) return x->ky < y->ky ? -1 : 1;
    return x->slot < y->slot ? -1 : (x->slot > y->slot);
}

int64_t swo_mesh_uniform(int depth, int8_t *plotter, double *coords, int32_t *nbr) {
    int64_t n = (int64_t)1 << (depth + 1);
    mesh_ctx m = {plotter, coords, 0};
    const double r1[6] = {0.0, 0.0, 1.0, 0.0, 1.0, 1.0};
    const double r2[6] = {1.0, 1.0, 0.0, 1.0, 0.0, 0.0};
    recurse(&m, -2, r1, depth);
    recurse(&m, -6, r2, depth);

    /* neighbours: match edge midpoints (dyadic, exact in integers) */
    edge_key *keys = (edge_key *)malloc((size_t)(3 * n) * sizeof(edge_key));
    if (!keys) return -1;
    const double scale = ldexp(1.0, depth / 2 + 3);
    for (int64_t c = 0; c < n; ++c) {
        const double *x = coords + 6 * c;
        for (int k = 0; k < 3; ++k) {
            int a = (k + 1) % 3, b = (k + 2) % 3; /* edge k opposite node k */
            double mx = 0.5 * (x[2 * a] + x[2 * b]);
            double my = 0.5 * (x[2 * a + 1] + x[2 * b + 1]);
            keys[3 * c + k].kx = (int64_t)llround(mx * scale);
            keys[3 * c + k].ky = (int64_t)llround(my * scale);
            keys[3 * c + k].slot = 3 * c + k;
        }
    }
    qsort(keys, (size_t)(3 * n), sizeof(edge_key), cmp_key);
    for (int64_t i = 0; i < 3 * n; ++i) nbr[i] = -1;
    for (int64_t i = 0; i + 1 < 3 * n; ++i) {
        if (keys[i].kx == keys[i + 1].kx && keys[i].ky == keys[i + 1].ky) {
            int64_t s0 = keys[i].slot, s1 = keys[i + 1].slot;
            nbr[s0] = (int32_t)(s1 / 3);
            nbr[s1] = (int32_t)(s0 / 3);
            ++i;
        }
    }
    free(keys);
    return n;
}

/* ---------------------------------------------------------------------- */
/* Euler step */

typedef struct {
    double n[17][3][2];
} normal_table;

static const normal_table *normals(void) {
    static normal_table t;
    static int init = 0;
    if (!init) {
        for (int p = -8; p <= 8; ++p)
            if (p != 0) swo_normals(p, t.n[p + 8]);
        init = 1;
    }
    return &t;
}

/* flux seen by cell c across its forward edge k (SWE_euler_timestep.f90:196-255,
 * scheduling Tools_adaptive_traversal.f90:606-617, boundary :177-206). */
static void edge_flux(int64_t c, int k, const int8_t *plotter, const int32_t *nbr, int bc,
                      int solver, int forward, double drytol, double g, const double *eta,
                      const double *hu, const double *hv, const double *b, double f[4]) {
    const normal_table *nt = normals();
    double Qc[4] = {eta[c], hu[c], hv[c], b[c]};
    int32_t o = nbr[3 * c + k];
    double fL[4], fR[4];
    if (o < 0) {
        const double *n = nt->n[plotter[c] + 8][k];
        double G[4] = {Qc[0], Qc[1], Qc[2], Qc[3]};
        if (bc == SWO_BC_WALL) {
            double dp = Qc[1] * n[0] + Qc[2] * n[1];
            G[1] = Qc[1] - (2.0 * dp) * n[0];
            G[2] = Qc[2] - (2.0 * dp) * n[1];
        }
        swo_flux(solver, n, Qc, G, drytol, g, fL, fR);
        memcpy(f, fL, sizeof fL);
        return;
    }
    double Qo[4] = {eta[o], hu[o], hv[o], b[o]};
    int c_is_left = forward ? (c < o) : (c > o);
    if (c_is_left) {
        const double *n = nt->n[plotter[c] + 8][k];
        swo_flux(solver, n, Qc, Qo, drytol, g, fL, fR);
        memcpy(f, fL, sizeof fL);
    } else {
        int ko = 0;
        while (ko < 3 && nbr[3 * (int64_t)o + ko] != (int32_t)c) ++ko;
        const double *n = nt->n[plotter[o] + 8][ko];
        swo_flux(solver, n, Qo, Qc, drytol, g, fL, fR);
        memcpy(f, fR, sizeof fR);
    }
}

/* cell_update_op + volume_op for one cell; returns the CFL candidate. */
static double cell_update(int64_t c, const int8_t *depth, const int8_t *plotter,
                          const int32_t *nbr, double S, double g, double drytol, int dmin,
                          int dmax, int bc, int solver, int forward, double dt,
                          const double *eta, const double *hu, const double *hv,
                          const double *b, double *eta_out, double *hu_out,
                          double *hv_out, int8_t *refine, int64_t *nref) {
    double f[3][4];
    for (int k = 0; k < 3; ++k)
        edge_flux(c, k, plotter, nbr, bc, solver, forward, drytol, g, eta, hu, hv, b, f[k]);

    int d = depth[c];
    /* volume_op, SWE_euler_timestep.f90:326-352 */
    double volume = (S * S) * volume_of(d);
    double len[3] = {S * leg(d), S * leg(d - 1), S * leg(d)};
    /* edge-index order of the current traversal: reversed cells swap legs */
    int ord[3] = {0, 1, 2};
    if (!forward) {
        ord[0] = 2;
        ord[2] = 0;
    }
    double dh = 0.0, dp1 = 0.0, dp2 = 0.0, sv = 0.0;
    for (int i = 0; i < 3; ++i) {
        dh = dh + len[i] * f[ord[i]][0];
        dp1 = dp1 + len[i] * f[ord[i]][1];
        dp2 = dp2 + len[i] * f[ord[i]][2];
        sv = sv + len[i] * f[ord[i]][3];
    }
    int8_t rf = 0;
    double dQ_norm = fabs(dh);
    double thr = 5.0 * S * leg(dmax);
    if (d < dmax && dQ_norm > thr) {
        rf = 1;
        ++*nref;
    } else if (d > dmin && dQ_norm < thr / 8.0) {
        rf = -1;
    }
    refine[c] = rf;
    double cand = volume / sv;
    double scale = -dt / volume;
    dh = scale * dh;
    dp1 = scale * dp1;
    dp2 = scale * dp2;

    /* cell_update_op, :271-283 */
    double qh = eta[c], q1 = hu[c], q2 = hv[c], qb = b[c];
    if (qh < qb + drytol && dh > 0.0) {
        qh = qb + drytol;
        q1 = 0.0;
        q2 = 0.0;
    }
    qh = qh + dh;
    q1 = q1 + dp1;
    q2 = q2 + dp2;
    if (qh < qb + drytol) {
        qh = dmin2(qb, 0.0);
        q1 = 0.0;
        q2 = 0.0;
    }
    eta_out[c] = qh;
    hu_out[c] = q1;
    hv_out[c] = q2;
    return cand;
}

double swo_step(int64_t n, const int8_t *depth, const int8_t *plotter,
                const int32_t *nbr, double scaling, double g, double drytol, int dmin,
                int dmax, int bc, int solver, int forward, double dt, const double *eta,
                const double *hu, const double *hv, const double *b, double *eta_out,
                double *hu_out, double *hv_out, int8_t *refine, int64_t *n_refine) {
    double dt_new = HUGE_VAL > 0 ? 1.7976931348623157e308 : 0.0; /* huge(1.0_SR) */
    int64_t nref = 0;
    for (int64_t c = 0; c < n; ++c) {
        double cand = cell_update(c, depth, plotter, nbr, scaling, g, drytol, dmin, dmax, bc,
                                  solver, forward, dt, eta, hu, hv, b, eta_out, hu_out,
                                  hv_out, refine, &nref);
        dt_new = dmin2(dt_new, cand);
    }
    *n_refine = nref;
    return dt_new;
}

double swo_step_omp(int64_t n, const int8_t *depth, const int8_t *plotter,
                    const int32_t *nbr, double scaling, double g, double drytol, int dmin,
                    int dmax, int bc, int solver, int forward, double dt,
                    const double *eta, const double *hu, const double *hv,
                    const double *b, double *eta_out, double *hu_out, double *hv_out,
                    int8_t *refine, int64_t *n_refine) {
    double dt_new = 1.7976931348623157e308;
    int64_t nref = 0;
    (void)normals(); /* initialise the static tables before the parallel region */
    (void)leg(0);
#pragma omp parallel for reduction(min : dt_new) reduction(+ : nref) schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        int64_t local = 0;
        double cand = cell_update(c, depth, plotter, nbr, scaling, g, drytol, dmin, dmax, bc,
                                  solver, forward, dt, eta, hu, hv, b, eta_out, hu_out,
                                  hv_out, refine, &local);
        nref += local;
        if (cand < dt_new) dt_new = cand;
    }
    *n_refine = nref;
    return dt_new;
}

/* SWE_initialize.f90:327-335 and :240 */
double swo_initial_dt(int64_t n, const int8_t *depth, const double *eta, const double *hu,
                      const double *hv, const double *b, double g, double courant) {
    double dt_new = 1.7976931348623157e308;
    for (int64_t c = 0; c < n; ++c) {
        double h = eta[c] - b[c];
        double s = 0.0;
        if (h > 0.0)
            s = sqrt(g * h) + sqrt((hu[c] * hu[c] + hv[c] * hv[c]) / (h * h));
        int d = depth[c];
        double sum_len = (leg(d) + leg(d - 1)) + leg(d);
        double cand = volume_of(d) / (sum_len * s);
        dt_new = dmin2(dt_new, cand);
    }
    return courant * dt_new;
}
