// swe_riemann_tol.hpp -- the GeoClaw solvers of swe_riemann.hpp in toleranced
// arithmetic (device only).
//
// Same scheme, same branches and the same operations as geoclaw<SOLVER> and
// riemanntype (c_bind_riemannsolvers.f90:16-176, riemannsolvers.f:2-258,
// 446-500, 507-618), but not rounded as the reference rounds them:
//   - divisions by a value used more than once share one reciprocal (a
//     cell's depth, det of the Cramer system, sE2 - sE1), and every
//     reciprocal is the hardware estimate refined by two Newton steps;
//   - square roots are the hardware reciprocal square root refined by one
//     Newton step;
//   - the velocities come from per-cell world velocities rotated into the
//     edge frame (u = n . (hu/h, hv/h)) instead of rotated momenta divided
//     by the depth, and momenta from h u.
// Sums and products round as written (no fused multiply-adds): the
// well-balancing terms (delphi and deldelphi of a lake at rest) cancel to
// their rounding errors, and rounding them as the exact path does keeps
// those errors, the largest part of a step's update in still water, the
// same.  Each other result stays within a few ulps of its correctly rounded
// value.
#pragma once

#include <hip/hip_runtime.h>

#include "swe_riemann.hpp"

namespace swe {
namespace tol {

// 1/x to about one ulp: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// sqrt(x) for x >= 0 to a few ulps (0 for x <= 0): the hardware reciprocal
// square root and one Newton step on the root and on half its reciprocal.
__device__ __forceinline__ double sqrt0(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r;
    double h = 0.5 * r;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    const double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x > 0.0 ? g : 0.0;
}

// riemanntype (riemannsolvers.f:507-618, maxiter = 1).
__device__ __forceinline__ void riemanntype(double hL, double hR, double uL, double uR,
                                            double sqL, double sqR, double drytol, double g,
                                            double inv16g, double &hm, double &s1m, double &s2m) {
    const bool r_min = hR < hL;
    const double h_min = r_min ? hR : hL;
    const double h_max = r_min ? hL : hR;
    const double sqh_min = r_min ? sqR : sqL;
    const double sqh_max = r_min ? sqL : sqR;
    const double delu = uR - uL;
    if (h_min <= drytol) {
        hm = 0.0;
        s1m = ((uR + uL) - 2.0 * sqR) + 2.0 * sqL;
        s2m = s1m;
        return;
    }
    const double F_min = delu + 2.0 * (sqh_min - sqh_max);
    const double F_max =
        delu + (h_max - h_min) * sqrt0(((0.5 * g) * (h_max + h_min)) * rcp(h_max * h_min));
    if (F_min > 0.0) {  // two rarefactions
        const double t = dmax2(0.0, -delu + 2.0 * (sqL + sqR));
        hm = inv16g * (t * t);
        const double sqm = sqrt0(g * hm);
        s1m = (uL + 2.0 * sqL) - 3.0 * sqm;
        s2m = (uR - 2.0 * sqR) + 3.0 * sqm;
    } else if (F_max <= 0.0) {  // two shocks: one Newton step on sqrt(h)
        const double rL = rcp(hL), rR = rcp(hR);
        const double h0 = h_max, r0 = r_min ? rL : rR;
        const double gL = sqrt0((0.5 * g) * (r0 + rL));
        const double gR = sqrt0((0.5 * g) * (r0 + rR));
        const double F0 = (delu + (h0 - hL) * gL) + (h0 - hR) * gR;
        const double q4 = 0.25 * (r0 * r0);
        const double dfdh = ((gL - (g * (h0 - hL)) * q4 * rcp(gL)) + gR) -
                            (g * (h0 - hR)) * q4 * rcp(gR);
        const double sq0 = sqrt0(h0);
        const double slope = (2.0 * sq0) * dfdh;
        const double q = sq0 - F0 * rcp(slope);
        hm = q * q;
        const double rm = rcp(hm);
        const double u1m = uL - (hm - hL) * sqrt0((0.5 * g) * (rm + rL));
        const double u2m = uR + (hm - hR) * sqrt0((0.5 * g) * (rm + rR));
        const double sqm = sqrt0(g * hm);
        s1m = u1m - sqm;
        s2m = u2m + sqm;
    } else {  // one shock, one rarefaction: one secant step from h_min
        const double F0 = F_min;
        hm = h_min - F0 * ((h_max - h_min) * rcp(F_max - F0));
        const double sqm = sqrt0(g * hm);
        if (hL > hR) {
            s1m = (uL + 2.0 * sqL) - 3.0 * sqm;
            s2m = (uL + 2.0 * sqL) - sqm;
        } else {
            s2m = (uR - 2.0 * sqR) + 3.0 * sqm;
            s1m = (uR - 2.0 * sqR) + sqm;
        }
    }
}

// wall_hstar of swe_riemann.hpp: hm of riemanntype(h, h, a, c) for the wall
// test.
__device__ __forceinline__ double wall_hstar(double h, double a, double c, double sq, double g,
                                             double inv16g) {
    const double delu = c - a;
    if (delu > 0.0) {
        const double t = dmax2(0.0, -delu + 4.0 * sq);
        return inv16g * (t * t);
    }
    const double gL = sqrt0(g * rcp(h));
    const double sq0 = sqrt0(h);
    const double q = sq0 - delu * rcp((4.0 * sq0) * gL);
    return q * q;
}

// geoclaw<SOLVER> in the edge frame from depths h, normal and tangential
// velocities u, v, bathymetry b and the wet sides' sqrt(g h) (0 on a dry
// side).  net[0..2] left-going, [3..5] right-going.
template <int SOLVER>
__device__ __forceinline__ void geoclaw(double hL, double hR, double uL, double uR, double vL,
                                        double vR, double bL, double bR, double sqL, double sqR,
                                        double drytol, double g, double inv16g, double netL[3],
                                        double netR[3], double &maxs) {
    netL[0] = netL[1] = netL[2] = 0.0;
    netR[0] = netR[1] = netR[2] = 0.0;
    maxs = 0.0;
    const bool dryR = hR <= drytol, dryL = hL <= drytol;
    if (dryL && dryR) return;  // :55-57
    if (dryR) {  // :60-69 (u, v and sqrt(g h) of a dry side are 0 already)
        hR = 0.0;
        uR = vR = sqR = 0.0;
    }
    if (dryL) {
        hL = 0.0;
        uL = vL = sqL = 0.0;
    }
    bool walled = false;
    double wall0 = 1.0, wall1 = 1.0, wall2 = 1.0;
    if (dryR || dryL) {  // :87-114, one side dry
        const double hw = dryR ? hL : hR;
        const double bw = dryR ? bL : bR, bd = dryR ? bR : bL;
        const double sqw = dryR ? sqL : sqR;
        const double hstar = wall_hstar(hw, dryR ? uL : -uR, dryR ? -uL : uR, sqw, g, inv16g);
        const double hstartest = dmax2(hw, hstar);
        if (hstartest + bw < bd) {
            walled = true;
            wall1 = 0.0;
            if (dryR) {
                wall2 = 0.0;
                hR = hL;
                bR = bL;
                uR = -uL;
                vR = vL;
                sqR = sqL;
            } else {
                wall0 = 0.0;
                hL = hR;
                bL = bR;
                uL = -uR;
                vL = vR;
                sqL = sqR;
            }
        } else if (hw + bw < bd) {
            if (dryR)
                bR = hL + bL;
            else
                bL = hR + bR;
        }
    }
    const double huL = hL * uL, huR = hR * uR;
    // :120
    const double delphi = (((huR - huL) * (uL + uR) - (uL * uR) * (hR - hL)) +
                           ((0.5 * g) * (((bR + hR) - bL) - hL)) * (hR + hL)) -
                          ((0.5 * g) * (hR + hL)) * (bR - bL);
    // :123-132
    const double sL = uL - sqL;
    const double sR = uR + sqR;
    const double uhat = (sqL * uL + sqR * uR) * rcp(sqR + sqL);
    const double chat = sqrt0((g * 0.5) * (hR + hL));
    double sE1 = dmin2(sL, uhat - chat);
    double sE2 = dmax2(sR, uhat + chat);

    double sw0, sw1, sw2;
    double f00, f10, f20, f01, f11, f21, f02, f12, f22;
    const double delhu = huR - huL;
    const double delb = bR - bL;
    if (SOLVER == 1) {  // riemann_fwave, riemannsolvers.f:446-500
        const double deldelphi = -(((g * 0.5) * (hR + hL)) * delb);
        const double dec = delphi - deldelphi;
        const double rdet = rcp(sE2 - sE1);
        const double beta1 = (sE2 * delhu - dec) * rdet;
        const double beta2 = (dec - sE1 * delhu) * rdet;
        sw0 = sE1;
        sw1 = 0.5 * (sE1 + sE2);
        sw2 = sE2;
        f00 = beta1;
        f10 = beta1 * sE1;
        f20 = beta1 * vL;
        f02 = beta2;
        f12 = beta2 * sE2;
        f22 = beta2 * vR;
        f01 = 0.0;
        f11 = 0.0;
        f21 = ((huR * vR - huL * vL) - f20) - f22;
    } else {  // riemann_aug_JCP, riemannsolvers.f:2-258
        const double delh = hR - hL;
        double hm, s1m, s2m;
        riemanntype(hL, hR, uL, uR, sqL, sqR, drytol, g, inv16g, hm, s1m, s2m);
        const double l0 = dmin2(sE1, s2m);  // :57-60
        const double l2 = dmax2(sE2, s1m);
        sE1 = l0;
        sE2 = l2;
        const double rspan = rcp(sE2 - sE1);
        const double hstarHLL = dmax2((((huL - huR) + sE2 * hR) - sE1 * hL) * rspan, 0.0);
        double l1 = 0.5 * (l0 + l2);  // :98-100
        if (fabs(l1) < 1.0e-20) l1 = 0.0;
        const double r20 = l0 * l0, r22 = l2 * l2;

        const double criticaltol = 1.0e-6;  // :108-185, one iteration
        const double hbar = dmax2(0.5 * (hL + hR), 0.0);
        const double s1s2bar = 0.25 * ((uL + uR) * (uL + uR)) - g * hbar;
        const double s1s2tilde = dmax2(0.0, uL * uR) - g * hbar;
        bool sonic = false;
        if (fabs(s1s2bar) <= criticaltol) sonic = true;
        if (s1s2bar * s1s2tilde <= criticaltol) sonic = true;
        if ((s1s2bar * sE1) * sE2 <= criticaltol) sonic = true;
        if (dmin2(fabs(sE1), fabs(sE2)) < criticaltol) sonic = true;
        if (sE1 < 0.0 && s1m > 0.0) sonic = true;
        if (sE2 > 0.0 && s2m < 0.0) sonic = true;
        if ((uL + sqL) * (uR + sqR) < 0.0) sonic = true;
        if ((uL - sqL) * (uR - sqR) < 0.0) sonic = true;

        // the steady-state terms divide as the exact path does: in still
        // water delphi - deldelphi and delh - deldelh cancel to rounding
        // errors, and those must be the exact path's
        double deldelh = sonic ? -delb : ((delb * g) * hbar) / s1s2bar;
        const double hspan = hstarHLL * (sE2 - sE1);
        if (sE1 < -criticaltol && sE2 > criticaltol) {
            deldelh = dmin2(deldelh, hspan * rcp(sE2));
            deldelh = dmax2(deldelh, hspan * rcp(sE1));
        } else if (sE1 >= criticaltol) {
            deldelh = dmin2(deldelh, hspan * rcp(sE1));
            deldelh = dmax2(deldelh, -hL);
        } else if (sE2 <= -criticaltol) {
            deldelh = dmin2(deldelh, hR);
            deldelh = dmax2(deldelh, hspan * rcp(sE2));
        }
        double deldelphi =
            sonic ? -((g * hbar) * delb) : -((((delb * g) * hbar) * s1s2tilde) / s1s2bar);
        const double pL = -(hL * delb), pR = -(hR * delb);
        deldelphi = dmin2(deldelphi, g * dmax2(pL, pR));
        deldelphi = dmax2(deldelphi, g * dmin2(pL, pR));

        const double d0 = delh - deldelh;
        const double d1 = delhu;
        const double d2 = delphi - deldelphi;

        // Cramer's rule with r = [(1,l0,l0^2), (0,0,1), (1,l2,l2^2)]
        const double rdet = rcp(-l2 + l0);
        const double beta0 = (-(d0 * l2) + d1) * rdet;
        const double beta1 =
            ((d1 * r22 - l2 * d2) - d0 * (l0 * r22 - l2 * r20) + (l0 * d2 - d1 * r20)) * rdet;
        const double beta2 = (-d1 + d0 * l0) * rdet;

        sw0 = l0;
        sw1 = l1;
        sw2 = l2;
        f00 = beta0 * l0;
        f10 = beta0 * r20;
        f20 = (beta0 * l0) * vL;
        f01 = 0.0;
        f11 = beta1;
        f02 = beta2 * l2;
        f12 = beta2 * r22;
        f22 = (beta2 * l2) * vR;
        f21 = ((huR * vR - huL * vL) - f20) - f22;  // :254
    }

    if (walled) {  // :148-153
        sw0 *= wall0;
        sw1 *= wall1;
        sw2 *= wall2;
        f00 *= wall0;
        f10 *= wall0;
        f20 *= wall0;
        f01 *= wall1;
        f11 *= wall1;
        f21 *= wall1;
        f02 *= wall2;
        f12 *= wall2;
        f22 *= wall2;
    }

    // :156-167 net updates by wave-speed sign, 50/50 at exactly zero
    const double a0 = sw0 < 0.0 ? 1.0 : (sw0 > 0.0 ? 0.0 : 0.5);
    const double a1 = sw1 < 0.0 ? 1.0 : (sw1 > 0.0 ? 0.0 : 0.5);
    const double a2 = sw2 < 0.0 ? 1.0 : (sw2 > 0.0 ? 0.0 : 0.5);
    const double c0 = sw0 > 0.0 ? 1.0 : (sw0 < 0.0 ? 0.0 : 0.5);
    const double c1 = sw1 > 0.0 ? 1.0 : (sw1 < 0.0 ? 0.0 : 0.5);
    const double c2 = sw2 > 0.0 ? 1.0 : (sw2 < 0.0 ? 0.0 : 0.5);
    netL[0] = (a0 * f00 + a1 * f01) + a2 * f02;
    netL[1] = (a0 * f10 + a1 * f11) + a2 * f12;
    netL[2] = (a0 * f20 + a1 * f21) + a2 * f22;
    netR[0] = (c0 * f00 + c1 * f01) + c2 * f02;
    netR[1] = (c0 * f10 + c1 * f11) + c2 * f12;
    netR[2] = (c0 * f20 + c1 * f21) + c2 * f22;
    maxs = dmax2(dmax2(fabs(sw0), fabs(sw1)), fabs(sw2));  // :174-175
}

// compute_geoclaw_flux (SWE_euler_timestep.f90:449-481) from per-cell
// quantities: depth h, world velocity (wu, wv) = (hu, hv)/h (0 when dry),
// bathymetry b and sqrt(g h) (0 when dry); n the left cell's outward normal.
template <int SOLVER>
__device__ __forceinline__ void edge_flux(double nx, double ny, double hL, double wuL,
                                          double wvL, double bL, double sqL, double hR,
                                          double wuR, double wvR, double bR, double sqR,
                                          double drytol, double g, double inv16g, Flux &fL,
                                          Flux &fR) {
    const double uL = nx * wuL + ny * wvL;
    const double vL = nx * wvL - ny * wuL;
    const double uR = nx * wuR + ny * wvR;
    const double vR = nx * wvR - ny * wuR;
    double netL[3], netR[3], s;
    geoclaw<SOLVER>(hL, hR, uL, uR, vL, vR, bL, bR, sqL, sqR, drytol, g, inv16g, netL, netR, s);
    fL.h = netL[0];
    fL.hu = netL[1] * nx - netL[2] * ny;
    fL.hv = netL[1] * ny + netL[2] * nx;
    fL.s = s;
    fR.h = netR[0];
    fR.hu = netR[1] * nx - netR[2] * ny;
    fR.hv = netR[1] * ny + netR[2] * nx;
    fR.s = s;
}

// Per-cell quantities of the toleranced solver from a cell's state: depth,
// world velocity and sqrt(g h) (velocity and root 0 on a dry cell,
// h <= drytol).
struct CellQ {
    double h, wu, wv, b, sq;
};

__device__ __forceinline__ CellQ cell_q(double eta, double hu, double hv, double b, double drytol,
                                        double g) {
    CellQ c;
    c.h = eta - b;
    c.b = b;
    const bool wet = c.h > drytol;
    const double r = wet ? rcp(c.h) : 0.0;
    c.wu = hu * r;
    c.wv = hv * r;
    c.sq = wet ? sqrt0(g * c.h) : 0.0;
    return c;
}

}  // namespace tol
}  // namespace swe
