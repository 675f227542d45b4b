// swe_riemann.hpp -- GeoClaw Riemann solvers as FP64 device functions.
//
// Restates c_bind_riemannsolvers.f90:16-176 (wet/dry handling, wall ghost,
// Einfeldt/Roe speeds, net-update split), riemann_aug_JCP
// (riemannsolvers.f:2-258, maxiter = 1), riemann_fwave (:446-500) and
// riemanntype (:507-618).  Every expression keeps the reference's operation
// order; the library is compiled with -ffp-contract=off, so each result
// rounds exactly as the reference (built with the same flag) rounds it.
//
// The 3x3 Cramer system of riemannsolvers.f:188-205 has a fixed sparsity
// (eigenvector 2 is (0,0,1) because the rarefaction corrector is disabled,
// :64-65 and :97-104).  Expanding the determinants with that structure
// leaves, for each beta, the same sequence of non-trivial roundings the
// reference performs; the terms dropped are exact zeros (0*x), so the results
// are identical up to the sign of a zero.
#pragma once

#include <hip/hip_runtime.h>

// Every solver function is usable on the host too, so that the adjoint can
// be checked on the CPU; host sqrt is correctly rounded like the device's.
#define SWE_HD __host__ __device__ __forceinline__

namespace swe {

struct Flux {
    double h, hu, hv, s;  // net update (h, hu, hv) and max wave speed
};

SWE_HD double dmin2(double a, double b) { return a < b ? a : b; }
SWE_HD double dmax2(double a, double b) { return a > b ? a : b; }

// Smallest dry tolerance the solver accepts: above it every depth the solver
// divides by is a normal number whose square and reciprocal are finite, which
// the exact-zero arguments of wall_hstar and riemanntype rely on.
constexpr double MIN_DRY_TOLERANCE = 1.0e-100;
// Range of the gravitational constant the library accepts: with it and the
// dry tolerance above, g h of a wet side is at least 1e-110.
constexpr double MIN_G = 1.0e-10, MAX_G = 1.0e10;

// sqrt(x) for x == 0, x == +inf or x >= 2^-767: the compiler's correctly
// rounded FP64 square root without its range scaling.  The compiler scales
// an argument below 2^-767 by 2^256 before its refinement (and the root back
// by 2^-128), and returns zeros and infinities as they are; for the
// arguments allowed here the scaling is by 2^0, so the refinement below --
// the same operations in the same order -- gives the same bits.  Used for
// sqrt(g h) of a side whose depth is 0 or above the dry tolerance.
SWE_HD double sqrt_wet(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double g0 = __builtin_amdgcn_rsq(x);
    const double s0 = x * g0;
    const double h0 = g0 * 0.5;
    const double r0 = __builtin_fma(-h0, s0, 0.5);
    const double s1 = __builtin_fma(s0, r0, s0);
    const double h1 = __builtin_fma(h0, r0, h0);
    const double d0 = __builtin_fma(-s1, s1, x);
    const double s2 = __builtin_fma(d0, h1, s1);
    const double d1 = __builtin_fma(-s2, s2, x);
    const double s3 = __builtin_fma(d1, h1, s2);
    return __builtin_amdgcn_class(x, 0x260) ? x : s3;  // -0, +0, +inf
#else
    return sqrt(x);
#endif
}

// riemannsolvers.f:507-618 (maxiter = 1); returns hm, s1m, s2m.
// sqL = sqrt(g*hL) and sqR = sqrt(g*hR) are passed in: the reference
// evaluates the same expressions here, so reusing the caller's values is
// exact.  inv16g = 1/(16 g) (:556), computed once on the host.
SWE_HD void riemanntype(double hL, double hR, double uL, double uR,
                                            double sqL, double sqR, double drytol, double g,
                                            double inv16g, double &hm, double &s1m,
                                            double &s2m) {
    const bool r_min = hR < hL;  // dmin2(hR, hL) picks hR
    const double h_min = r_min ? hR : hL;
    const double h_max = hR > hL ? hR : hL;
    const double sqh_min = r_min ? sqR : sqL;
    const double sqh_max = hR > hL ? sqR : sqL;
    const double delu = uR - uL;
    if (h_min <= drytol) {
        hm = 0.0;
        s1m = ((uR + uL) - 2.0 * sqR) + 2.0 * sqL;
        s2m = s1m;
        return;
    }
    const double F_min = delu + 2.0 * (sqh_min - sqh_max);
    const double F_max =
        delu + (h_max - h_min) * sqrt(((0.5 * g) * (h_max + h_min)) / (h_max * h_min));
    if (F_min > 0.0) {  // two rarefactions
        const double t = dmax2(0.0, -delu + 2.0 * (sqL + sqR));
        hm = inv16g * (t * t);
        const double sqm = sqrt(g * hm);
        s1m = (uL + 2.0 * sqL) - 3.0 * sqm;
        s2m = (uR - 2.0 * sqR) + 3.0 * sqm;
    } else if (F_max <= 0.0) {  // two shocks: one Newton step on sqrt(h)
        const double h0 = h_max;
        const double gL = sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hL));
        const double gR = sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hR));
        const double F0 = (delu + (h0 - hL) * gL) + (h0 - hR) * gR;
        const double h02 = h0 * h0;
        const double dfdh = ((gL - (g * (h0 - hL)) / ((4.0 * h02) * gL)) + gR) -
                            (g * (h0 - hR)) / ((4.0 * h02) * gR);
        const double sq0 = sqrt(h0);
        const double slope = (2.0 * sq0) * dfdh;
        const double q = sq0 - F0 / slope;
        hm = q * q;
        const double u1m = uL - (hm - hL) * sqrt((0.5 * g) * (1.0 / hm + 1.0 / hL));
        const double u2m = uR + (hm - hR) * sqrt((0.5 * g) * (1.0 / hm + 1.0 / hR));
        const double sqm = sqrt(g * hm);
        s1m = u1m - sqm;
        s2m = u2m + sqm;
    } else {  // one shock, one rarefaction: one secant step
        // The reference evaluates F0 at h0 = h_min as
        //   (delu + 2 (sqrt(g h_min) - sqrt(g h_max))) + (h0 - h_min) sqrt(g/2 (1/h0 + 1/h_min)),
        // whose second term is (+0) * (a finite value, since h_min > drytol >=
        // MIN_DRY_TOLERANCE) = +0; adding +0 is exact, so F0 = F_min + 0.
        const double h0 = h_min;
        const double F0 = F_min + 0.0;
        const double slope = (F_max - F0) / (h_max - h_min);
        hm = h0 - F0 / slope;
        const double sqm = sqrt(g * hm);
        if (hL > hR) {
            s1m = (uL + 2.0 * sqL) - 3.0 * sqm;
            s2m = (uL + 2.0 * sqL) - sqm;
        } else {
            s2m = (uR - 2.0 * sqR) + 3.0 * sqm;
            s1m = (uR - 2.0 * sqR) + sqm;
        }
    }
}

// hm of riemanntype(h, h, a, c) for the wall test of a wet cell of depth h
// against a dry neighbour (c_bind_riemannsolvers.f90:88,102).  With equal
// depths both sides' sqrt(g h) are equal, so the reference's terms
// 2 (sq - sq), (h - h) * sqrt(...) and g (h0 - h) / (...) are exact zeros
// (their other factors are finite because h > drytol >= MIN_DRY_TOLERANCE):
// F_min = F_max = delu + 0, the shock-rarefaction branch is unreachable, and
// in the two-shock branch gL = gR, F0 = delu + 0, dfdh = gL + gL.  s1m and
// s2m are not needed by the wall test.
SWE_HD double wall_hstar(double h, double a, double c, double sq, double g, double inv16g) {
    const double delu = c - a;
    const double F = delu + 0.0;
    if (F > 0.0) {  // two rarefactions
        const double t = dmax2(0.0, -delu + 2.0 * (sq + sq));
        return inv16g * (t * t);
    }
    const double inv = 1.0 / h;  // two shocks, h0 = hL = hR = h
    const double gL = sqrt((0.5 * g) * (inv + inv));
    const double dfdh = gL + gL;
    const double sq0 = sqrt(h);
    const double slope = (2.0 * sq0) * dfdh;
    const double q = sq0 - F / slope;
    return q * q;
}

// c_bind_geoclaw_solver_dp (c_bind_riemannsolvers.f90:16-176) in the edge
// frame: h = depth, hu = normal momentum, hv = tangential momentum.
// net[0..2] left-going (A^- dQ), net[3..5] right-going (A^+ dQ).
template <int SOLVER>
SWE_HD void geoclaw(double hL, double hR, double huL, double huR,
                                        double hvL, double hvR, double bL, double bR,
                                        double drytol, double g, double inv16g, double netL[3],
                                        double netR[3], double &maxs) {
    netL[0] = netL[1] = netL[2] = 0.0;
    netR[0] = netR[1] = netR[2] = 0.0;
    maxs = 0.0;
    if (hL <= drytol && hR <= drytol) return;  // :55-57

    double uL, uR, vL, vR;
    if (hR > drytol) {  // :60-69
        uR = huR / hR;
        vR = hvR / hR;
    } else {
        hR = 0.0;
        huR = 0.0;
        hvR = 0.0;
        uR = 0.0;
        vR = 0.0;
    }
    if (hL > drytol) {  // :71-80
        uL = huL / hL;
        vL = hvL / hL;
    } else {
        hL = 0.0;
        huL = 0.0;
        hvL = 0.0;
        uL = 0.0;
        vL = 0.0;
    }

    double wall0 = 1.0, wall1 = 1.0, wall2 = 1.0;
    if (hR <= drytol || hL <= drytol) {
        // :87-114, exactly one side dry (dry/dry returned above): wall test
        // riemanntype(hL, hL, uL, -uL) if the right cell is dry, else
        // riemanntype(hR, hR, -uR, uR); both branches share one code path.
        const bool rdry = hR <= drytol;
        const double hw = rdry ? hL : hR;
        const double bw = rdry ? bL : bR, bd = rdry ? bR : bL;
        const double sq = sqrt(g * hw);
        const double hstar = wall_hstar(hw, rdry ? uL : -uR, rdry ? -uL : uR, sq, g, inv16g);
        const double hstartest = dmax2(hw, hstar);
        if (hstartest + bw < bd) {
            wall1 = 0.0;
            if (rdry) {
                wall2 = 0.0;
                hR = hL;
                huR = -huL;
                bR = bL;
                uR = -uL;
                vR = vL;
            } else {
                wall0 = 0.0;
                hL = hR;
                huL = -huR;
                bL = bR;
                uL = -uR;
                vL = vR;
            }
        } else if (hw + bw < bd) {
            if (rdry)
                bR = hL + bL;
            else
                bL = hR + bR;
        }
    }

    // :120
    const double delphi = (((huR - huL) * (uL + uR) - (uL * uR) * (hR - hL)) +
                           ((0.5 * g) * (((bR + hR) - bL) - hL)) * (hR + hL)) -
                          ((0.5 * g) * (hR + hL)) * (bR - bL);
    // :123-132 (a side's depth is 0 or above the dry tolerance here)
    const double sqL = sqrt_wet(g * hL), sqR = sqrt_wet(g * hR);
    const double sL = uL - sqL;
    const double sR = uR + sqR;
    const double uhat = (sqL * uL + sqR * uR) / (sqR + sqL);
    const double chat = sqrt_wet((g * 0.5) * (hR + hL));
    double sE1 = dmin2(sL, uhat - chat);
    double sE2 = dmax2(sR, uhat + chat);

    double sw0, sw1, sw2;
    double f00, f10, f20, f01, f11, f21, f02, f12, f22;  // f<eq><wave>
    const double delhu = huR - huL;
    const double delb = bR - bL;
    if (SOLVER == 1) {  // riemann_fwave, riemannsolvers.f:446-500
        const double deldelphi = -(((g * 0.5) * (hR + hL)) * delb);
        const double dec = delphi - deldelphi;
        const double beta1 = (sE2 * delhu - dec) / (sE2 - sE1);
        const double beta2 = (dec - sE1 * delhu) / (sE2 - sE1);
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
        const double hstarHLL = dmax2((((huL - huR) + sE2 * hR) - sE1 * hL) / (sE2 - sE1), 0.0);
        double l1 = 0.5 * (l0 + l2);  // :98-100
        if (fabs(l1) < 1.0e-20) l1 = 0.0;
        const double r20 = l0 * l0, r22 = l2 * l2;  // r(3,1), r(3,3)

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

        double deldelh = sonic ? -delb : ((delb * g) * hbar) / s1s2bar;
        if (sE1 < -criticaltol && sE2 > criticaltol) {
            deldelh = dmin2(deldelh, (hstarHLL * (sE2 - sE1)) / sE2);
            deldelh = dmax2(deldelh, (hstarHLL * (sE2 - sE1)) / sE1);
        } else if (sE1 >= criticaltol) {
            deldelh = dmin2(deldelh, (hstarHLL * (sE2 - sE1)) / sE1);
            deldelh = dmax2(deldelh, -hL);
        } else if (sE2 <= -criticaltol) {
            deldelh = dmin2(deldelh, hR);
            deldelh = dmax2(deldelh, (hstarHLL * (sE2 - sE1)) / sE2);
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
        const double det = -l2 + l0;
        const double beta0 = (-(d0 * l2) + d1) / det;
        const double beta1 =
            ((d1 * r22 - l2 * d2) - d0 * (l0 * r22 - l2 * r20) + (l0 * d2 - d1 * r20)) / det;
        const double beta2 = (-d1 + d0 * l0) / det;

        sw0 = l0;
        sw1 = l1;
        sw2 = l2;
        f00 = beta0 * l0;
        f10 = beta0 * r20;
        f20 = (beta0 * l0) * vL;
        f01 = beta1 * 0.0;
        f11 = beta1;
        f02 = beta2 * l2;
        f12 = beta2 * r22;
        f22 = (beta2 * l2) * vR;
        f21 = (((hR * uR) * vR - (hL * uL) * vL) - f20) - f22;  // :254
    }

    // :148-153 wall masking (a factor 1.0 leaves every value as it is)
    if (wall0 != 1.0 || wall1 != 1.0 || wall2 != 1.0) {
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
#define SWE_SPLIT(sw, fa, fb, fc)          \
    if (sw < 0.0) {                        \
        netL[0] += fa;                     \
        netL[1] += fb;                     \
        netL[2] += fc;                     \
    } else if (sw > 0.0) {                 \
        netR[0] += fa;                     \
        netR[1] += fb;                     \
        netR[2] += fc;                     \
    } else {                               \
        netL[0] += 0.5 * fa;               \
        netL[1] += 0.5 * fb;               \
        netL[2] += 0.5 * fc;               \
        netR[0] += 0.5 * fa;               \
        netR[1] += 0.5 * fb;               \
        netR[2] += 0.5 * fc;               \
    }
    SWE_SPLIT(sw0, f00, f10, f20)
    SWE_SPLIT(sw1, f01, f11, f21)
    SWE_SPLIT(sw2, f02, f12, f22)
#undef SWE_SPLIT
    // :174-175
    maxs = dmax2(dmax2(fabs(sw0), fabs(sw1)), fabs(sw2));
}

// Solver ids beyond GeoClaw's: compute_lf_flux (SWE_euler_timestep.f90:359-446)
// under _LF_FLUX, _LLF_FLUX, _LF_BATH_FLUX and _LLF_BATH_FLUX.
constexpr int SOLVER_LF = 2, SOLVER_LLF = 3, SOLVER_LF_BATH = 4, SOLVER_LLF_BATH = 5;

// compute_lf_flux in world coordinates: Q = (eta, hu, hv, b), n the left
// cell's outward normal; each side gets its own wave speed.  Fortran's vector
// expressions are written out per component in the source's order (left to
// right; a unary minus binds looser than a product, which changes no value).
template <int SOLVER>
SWE_HD void lf_flux(double nx, double ny, double etaL, double huL, double hvL, double bL,
                    double etaR, double huR, double hvR, double bR, double drytol, double g,
                    Flux &fL, Flux &fR) {
    constexpr bool llf = SOLVER == SOLVER_LLF || SOLVER == SOLVER_LLF_BATH;
    if (SOLVER == SOLVER_LF_BATH || SOLVER == SOLVER_LLF_BATH) {
        const bool dryL = etaL - bL < drytol, dryR = etaR - bR < drytol;
        if (dryL || dryR) {  // :367-392, a wall against the dry side
            fL = Flux{0.0, 0.0, 0.0, 0.0};
            fR = Flux{0.0, 0.0, 0.0, 0.0};
            if (dryL && dryR) return;
            if (dryL) {
                const double hR = dmax2(etaR - bR, 0.0);
                const double vR = nx * (huR / (etaR - bR)) + ny * (hvR / (etaR - bR));
                fR.s = sqrt(g * hR) + fabs(vR);
                const double pr = ((0.5 * g) * hR) * hR;
                fR.hu = (-((0.5 * vR) * huR) - pr * nx) + (0.5 * fR.s) * huR;
                fR.hv = (-((0.5 * vR) * hvR) - pr * ny) + (0.5 * fR.s) * hvR;
            } else {
                const double hL = dmax2(etaL - bL, 0.0);
                const double vL = nx * (huL / (etaL - bL)) + ny * (hvL / (etaL - bL));
                fL.s = sqrt(g * hL) + fabs(vL);
                const double pr = ((0.5 * g) * hL) * hL;
                fL.hu = ((0.5 * vL) * huL + pr * nx) + (0.5 * fL.s) * huL;
                fL.hv = ((0.5 * vL) * hvL + pr * ny) + (0.5 * fL.s) * hvL;
            }
            return;
        }
        const double hL = dmax2(etaL - bL, 0.0), hR = dmax2(etaR - bR, 0.0);  // :394-418
        const double vL = nx * (huL / (etaL - bL)) + ny * (hvL / (etaL - bL));
        const double vR = nx * (huR / (etaR - bR)) + ny * (hvR / (etaR - bR));
        fL.s = sqrt(g * hL) + fabs(vL);
        fR.s = sqrt(g * hR) + fabs(vR);
        const double alpha = llf ? dmax2(fL.s, fR.s) : 100.0;
        fL.h = 0.5 * ((hL * vL + hR * vR) + alpha * (etaL - etaR));
        fR.h = -fL.h;
        const double pr = (0.5 * g) * (hL * hL + hR * hR);
        const double u = 0.5 * (((huL * vL + huR * vR) + pr * nx) + alpha * (huL - huR));
        const double v = 0.5 * (((hvL * vL + hvR * vR) + pr * ny) + alpha * (hvL - hvR));
        const double src = ((0.25 * g) * (hL + hR)) * (bR - bL);  // well-balancing, :414-418
        fL.hu = u + src * nx;
        fL.hv = v + src * ny;
        fR.hu = -u + src * nx;
        fR.hv = -v + src * ny;
    } else {  // :422-445, the constant bed b = 0
        const double hL = dmax2(etaL - 0.0, 0.0), hR = dmax2(etaR - 0.0, 0.0);
        const double vL = nx * (huL / (etaL - 0.0)) + ny * (hvL / (etaL - 0.0));
        const double vR = nx * (huR / (etaR - 0.0)) + ny * (hvR / (etaR - 0.0));
        fL.s = sqrt(g * (etaL - 0.0)) + fabs(vL);
        fR.s = sqrt(g * (etaR - 0.0)) + fabs(vR);
        const double alpha = llf ? dmax2(fL.s, fR.s) : 100.0;
        fL.h = 0.5 * ((vL * hL + vR * hR) + alpha * (etaL - etaR));
        fR.h = -fL.h;
        // vL vL hL + vR vR hR is a scalar, added to both components
        const double k = (vL * vL) * hL + (vR * vR) * hR;
        const double pr = (0.5 * g) * (hL * hL + hR * hR);
        fL.hu = 0.5 * ((k + pr * nx) + alpha * (huL - huR));
        fL.hv = 0.5 * ((k + pr * ny) + alpha * (hvL - hvR));
        fR.hu = -fL.hu;
        fR.hv = -fL.hv;
    }
}

// compute_geoclaw_flux, SWE_euler_timestep.f90:449-481.
// QL/QR = (eta, hu, hv, b) in world coordinates, n the left cell's outward
// normal.
template <int SOLVER>
SWE_HD void edge_flux(double nx, double ny, double etaL, double huL,
                                          double hvL, double bL, double etaR, double huR,
                                          double hvR, double bR, double drytol, double g,
                                          double inv16g, Flux &fL, Flux &fR) {
    if constexpr (SOLVER >= SOLVER_LF) {
        lf_flux<SOLVER>(nx, ny, etaL, huL, hvL, bL, etaR, huR, hvR, bR, drytol, g, fL, fR);
        return;
    }
    const double pL1 = nx * huL + ny * hvL;
    const double pL2 = (-ny) * huL + nx * hvL;
    const double pR1 = nx * huR + ny * hvR;
    const double pR2 = (-ny) * huR + nx * hvR;
    double netL[3], netR[3], s;
    geoclaw<SOLVER>(etaL - bL, etaR - bR, pL1, pR1, pL2, pR2, bL, bR, drytol, g, inv16g, netL,
                    netR, s);
    fL.h = netL[0];
    fL.hu = netL[1] * nx + netL[2] * (-ny);
    fL.hv = netL[1] * ny + netL[2] * nx;
    fL.s = s;
    fR.h = netR[0];
    fR.hu = netR[1] * nx + netR[2] * (-ny);
    fR.hv = netR[1] * ny + netR[2] * nx;
    fR.s = s;
}

}  // namespace swe
