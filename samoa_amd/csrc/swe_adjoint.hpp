// swe_adjoint.hpp -- discrete adjoint (reverse mode) of the edge flux.
//
// edge_flux_adj<SOLVER>() returns the transpose Jacobian-vector product of
// edge_flux<SOLVER>() (swe_riemann.hpp): given seeds on the left and right net
// updates (h, hu, hv), it accumulates the gradients with respect to the left
// and right states (eta, hu, hv, b).  It re-evaluates the forward solve with
// the same branch decisions (wet/dry zeroing, wall ghosting, Riemann
// structure, sonic and bound selections, net-update split), then runs the
// statements backwards.  Conventions at non-smooth points: min/max and
// selections pass the gradient to the operand that was selected; comparisons
// carry none; sqrt has derivative 0 where its value is exactly 0 (a dry
// side).  Wave speeds only steer the split and the CFL bound (dt is frozen),
// so they carry no gradient.
#pragma once

#include "swe_riemann.hpp"

namespace swe {

// Reciprocal for the reverse sweep, whose products need about 1e-15
// relative accuracy rather than correct rounding: on the device the hardware
// estimate refined by two Newton-Raphson steps (r <- r + r (1 - x r)), five
// instructions instead of a correctly rounded division's eleven.  Values that
// steer a branch keep the reference's exact division.
SWE_HD double rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}

// d sqrt(x)/dx with value y = sqrt(x), 0 at y == 0.
SWE_HD double dsqrt(double y) { return y > 0.0 ? 0.5 * rcp(y) : 0.0; }

// Reverse of riemanntype (main call of riemann_aug_JCP): seeds gs1m, gs2m;
// accumulates into ghL, ghR, guL, guR, gsqL, gsqR.
SWE_HD void riemanntype_adj(double hL, double hR, double uL, double uR, double sqL, double sqR,
                            double drytol, double g, double inv16g, double gs1m, double gs2m,
                            double &ghL, double &ghR, double &guL, double &guR, double &gsqL,
                            double &gsqR) {
    const bool r_min = hR < hL;
    const bool r_max = hR > hL;
    const double h_min = r_min ? hR : hL;
    const double h_max = r_max ? hR : hL;
    const double sqh_min = r_min ? sqR : sqL;
    const double sqh_max = r_max ? sqR : sqL;
    const double delu = uR - uL;
    double gdelu = 0.0, ghmin = 0.0, ghmax = 0.0, gsqmin = 0.0, gsqmax = 0.0;
    if (h_min <= drytol) {
        const double gs = gs1m + gs2m;
        guR += gs;
        guL += gs;
        gsqR -= 2.0 * gs;
        gsqL += 2.0 * gs;
        return;
    }
    const double F_min = delu + 2.0 * (sqh_min - sqh_max);
    const double Wn = (0.5 * g) * (h_max + h_min);
    const double Wd = h_max * h_min;
    const double Y = Wn / Wd;
    const double sqY = sqrt(Y);
    const double F_max = delu + (h_max - h_min) * sqY;
    if (F_min > 0.0) {  // two rarefactions
        const double tt = -delu + 2.0 * (sqL + sqR);
        const double t = tt > 0.0 ? tt : 0.0;
        const double hm = inv16g * (t * t);
        const double sqm = sqrt(g * hm);
        // s1m = (uL + 2 sqL) - 3 sqm; s2m = (uR - 2 sqR) + 3 sqm
        guL += gs1m;
        gsqL += 2.0 * gs1m;
        guR += gs2m;
        gsqR -= 2.0 * gs2m;
        const double gsqm = -3.0 * gs1m + 3.0 * gs2m;
        const double ghm = gsqm * g * dsqrt(sqm);
        const double gt = ghm * inv16g * 2.0 * t;
        if (tt > 0.0) {
            gdelu -= gt;
            gsqL += 2.0 * gt;
            gsqR += 2.0 * gt;
        }
    } else if (F_max <= 0.0) {  // two shocks
        const double h0 = h_max;
        const double ih0 = rcp(h0), ihL = rcp(hL), ihR = rcp(hR);
        const double iL = ih0 + ihL, iR = ih0 + ihR;
        const double gL = sqrt((0.5 * g) * iL), gR = sqrt((0.5 * g) * iR);
        const double F0 = (delu + (h0 - hL) * gL) + (h0 - hR) * gR;
        const double h02 = h0 * h0;
        const double e1 = g * (h0 - hL), e2 = (4.0 * h02) * gL, ie2 = rcp(e2), t1 = e1 * ie2;
        const double e3 = g * (h0 - hR), e4 = (4.0 * h02) * gR, ie4 = rcp(e4), t2 = e3 * ie4;
        const double dfdh = ((gL - t1) + gR) - t2;
        const double sq0 = sqrt(h0);
        const double slope = (2.0 * sq0) * dfdh;
        const double islope = rcp(slope);
        const double fs = F0 * islope;
        const double q = sq0 - fs;
        const double hm = q * q;
        const double ihm = rcp(hm);
        const double sL_ = sqrt((0.5 * g) * (ihm + ihL));
        const double sR_ = sqrt((0.5 * g) * (ihm + ihR));
        const double sqm = sqrt(g * hm);
        // s1m = u1m - sqm, s2m = u2m + sqm
        const double gu1m = gs1m, gu2m = gs2m;
        const double gsqm = -gs1m + gs2m;
        double ghm = gsqm * g * dsqrt(sqm);
        double gh0 = 0.0, ggL = 0.0, ggR = 0.0;
        // u2m = uR + (hm - hR) sR_
        guR += gu2m;
        ghm += gu2m * sR_;
        ghR -= gu2m * sR_;
        {
            const double gjR = gu2m * (hm - hR) * dsqrt(sR_) * (0.5 * g);
            ghm -= gjR * (ihm * ihm);
            ghR -= gjR * (ihR * ihR);
        }
        // u1m = uL - (hm - hL) sL_
        guL += gu1m;
        ghm -= gu1m * sL_;
        ghL += gu1m * sL_;
        {
            const double gjL = -gu1m * (hm - hL) * dsqrt(sL_) * (0.5 * g);
            ghm -= gjL * (ihm * ihm);
            ghL -= gjL * (ihL * ihL);
        }
        const double gq = 2.0 * q * ghm;
        double gsq0 = gq;
        const double gfs = -gq;
        double gF0 = gfs * islope;
        const double gslope = -gfs * fs * islope;
        gsq0 += gslope * 2.0 * dfdh;
        const double gdfdh = gslope * (2.0 * sq0);
        ggL += gdfdh;
        ggR += gdfdh;
        const double gt1 = -gdfdh, gt2 = -gdfdh;
        double gh02 = 0.0;
        {
            const double ge3 = gt2 * ie4, ge4 = -gt2 * t2 * ie4;
            gh0 += ge3 * g;
            ghR -= ge3 * g;
            gh02 += ge4 * 4.0 * gR;
            ggR += ge4 * (4.0 * h02);
            const double ge1 = gt1 * ie2, ge2 = -gt1 * t1 * ie2;
            gh0 += ge1 * g;
            ghL -= ge1 * g;
            gh02 += ge2 * 4.0 * gL;
            ggL += ge2 * (4.0 * h02);
        }
        gh0 += 2.0 * h0 * gh02;
        gh0 += gsq0 * dsqrt(sq0);
        // F0 = (delu + (h0 - hL) gL) + (h0 - hR) gR
        gdelu += gF0;
        gh0 += gF0 * gL + gF0 * gR;
        ghL -= gF0 * gL;
        ghR -= gF0 * gR;
        ggL += gF0 * (h0 - hL);
        ggR += gF0 * (h0 - hR);
        {
            const double giR = ggR * dsqrt(gR) * (0.5 * g);
            gh0 -= giR * (ih0 * ih0);
            ghR -= giR * (ihR * ihR);
            const double giL = ggL * dsqrt(gL) * (0.5 * g);
            gh0 -= giL * (ih0 * ih0);
            ghL -= giL * (ihL * ihL);
        }
        ghmax += gh0;
        (void)gF0;
    } else {  // one shock, one rarefaction
        const double h0 = h_min;
        const double F0 = F_min + 0.0;  // the reference's second term is an exact zero
        const double den = h_max - h_min, iden = rcp(den);
        const double slope = (F_max - F0) * iden;
        const double islope = rcp(slope);
        const double fs = F0 * islope;
        const double hm = h0 - fs;
        const double sqm = sqrt(g * hm);
        double gsqm;
        if (hL > hR) {
            const double gs = gs1m + gs2m;
            guL += gs;
            gsqL += 2.0 * gs;
            gsqm = -3.0 * gs1m - gs2m;
        } else {
            const double gs = gs1m + gs2m;
            guR += gs;
            gsqR -= 2.0 * gs;
            gsqm = 3.0 * gs2m + gs1m;
        }
        const double ghm = gsqm * g * dsqrt(sqm);
        double gh0 = ghm;
        const double gfs = -ghm;
        double gF0 = gfs * islope;
        const double gslope = -gfs * fs * islope;
        const double gnum = gslope * iden, gden = -gslope * slope * iden;
        const double gFmax = gnum;
        gF0 -= gnum;
        ghmax += gden;
        ghmin -= gden;
        // F_max = delu + (h_max - h_min) sqrt(Wn / Wd)
        gdelu += gFmax;
        ghmax += gFmax * sqY;
        ghmin -= gFmax * sqY;
        const double gY = gFmax * (h_max - h_min) * dsqrt(sqY);
        const double iWd = rcp(Wd);
        const double gWn = gY * iWd, gWd = -gY * Y * iWd;
        ghmax += gWn * (0.5 * g) + gWd * h_min;
        ghmin += gWn * (0.5 * g) + gWd * h_max;
        // F0 = delu + 2 (sqmin - sqmax) (+ an exact zero)
        gdelu += gF0;
        gsqmin += 2.0 * gF0;
        gsqmax -= 2.0 * gF0;
        ghmin += gh0;
    }
    guR += gdelu;
    guL -= gdelu;
    // selects, not branches: accumulating into a variable chosen by a
    // condition makes the compiler index a private array in scratch
    ghR += (r_min ? ghmin : 0.0) + (r_max ? ghmax : 0.0);
    gsqR += (r_min ? gsqmin : 0.0) + (r_max ? gsqmax : 0.0);
    ghL += (r_min ? 0.0 : ghmin) + (r_max ? 0.0 : ghmax);
    gsqL += (r_min ? 0.0 : gsqmin) + (r_max ? 0.0 : gsqmax);
}

// Reverse of geoclaw<SOLVER> (c_bind_riemannsolvers.f90:16-176).  Inputs as
// the forward; seeds gnL/gnR on the net updates; out: gradients with respect
// to (hL, hR, huL, huR, hvL, hvR, bL, bR), accumulated into gin.
template <int SOLVER>
SWE_HD void geoclaw_adj(double hL, double hR, double huL, double huR, double hvL, double hvR,
                        double bL, double bR, double drytol, double g, double inv16g,
                        const double gnL[3], const double gnR[3], double gin[8]) {
    if (hL <= drytol && hR <= drytol) return;  // zero updates

    // stage A: wet/dry zeroing (:60-80)
    const bool wetR = hR > drytol, wetL = hL > drytol;
    const double hR1 = wetR ? hR : 0.0, huR1 = wetR ? huR : 0.0;
    // u steers the wall test and the Riemann structure: the reference's
    // division; v only scales f-waves: a refined reciprocal, shared with the
    // reverse of stage A
    const double ihR = wetR ? rcp(hR) : 0.0, ihL = wetL ? rcp(hL) : 0.0;
    const double uR1 = wetR ? huR / hR : 0.0, vR1 = wetR ? hvR * ihR : 0.0;
    const double hL1 = wetL ? hL : 0.0, huL1 = wetL ? huL : 0.0;
    const double uL1 = wetL ? huL / hL : 0.0, vL1 = wetL ? hvL * ihL : 0.0;

    // stage B: wall ghosting (:87-115)
    // (the decisions only: hstar carries no gradient; one code path for both
    // sides, with the closed-form hstar of the forward solver)
    bool wallR = false, clipR = false, wallL = false, clipL = false;
    if (hR1 <= drytol || hL1 <= drytol) {
        const bool rdry = hR1 <= drytol;
        const double hw = rdry ? hL1 : hR1;
        const double bw = rdry ? bL : bR, bd = rdry ? bR : bL;
        const double hstar =
            wall_hstar(hw, rdry ? uL1 : -uR1, rdry ? -uL1 : uR1, sqrt(g * hw), g, inv16g);
        const bool wall = dmax2(hw, hstar) + bw < bd;
        const bool clip = !wall && (hw + bw < bd);
        wallR = rdry && wall;
        clipR = rdry && clip;
        wallL = !rdry && wall;
        clipL = !rdry && clip;
    }
    const double hR2 = wallR ? hL1 : hR1, huR2 = wallR ? -huL1 : huR1;
    const double bR2 = wallR ? bL : (clipR ? hL1 + bL : bR);
    const double uR2 = wallR ? -uL1 : uR1, vR2 = wallR ? vL1 : vR1;
    const double hL2 = wallL ? hR1 : hL1, huL2 = wallL ? -huR1 : huL1;
    const double bL2 = wallL ? bR : (clipL ? hR1 + bR : bL);
    const double uL2 = wallL ? -uR1 : uL1, vL2 = wallL ? vR1 : vL1;
    const double wall0 = wallL ? 0.0 : 1.0, wall1 = (wallL || wallR) ? 0.0 : 1.0,
                 wall2 = wallR ? 0.0 : 1.0;

    // stage C: speeds (:120-132)
    const double sqL = sqrt(g * hL2), sqR = sqrt(g * hR2);
    const double sL = uL2 - sqL, sR = uR2 + sqR;
    const double num = sqL * uL2 + sqR * uR2, den = sqR + sqL;
    const double uhat = num / den;
    const double chat = sqrt((g * 0.5) * (hR2 + hL2));
    const double a1 = uhat - chat, a2 = uhat + chat;
    const bool selE1 = sL < a1, selE2 = sR > a2;
    const double sE1 = selE1 ? sL : a1, sE2 = selE2 ? sR : a2;
    const double delhu = huR2 - huL2, delb = bR2 - bL2;
    const double delphi = (((huR2 - huL2) * (uL2 + uR2) - (uL2 * uR2) * (hR2 - hL2)) +
                           ((0.5 * g) * (((bR2 + hR2) - bL2) - hL2)) * (hR2 + hL2)) -
                          ((0.5 * g) * (hR2 + hL2)) * (bR2 - bL2);

    // gradients of the stage-C/D variables
    double ghL2 = 0, ghR2 = 0, ghuL2 = 0, ghuR2 = 0, guL2 = 0, guR2 = 0, gvL2 = 0, gvR2 = 0;
    double gbL2 = 0, gbR2 = 0, gdelphi = 0, gsE1 = 0, gsE2 = 0, gsqL = 0, gsqR = 0;
    double gdelhu = 0, gdelb = 0;

    // seeds on the f-waves from the net-update split (:148-167)
    double sw[3], wall[3] = {wall0, wall1, wall2};
    double gf[3][3];  // gf[eq][wave]

    if (SOLVER == 1) {  // riemann_fwave, riemannsolvers.f:446-500
        sw[0] = sE1 * wall0;
        sw[1] = (0.5 * (sE1 + sE2)) * wall1;
        sw[2] = sE2 * wall2;
        for (int mw = 0; mw < 3; ++mw)
            for (int m = 0; m < 3; ++m)
                gf[m][mw] = (sw[mw] < 0.0 ? gnL[m] : (sw[mw] > 0.0 ? gnR[m] : 0.5 * (gnL[m] + gnR[m]))) * wall[mw];
        const double K = (g * 0.5) * (hR2 + hL2);
        const double deldelphi = -(K * delb);
        const double dec = delphi - deldelphi;
        const double dd = sE2 - sE1;
        const double idd = rcp(dd);  // the wave strengths steer nothing
        const double beta1 = (sE2 * delhu - dec) * idd;
        const double beta2 = (dec - sE1 * delhu) * idd;
        const double f20 = beta1 * vL2, f22 = beta2 * vR2;
        // f21 = ((huR vR - huL vL) - f20) - f22
        const double gA = gf[2][1];
        ghuR2 += gA * vR2;
        gvR2 += gA * huR2;
        ghuL2 -= gA * vL2;
        gvL2 -= gA * huL2;
        const double gf20 = gf[2][0] - gA, gf22 = gf[2][2] - gA;
        const double gb2 = gf[0][2] + gf[1][2] * sE2 + gf22 * vR2;
        gsE2 += gf[1][2] * beta2;
        gvR2 += gf22 * beta2;
        const double gb1 = gf[0][0] + gf[1][0] * sE1 + gf20 * vL2;
        gsE1 += gf[1][0] * beta1;
        gvL2 += gf20 * beta1;
        (void)f20;
        (void)f22;
        const double gn2 = gb2 * idd, gn1 = gb1 * idd;
        const double gdd = -(gb1 * beta1 + gb2 * beta2) * idd;
        double gdec = gn2 - gn1;
        gsE1 -= gn2 * delhu;
        gdelhu -= gn2 * sE1;
        gsE2 += gn1 * delhu;
        gdelhu += gn1 * sE2;
        gsE2 += gdd;
        gsE1 -= gdd;
        gdelphi += gdec;
        const double gddp = -gdec;
        const double gK = -gddp * delb;
        gdelb -= gddp * K;
        ghR2 += gK * (g * 0.5);
        ghL2 += gK * (g * 0.5);
    } else {  // riemann_aug_JCP, riemannsolvers.f:2-258
        const double delh = hR2 - hL2;
        double hm, s1m, s2m;
        riemanntype(hL2, hR2, uL2, uR2, sqL, sqR, drytol, g, inv16g, hm, s1m, s2m);
        const bool sel0 = sE1 < s2m, sel2 = sE2 > s1m;
        const double l0 = sel0 ? sE1 : s2m, l2 = sel2 ? sE2 : s1m;
        const double dl = l2 - l0;
        const double xh = ((huL2 - huR2) + l2 * hR2) - l0 * hL2;
        const double xq = xh / dl;
        const bool hpos = xq > 0.0;
        const double hstarHLL = hpos ? xq : 0.0;
        double l1 = 0.5 * (l0 + l2);
        if (fabs(l1) < 1.0e-20) l1 = 0.0;
        const double r20 = l0 * l0, r22 = l2 * l2;
        const double criticaltol = 1.0e-6;
        const double hs = 0.5 * (hL2 + hR2);
        const bool hbpos = hs > 0.0;
        const double hbar = hbpos ? hs : 0.0;
        const double us = uL2 + uR2;
        const double s1s2bar = 0.25 * (us * us) - g * hbar;
        const double uu = uL2 * uR2;
        const bool uneg = 0.0 > uu;  // dmax2(0, uu) picks 0
        const double s1s2tilde = (uneg ? 0.0 : uu) - g * hbar;
        bool sonic = false;
        if (fabs(s1s2bar) <= criticaltol) sonic = true;
        if (s1s2bar * s1s2tilde <= criticaltol) sonic = true;
        if ((s1s2bar * l0) * l2 <= criticaltol) sonic = true;
        if (dmin2(fabs(l0), fabs(l2)) < criticaltol) sonic = true;
        if (l0 < 0.0 && s1m > 0.0) sonic = true;
        if (l2 > 0.0 && s2m < 0.0) sonic = true;
        if ((uL2 + sqL) * (uR2 + sqR) < 0.0) sonic = true;
        if ((uL2 - sqL) * (uR2 - sqR) < 0.0) sonic = true;

        // deldelh with its bounds; remember which operand each min/max took
        const double M = (delb * g) * hbar;
        const double dd0 = sonic ? -delb : M / s1s2bar;
        const double B2 = (hstarHLL * dl) / l2, B1 = (hstarHLL * dl) / l0;
        int ddcase = 0;  // 0: none, 1: (sE1<-tol, sE2>tol), 2: sE1>=tol, 3: sE2<=-tol
        bool sa = true, sb = true;
        double deldelh = dd0;
        if (l0 < -criticaltol && l2 > criticaltol) {
            ddcase = 1;
            sa = dd0 < B2;
            const double t1 = sa ? dd0 : B2;
            sb = t1 > B1;
            deldelh = sb ? t1 : B1;
        } else if (l0 >= criticaltol) {
            ddcase = 2;
            sa = dd0 < B1;
            const double t1 = sa ? dd0 : B1;
            sb = t1 > -hL2;
            deldelh = sb ? t1 : -hL2;
        } else if (l2 <= -criticaltol) {
            ddcase = 3;
            sa = dd0 < hR2;
            const double t1 = sa ? dd0 : hR2;
            sb = t1 > B2;
            deldelh = sb ? t1 : B2;
        }
        const double Kp = M * s1s2tilde;
        const double p0 = sonic ? -((g * hbar) * delb) : -(Kp / s1s2bar);
        const double pL = -(hL2 * delb), pR = -(hR2 * delb);
        const bool hiL = pL > pR, loL = pL < pR;
        const double hi = g * (hiL ? pL : pR), lo = g * (loL ? pL : pR);
        const bool sc_ = p0 < hi;
        const double t2 = sc_ ? p0 : hi;
        const bool sd_ = t2 > lo;
        const double deldelphi = sd_ ? t2 : lo;
        const double d0 = delh - deldelh, d1 = delhu, d2 = delphi - deldelphi;
        const double det = -l2 + l0;
        const double idet = rcp(det);  // the wave strengths steer nothing
        const double beta0 = (-(d0 * l2) + d1) * idet;
        const double Aq = d1 * r22 - l2 * d2, Bq = l0 * r22 - l2 * r20, Cq = l0 * d2 - d1 * r20;
        const double beta1 = ((Aq - d0 * Bq) + Cq) * idet;
        const double beta2 = (-d1 + d0 * l0) * idet;
        const double f00 = beta0 * l0, f02 = beta2 * l2;

        sw[0] = l0 * wall0;
        sw[1] = l1 * wall1;
        sw[2] = l2 * wall2;
        for (int mw = 0; mw < 3; ++mw)
            for (int m = 0; m < 3; ++m)
                gf[m][mw] = (sw[mw] < 0.0 ? gnL[m] : (sw[mw] > 0.0 ? gnR[m] : 0.5 * (gnL[m] + gnR[m]))) * wall[mw];

        // f21 = ((hR uR vR - hL uL vL) - f20) - f22
        const double gA = gf[2][1];
        {
            const double huR_ = hR2 * uR2, huL_ = hL2 * uL2;
            const double ghuR_ = gA * vR2, ghuL_ = -gA * vL2;
            gvR2 += gA * huR_;
            gvL2 -= gA * huL_;
            ghR2 += ghuR_ * uR2;
            guR2 += ghuR_ * hR2;
            ghL2 += ghuL_ * uL2;
            guL2 += ghuL_ * hL2;
        }
        const double gf20 = gf[2][0] - gA, gf22 = gf[2][2] - gA;
        // f22 = f02 vR, f20 = f00 vL
        const double gf02 = gf[0][2] + gf22 * vR2;
        gvR2 += gf22 * f02;
        const double gf00 = gf[0][0] + gf20 * vL2;
        gvL2 += gf20 * f00;
        double gbeta2 = gf[1][2] * r22 + gf02 * l2;
        double gr22 = gf[1][2] * beta2;
        double gl2 = gf02 * beta2;
        double gbeta1 = gf[1][1];
        double gbeta0 = gf[1][0] * r20 + gf00 * l0;
        double gr20 = gf[1][0] * beta0;
        double gl0 = gf00 * beta0;
        const double gn0 = gbeta0 * idet, gn1 = gbeta1 * idet, gn2 = gbeta2 * idet;
        const double gdet = -(gbeta0 * beta0 + gbeta1 * beta1 + gbeta2 * beta2) * idet;
        double gd0 = 0, gd1 = 0, gd2 = 0;
        // n2 = -d1 + d0 l0
        gd1 -= gn2;
        gd0 += gn2 * l0;
        gl0 += gn2 * d0;
        // n1 = (A - d0 B) + C
        {
            const double gAq = gn1, gBq = -gn1 * d0, gCq = gn1;
            gd0 -= gn1 * Bq;
            gd1 += gAq * r22;
            gr22 += gAq * d1;
            gl2 -= gAq * d2;
            gd2 -= gAq * l2;
            gl0 += gBq * r22;
            gr22 += gBq * l0;
            gl2 -= gBq * r20;
            gr20 -= gBq * l2;
            gl0 += gCq * d2;
            gd2 += gCq * l0;
            gd1 -= gCq * r20;
            gr20 -= gCq * d1;
        }
        // n0 = -(d0 l2) + d1
        gd0 -= gn0 * l2;
        gl2 -= gn0 * d0;
        gd1 += gn0;
        // det = -l2 + l0
        gl2 -= gdet;
        gl0 += gdet;
        // d2 = delphi - deldelphi; d1 = delhu; d0 = delh - deldelh
        gdelphi += gd2;
        const double gddphi = -gd2;
        gdelhu += gd1;
        const double gdelh = gd0;
        const double gddh = -gd0;
        // deldelphi bounds
        const double gt2 = sd_ ? gddphi : 0.0, glo = sd_ ? 0.0 : gddphi;
        const double gp0 = sc_ ? gt2 : 0.0, ghi = sc_ ? 0.0 : gt2;
        double gpL = (hiL ? ghi * g : 0.0) + (loL ? glo * g : 0.0);
        double gpR = (hiL ? 0.0 : ghi * g) + (loL ? 0.0 : glo * g);
        ghL2 -= gpL * delb;
        gdelb -= gpL * hL2;
        ghR2 -= gpR * delb;
        gdelb -= gpR * hR2;
        double ghbar = 0.0, gs1s2bar = 0.0, gs1s2tilde = 0.0, gM = 0.0;
        if (sonic) {  // p0 = -((g hbar) delb)
            ghbar -= gp0 * g * delb;
            gdelb -= gp0 * (g * hbar);
        } else {  // p0 = -(Kp / s1s2bar), Kp = M s1s2tilde
            const double ib = rcp(s1s2bar);
            const double gKp = -gp0 * ib;
            gs1s2bar += gp0 * Kp * (ib * ib);
            gM += gKp * s1s2tilde;
            gs1s2tilde += gKp * M;
        }
        // deldelh bounds
        double gdd0 = 0.0, gB1 = 0.0, gB2 = 0.0;
        if (ddcase == 0) {
            gdd0 = gddh;
        } else {
            const double gt1 = sb ? gddh : 0.0, gout = sb ? 0.0 : gddh;
            const double gin_a = sa ? gt1 : 0.0, gin_b = sa ? 0.0 : gt1;
            gdd0 = gin_a;
            if (ddcase == 1) {
                gB2 += gin_b;
                gB1 += gout;
            } else if (ddcase == 2) {
                gB1 += gin_b;
                ghL2 -= gout;
            } else {
                ghR2 += gin_b;
                gB2 += gout;
            }
        }
        double gH = 0.0, gdl = 0.0;
        // B2 = (H dl) / l2, B1 = (H dl) / l0 (only where a bound was taken:
        // an unused bound may divide by zero)
        if (gB2 != 0.0) {
            const double il2 = rcp(l2);
            gH += gB2 * dl * il2;
            gdl += gB2 * hstarHLL * il2;
            gl2 -= gB2 * B2 * il2;
        }
        if (gB1 != 0.0) {
            const double il0 = rcp(l0);
            gH += gB1 * dl * il0;
            gdl += gB1 * hstarHLL * il0;
            gl0 -= gB1 * B1 * il0;
        }
        if (sonic) {
            gdelb -= gdd0;
        } else {  // dd0 = M / s1s2bar
            const double ib = rcp(s1s2bar);
            gM += gdd0 * ib;
            gs1s2bar -= gdd0 * dd0 * ib;
        }
        // M = (delb g) hbar
        gdelb += gM * g * hbar;
        ghbar += gM * (delb * g);
        // s1s2tilde = max(0, uu) - g hbar
        if (!uneg) {
            guL2 += gs1s2tilde * uR2;
            guR2 += gs1s2tilde * uL2;
        }
        ghbar -= g * gs1s2tilde;
        // s1s2bar = 0.25 us^2 - g hbar
        {
            const double gus = gs1s2bar * 0.25 * 2.0 * us;
            guL2 += gus;
            guR2 += gus;
            ghbar -= g * gs1s2bar;
        }
        if (hbpos) {
            ghL2 += 0.5 * ghbar;
            ghR2 += 0.5 * ghbar;
        }
        gl2 += 2.0 * l2 * gr22;
        gl0 += 2.0 * l0 * gr20;
        // hstarHLL = max(xh / dl, 0)
        if (hpos) {
            const double idl = rcp(dl);
            const double gxh = gH * idl;
            gdl -= gH * xq * idl;
            ghuL2 += gxh;
            ghuR2 -= gxh;
            gl2 += gxh * hR2;
            ghR2 += gxh * l2;
            gl0 -= gxh * hL2;
            ghL2 -= gxh * l0;
        }
        gl2 += gdl;
        gl0 -= gdl;
        double gs1m = 0.0, gs2m = 0.0;
        if (sel2)
            gsE2 += gl2;
        else
            gs1m += gl2;
        if (sel0)
            gsE1 += gl0;
        else
            gs2m += gl0;
        riemanntype_adj(hL2, hR2, uL2, uR2, sqL, sqR, drytol, g, inv16g, gs1m, gs2m, ghL2, ghR2,
                        guL2, guR2, gsqL, gsqR);
        // delh = hR - hL
        ghR2 += gdelh;
        ghL2 -= gdelh;
        (void)hm;
    }
    // delb, delhu
    gbR2 += gdelb;
    gbL2 -= gdelb;
    ghuR2 += gdelhu;
    ghuL2 -= gdelhu;

    // stage C reverse
    {
        const double gsR = selE2 ? gsE2 : 0.0, ga2 = selE2 ? 0.0 : gsE2;
        const double gsL = selE1 ? gsE1 : 0.0, ga1 = selE1 ? 0.0 : gsE1;
        const double guhat = ga1 + ga2, gchat = ga2 - ga1;
        const double gcx = gchat * dsqrt(chat);
        ghR2 += gcx * (g * 0.5);
        ghL2 += gcx * (g * 0.5);
        const double iden = rcp(den);
        const double gnum = guhat * iden, gden = -guhat * uhat * iden;
        gsqL += gnum * uL2 + gden;
        guL2 += gnum * sqL;
        gsqR += gnum * uR2 + gden;
        guR2 += gnum * sqR;
        guR2 += gsR;
        gsqR += gsR;
        guL2 += gsL;
        gsqL -= gsL;
        // delphi = ((T1 - T2) + T3) - T4
        const double gT1 = gdelphi, gT2 = -gdelphi, gT3 = gdelphi, gT4 = -gdelphi;
        const double du = uL2 + uR2, dhu = huR2 - huL2;
        ghuR2 += gT1 * du;
        ghuL2 -= gT1 * du;
        guL2 += gT1 * dhu;
        guR2 += gT1 * dhu;
        const double uu2 = uL2 * uR2, dh = hR2 - hL2;
        const double gw = gT2 * dh;
        guL2 += gw * uR2;
        guR2 += gw * uL2;
        ghR2 += gT2 * uu2;
        ghL2 -= gT2 * uu2;
        const double E = ((bR2 + hR2) - bL2) - hL2;
        const double a3 = (0.5 * g) * E, sh = hR2 + hL2;
        const double gE = gT3 * sh * (0.5 * g);
        ghR2 += gT3 * a3 + gE;
        ghL2 += gT3 * a3 - gE;
        gbR2 += gE;
        gbL2 -= gE;
        const double c4 = (0.5 * g) * sh, db = bR2 - bL2;
        const double gc4 = gT4 * db;
        gbR2 += gT4 * c4;
        gbL2 -= gT4 * c4;
        ghR2 += gc4 * (0.5 * g);
        ghL2 += gc4 * (0.5 * g);
        ghR2 += gsqR * g * dsqrt(sqR);
        ghL2 += gsqL * g * dsqrt(sqL);
    }

    // stage B reverse (walls)
    double ghL1 = 0, ghR1 = 0, ghuL1 = 0, ghuR1 = 0, guL1 = 0, guR1 = 0, gvL1 = 0, gvR1 = 0;
    double gbL = 0, gbR = 0;
    if (wallR) {
        ghL1 += ghR2;
        ghuL1 -= ghuR2;
        gbL += gbR2;
        guL1 -= guR2;
        gvL1 += gvR2;
    } else {
        ghR1 += ghR2;
        ghuR1 += ghuR2;
        guR1 += guR2;
        gvR1 += gvR2;
        if (clipR) {
            ghL1 += gbR2;
            gbL += gbR2;
        } else {
            gbR += gbR2;
        }
    }
    if (wallL) {
        ghR1 += ghL2;
        ghuR1 -= ghuL2;
        gbR += gbL2;
        guR1 -= guL2;
        gvR1 += gvL2;
    } else {
        ghL1 += ghL2;
        ghuL1 += ghuL2;
        guL1 += guL2;
        gvL1 += gvL2;
        if (clipL) {
            ghR1 += gbL2;
            gbR += gbL2;
        } else {
            gbL += gbL2;
        }
    }
    // stage A reverse
    if (wetR) {
        const double ih = ihR;
        gin[1] += ghR1 - guR1 * uR1 * ih - gvR1 * vR1 * ih;
        gin[3] += ghuR1 + guR1 * ih;
        gin[5] += gvR1 * ih;
    }
    if (wetL) {
        const double ih = ihL;
        gin[0] += ghL1 - guL1 * uL1 * ih - gvL1 * vL1 * ih;
        gin[2] += ghuL1 + guL1 * ih;
        gin[4] += gvL1 * ih;
    }
    gin[6] += gbL;
    gin[7] += gbR;
}

// Reverse of edge_flux<SOLVER>: seeds gfL/gfR on (h, hu, hv) of the left and
// right fluxes; accumulates gradients into gQL/gQR = (eta, hu, hv, b).
template <int SOLVER>
SWE_HD void edge_flux_adj(double nx, double ny, double etaL, double huL, double hvL, double bL,
                          double etaR, double huR, double hvR, double bR, double drytol,
                          double g, double inv16g, const double gfL[3], const double gfR[3],
                          double gQL[4], double gQR[4]) {
    const double pL1 = nx * huL + ny * hvL;
    const double pL2 = (-ny) * huL + nx * hvL;
    const double pR1 = nx * huR + ny * hvR;
    const double pR2 = (-ny) * huR + nx * hvR;
    // f.hu = n1 nx - n2 ny, f.hv = n1 ny + n2 nx  ->  transpose
    const double gnL[3] = {gfL[0], gfL[1] * nx + gfL[2] * ny, -gfL[1] * ny + gfL[2] * nx};
    const double gnR[3] = {gfR[0], gfR[1] * nx + gfR[2] * ny, -gfR[1] * ny + gfR[2] * nx};
    double gin[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    geoclaw_adj<SOLVER>(etaL - bL, etaR - bR, pL1, pR1, pL2, pR2, bL, bR, drytol, g, inv16g, gnL,
                        gnR, gin);
    // h = eta - b
    gQL[0] += gin[0];
    gQL[3] += gin[6] - gin[0];
    gQR[0] += gin[1];
    gQR[3] += gin[7] - gin[1];
    // p1 = nx hu + ny hv, p2 = -ny hu + nx hv
    gQL[1] += gin[2] * nx - gin[4] * ny;
    gQL[2] += gin[2] * ny + gin[4] * nx;
    gQR[1] += gin[3] * nx - gin[5] * ny;
    gQR[2] += gin[3] * ny + gin[5] * nx;
}

}  // namespace swe
