"""Plain FP64 PyTorch restatement of the SWE Euler step (TEST INFRASTRUCTURE).

The same discrete scheme as the reference path (SWE_euler_timestep.f90
:257-355,449-481; c_bind_riemannsolvers.f90:16-176; riemannsolvers.f:2-258,
446-500,507-618), vectorised over edges with torch.where for every branch,
so that torch.autograd yields the exact discrete adjoint of the scheme with
the branch decisions of the forward pass (the "frozen masks").  Used only by
tests to check the hand-written adjoint kernels, on small meshes.

Branches are evaluated on guarded inputs (denominators replaced by 1 where a
branch is not taken) so that untaken branches cannot inject inf/NaN into the
gradients.
"""
from __future__ import annotations

import numpy as np
import torch

HUGE = 1.7976931348623157e308


def _w(c, a, b):
    return torch.where(c, a, b)


class _Sqrt(torch.autograd.Function):
    """Correctly rounded sqrt (torch's vectorised CPU sqrt is not always) with
    the derivative taken as 0 where the value is exactly 0 (a dry side)."""

    @staticmethod
    def forward(ctx, x):
        y = torch.from_numpy(np.sqrt(x.detach().cpu().numpy())).to(x.device)
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, gy):
        (y,) = ctx.saved_tensors
        pos = y > 0.0
        return torch.where(pos, gy / torch.where(pos, 2.0 * y, torch.ones_like(y)), torch.zeros_like(y))


def _sqrt(x):
    return _Sqrt.apply(x)


def _safe(x, ok):
    """x where ok, else 1 (a harmless value for untaken branches)."""
    return torch.where(ok, x, torch.ones_like(x))


def riemanntype(hL, hR, uL, uR, sqL, sqR, drytol, g):
    """riemannsolvers.f:507-618 (maxiter = 1); returns s1m, s2m, hm."""
    r_min = hR < hL
    h_min = _w(r_min, hR, hL)
    h_max = _w(hR > hL, hR, hL)
    sqmin = _w(r_min, sqR, sqL)
    sqmax = _w(hR > hL, sqR, sqL)
    delu = uR - uL
    dry = h_min <= drytol
    wet = ~dry
    hmin_s = _safe(h_min, wet)
    hmax_s = _safe(h_max, wet)
    F_min = delu + 2.0 * (sqmin - sqmax)
    F_max = delu + (h_max - h_min) * _sqrt(((0.5 * g) * (hmax_s + hmin_s)) / (hmax_s * hmin_s))
    rare2 = wet & (F_min > 0.0)
    shock2 = wet & ~(F_min > 0.0) & (F_max <= 0.0)
    mixed = wet & ~(F_min > 0.0) & ~(F_max <= 0.0)
    # dry
    s_dry = ((uR + uL) - 2.0 * sqR) + 2.0 * sqL
    # two rarefactions
    tt = -delu + 2.0 * (sqL + sqR)
    t = _w(tt > 0.0, tt, torch.zeros_like(tt))
    hm_r = (1.0 / (16.0 * g)) * (t * t)
    sqm_r = _sqrt(_w(rare2, g * hm_r, torch.ones_like(hm_r)))
    s1_r = (uL + 2.0 * sqL) - 3.0 * sqm_r
    s2_r = (uR - 2.0 * sqR) + 3.0 * sqm_r
    # two shocks
    h0 = _safe(h_max, shock2)
    hLs, hRs = _safe(hL, shock2), _safe(hR, shock2)
    gL = _sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hLs))
    gR = _sqrt((0.5 * g) * (1.0 / h0 + 1.0 / hRs))
    F0 = (delu + (h0 - hLs) * gL) + (h0 - hRs) * gR
    h02 = h0 * h0
    dfdh = ((gL - (g * (h0 - hLs)) / ((4.0 * h02) * gL)) + gR) - (g * (h0 - hRs)) / ((4.0 * h02) * gR)
    sq0 = _sqrt(h0)
    slope = (2.0 * sq0) * dfdh
    q = sq0 - F0 / _safe(slope, shock2)
    hm_s = _safe(q * q, shock2)
    u1m = uL - (hm_s - hLs) * _sqrt((0.5 * g) * (1.0 / hm_s + 1.0 / hLs))
    u2m = uR + (hm_s - hRs) * _sqrt((0.5 * g) * (1.0 / hm_s + 1.0 / hRs))
    sqm_s = _sqrt(g * hm_s)
    s1_s = u1m - sqm_s
    s2_s = u2m + sqm_s
    # one shock, one rarefaction
    hmin_m = _safe(h_min, mixed)
    hmax_m = _safe(h_max, mixed)
    F0m = (delu + 2.0 * (sqmin - sqmax)) + (hmin_m - hmin_m) * _sqrt(
        (0.5 * g) * (1.0 / hmin_m + 1.0 / hmin_m))
    den = _safe(hmax_m - hmin_m, mixed)
    slope_m = (F_max - F0m) / den
    hm_m = hmin_m - F0m / _safe(slope_m, mixed)
    sqm_m = _sqrt(_w(mixed, g * hm_m, torch.ones_like(hm_m)))
    lgt = hL > hR
    s1_m = _w(lgt, (uL + 2.0 * sqL) - 3.0 * sqm_m, (uR - 2.0 * sqR) + sqm_m)
    s2_m = _w(lgt, (uL + 2.0 * sqL) - sqm_m, (uR - 2.0 * sqR) + 3.0 * sqm_m)
    s1 = _w(dry, s_dry, _w(rare2, s1_r, _w(shock2, s1_s, s1_m)))
    s2 = _w(dry, s_dry, _w(rare2, s2_r, _w(shock2, s2_s, s2_m)))
    hm = _w(dry, torch.zeros_like(hL), _w(rare2, hm_r, _w(shock2, hm_s, hm_m)))
    return s1, s2, hm


def geoclaw(hL, hR, huL, huR, hvL, hvR, bL, bR, drytol, g, solver=0):
    """c_bind_riemannsolvers.f90:16-176 in the edge frame; returns netL, netR
    (each a tuple of 3) and the max wave speed."""
    z = torch.zeros_like(hL)
    active = ~((hL <= drytol) & (hR <= drytol))
    wetR = hR > drytol
    wetL = hL > drytol
    hRs, hLs = _safe(hR, wetR), _safe(hL, wetL)
    uR = _w(wetR, huR / hRs, z)
    vR = _w(wetR, hvR / hRs, z)
    hR = _w(wetR, hR, z)
    huR = _w(wetR, huR, z)
    hvR = _w(wetR, hvR, z)
    uL = _w(wetL, huL / hLs, z)
    vL = _w(wetL, hvL / hLs, z)
    hL = _w(wetL, hL, z)
    huL = _w(wetL, huL, z)
    hvL = _w(wetL, hvL, z)
    one = torch.ones_like(hL)
    # wall tests (right dry, else left dry)
    caseR = active & (hR <= drytol)
    caseL = active & ~(hR <= drytol) & (hL <= drytol)
    sqa = _sqrt(g * hL)
    _, _, hstarR = riemanntype(hL, hL, uL, -uL, sqa, sqa, drytol, g)
    sqb = _sqrt(g * hR)
    _, _, hstarL = riemanntype(hR, hR, -uR, uR, sqb, sqb, drytol, g)
    wallR = caseR & (_w(hL > hstarR, hL, hstarR) + bL < bR)
    clipR = caseR & ~wallR & (hL + bL < bR)
    wallL = caseL & (_w(hR > hstarL, hR, hstarL) + bR < bL)
    clipL = caseL & ~wallL & (hR + bR < bL)
    w0 = _w(wallL, z, one)
    w1 = _w(wallL | wallR, z, one)
    w2 = _w(wallR, z, one)
    hR2 = _w(wallR, hL, hR)
    huR2 = _w(wallR, -huL, huR)
    bR2 = _w(wallR, bL, _w(clipR, hL + bL, bR))
    uR2 = _w(wallR, -uL, uR)
    vR2 = _w(wallR, vL, vR)
    hL2 = _w(wallL, hR, hL)
    huL2 = _w(wallL, -huR, huL)
    bL2 = _w(wallL, bR, _w(clipL, hR + bR, bL))
    uL2 = _w(wallL, -uR, uL)
    vL2 = _w(wallL, vR, vL)
    hR, huR, bR, uR, vR = hR2, huR2, bR2, uR2, vR2
    hL, huL, bL, uL, vL = hL2, huL2, bL2, uL2, vL2

    delphi = (((huR - huL) * (uL + uR) - (uL * uR) * (hR - hL)) +
              ((0.5 * g) * (((bR + hR) - bL) - hL)) * (hR + hL)) - ((0.5 * g) * (hR + hL)) * (bR - bL)
    sqL = _sqrt(g * hL)
    sqR = _sqrt(g * hR)
    sL = uL - sqL
    sR = uR + sqR
    den = _safe(sqR + sqL, active)
    uhat = (sqL * uL + sqR * uR) / den
    chat = _sqrt((g * 0.5) * (hR + hL))
    sE1 = _w(sL < uhat - chat, sL, uhat - chat)
    sE2 = _w(sR > uhat + chat, sR, uhat + chat)
    delhu = huR - huL
    delb = bR - bL
    if solver == 1:
        deldelphi = -(((g * 0.5) * (hR + hL)) * delb)
        dec = delphi - deldelphi
        dd = _safe(sE2 - sE1, active)
        beta1 = (sE2 * delhu - dec) / dd
        beta2 = (dec - sE1 * delhu) / dd
        sw = (sE1, 0.5 * (sE1 + sE2), sE2)
        f0 = (beta1, beta1 * sE1, beta1 * vL)
        f2 = (beta2, beta2 * sE2, beta2 * vR)
        f1 = (z, z, ((huR * vR - huL * vL) - f0[2]) - f2[2])
    else:
        delh = hR - hL
        s1m, s2m, _ = riemanntype(hL, hR, uL, uR, sqL, sqR, drytol, g)
        l0 = _w(sE1 < s2m, sE1, s2m)
        l2 = _w(sE2 > s1m, sE2, s1m)
        sE1, sE2 = l0, l2
        dd = _safe(sE2 - sE1, active)
        x = (((huL - huR) + sE2 * hR) - sE1 * hL) / dd
        hstarHLL = _w(x > 0.0, x, z)
        l1 = 0.5 * (l0 + l2)
        l1 = _w(torch.abs(l1) < 1.0e-20, z, l1)
        r20, r22 = l0 * l0, l2 * l2
        tol = 1.0e-6
        hbar = _w(0.5 * (hL + hR) > 0.0, 0.5 * (hL + hR), z)
        s1s2bar = 0.25 * ((uL + uR) * (uL + uR)) - g * hbar
        s1s2tilde = _w(uL * uR > 0.0, uL * uR, z) - g * hbar
        sonic = torch.abs(s1s2bar) <= tol
        sonic = sonic | (s1s2bar * s1s2tilde <= tol)
        sonic = sonic | ((s1s2bar * sE1) * sE2 <= tol)
        sonic = sonic | (_w(torch.abs(sE1) < torch.abs(sE2), torch.abs(sE1), torch.abs(sE2)) < tol)
        sonic = sonic | ((sE1 < 0.0) & (s1m > 0.0))
        sonic = sonic | ((sE2 > 0.0) & (s2m < 0.0))
        sonic = sonic | ((uL + sqL) * (uR + sqR) < 0.0)
        sonic = sonic | ((uL - sqL) * (uR - sqR) < 0.0)
        sb = _safe(s1s2bar, ~sonic & active)
        deldelh = _w(sonic, -delb, ((delb * g) * hbar) / sb)
        c1 = (sE1 < -tol) & (sE2 > tol)
        c2 = ~c1 & (sE1 >= tol)
        c3 = ~c1 & ~c2 & (sE2 <= -tol)
        sE1s = _safe(sE1, c1 | c2)
        sE2s = _safe(sE2, c1 | c3)
        b_hi = (hstarHLL * (sE2 - sE1)) / sE2s
        b_lo = (hstarHLL * (sE2 - sE1)) / sE1s
        m1 = _w(deldelh < b_hi, deldelh, b_hi)
        m1 = _w(m1 > b_lo, m1, b_lo)
        m2 = _w(deldelh < b_lo, deldelh, b_lo)
        m2 = _w(m2 > -hL, m2, -hL)
        m3 = _w(deldelh < hR, deldelh, hR)
        m3 = _w(m3 > b_hi, m3, b_hi)
        deldelh = _w(c1, m1, _w(c2, m2, _w(c3, m3, deldelh)))
        deldelphi = _w(sonic, -((g * hbar) * delb), -((((delb * g) * hbar) * s1s2tilde) / sb))
        pL, pR = -(hL * delb), -(hR * delb)
        hi = g * _w(pL > pR, pL, pR)
        lo = g * _w(pL < pR, pL, pR)
        deldelphi = _w(deldelphi < hi, deldelphi, hi)
        deldelphi = _w(deldelphi > lo, deldelphi, lo)
        d0 = delh - deldelh
        d1 = delhu
        d2 = delphi - deldelphi
        det = _safe(-l2 + l0, active)
        beta0 = (-(d0 * l2) + d1) / det
        beta1 = ((d1 * r22 - l2 * d2) - d0 * (l0 * r22 - l2 * r20) + (l0 * d2 - d1 * r20)) / det
        beta2 = (-d1 + d0 * l0) / det
        sw = (l0, l1, l2)
        f0 = (beta0 * l0, beta0 * r20, (beta0 * l0) * vL)
        f2 = (beta2 * l2, beta2 * r22, (beta2 * l2) * vR)
        f1 = (beta1 * 0.0, beta1, (((hR * uR) * vR - (hL * uL) * vL) - f0[2]) - f2[2])
    walls = (w0, w1, w2)
    sw = tuple(s * w for s, w in zip(sw, walls))
    fws = [tuple(f * w for f in fw) for fw, w in zip((f0, f1, f2), walls)]
    netL = [z, z, z]
    netR = [z, z, z]
    for m in range(3):
        for mw in range(3):
            f = fws[mw][m]
            neg = sw[mw] < 0.0
            pos = sw[mw] > 0.0
            netL[m] = netL[m] + _w(neg, f, _w(pos, z, 0.5 * f))
            netR[m] = netR[m] + _w(pos, f, _w(neg, z, 0.5 * f))
    a0, a1, a2 = torch.abs(sw[0]), torch.abs(sw[1]), torch.abs(sw[2])
    mx = _w(a1 > a0, a1, a0)
    mx = _w(a2 > mx, a2, mx)
    netL = [_w(active, x, z) for x in netL]
    netR = [_w(active, x, z) for x in netR]
    mx = _w(active, mx, z)
    return netL, netR, mx


def edge_flux(nx, ny, QL, QR, drytol, g, solver=0):
    """compute_geoclaw_flux (SWE_euler_timestep.f90:449-481)."""
    eL, uL, vL, bL = QL
    eR, uR, vR, bR = QR
    pL1 = nx * uL + ny * vL
    pL2 = (-ny) * uL + nx * vL
    pR1 = nx * uR + ny * vR
    pR2 = (-ny) * uR + nx * vR
    netL, netR, s = geoclaw(eL - bL, eR - bR, pL1, pR1, pL2, pR2, bL, bR, drytol, g, solver)
    fL = (netL[0], netL[1] * nx + netL[2] * (-ny), netL[1] * ny + netL[2] * nx, s)
    fR = (netR[0], netR[1] * nx + netR[2] * (-ny), netR[1] * ny + netR[2] * nx, s)
    return fL, fR


class TorchMesh:
    """Edge lists of a mesh for the vectorised step."""

    def __init__(self, depth, plotter, nbr, normals, legs):
        n = len(depth)
        self.n = n
        self.depth = torch.tensor(np.asarray(depth, np.int64))
        nbr = np.asarray(nbr).reshape(n, 3)
        c = np.repeat(np.arange(n), 3)
        k = np.tile(np.arange(3), n)
        o = nbr.reshape(-1)
        nt = np.asarray(normals).reshape(17, 3, 2)
        nrm = nt[np.asarray(plotter, np.int64)[c] + 8, k]
        self.c, self.k, self.o = (torch.tensor(x) for x in (c, k, o))
        self.nx = torch.tensor(nrm[:, 0])
        self.ny = torch.tensor(nrm[:, 1])
        self.legs = torch.tensor(np.asarray(legs))


def step(tm: TorchMesh, eta, hu, hv, b, dt, forward, S, g, drytol, bc=0, solver=0):
    """One Euler traversal; every half-edge (c, k) is solved from c's side
    with the left/right choice of the traversal direction (both sides of an
    interior edge give the same fluxes).  Returns (eta', hu', hv', cand)."""
    c, o = tm.c, tm.o
    bnd = o < 0
    oo = torch.where(bnd, c, o)
    Qc = (eta[c], hu[c], hv[c], b[c])
    Qo = (eta[oo], hu[oo], hv[oo], b[oo])
    if bc == 1:
        dp = Qc[1] * tm.nx + Qc[2] * tm.ny
        gu = Qc[1] - (2.0 * dp) * tm.nx
        gv = Qc[2] - (2.0 * dp) * tm.ny
    else:
        gu, gv = Qc[1], Qc[2]
    Qo = (torch.where(bnd, Qc[0], Qo[0]), torch.where(bnd, gu, Qo[1]),
          torch.where(bnd, gv, Qo[2]), torch.where(bnd, Qc[3], Qo[3]))
    c_left = bnd | ((c < oo) if forward else (c > oo))
    sx = torch.where(c_left, tm.nx, -tm.nx)
    sy = torch.where(c_left, tm.ny, -tm.ny)
    QL = tuple(torch.where(c_left, x, y) for x, y in zip(Qc, Qo))
    QR = tuple(torch.where(c_left, y, x) for x, y in zip(Qc, Qo))
    fL, fR = edge_flux(sx, sy, QL, QR, drytol, g, solver)
    f = [torch.where(c_left, a, bb).reshape(-1, 3) for a, bb in zip(fL, fR)]
    d = tm.depth
    vol = (S * S) * torch.pow(torch.tensor(0.5, dtype=torch.float64), (d + 1).to(torch.float64))
    lleg = S * tm.legs[d + 1]
    lhyp = S * tm.legs[d]
    order = (0, 1, 2) if forward else (2, 1, 0)
    lens = (lleg, lhyp, lleg)
    acc = []
    for comp in range(4):
        x = torch.zeros_like(eta)
        for i, kk in enumerate(order):
            x = x + lens[i] * f[comp][:, kk]
        acc.append(x)
    dh, d1, d2, sv = acc
    cand = vol / sv
    # tensor/tensor: torch evaluates float/tensor as reciprocal(tensor)*float
    scale = torch.full_like(vol, -dt) / vol if not torch.is_tensor(dt) else -dt / vol
    dh, d1, d2 = scale * dh, scale * d1, scale * d2
    wet = (eta < b + drytol) & (dh > 0.0)
    zero = torch.zeros_like(eta)
    qh = torch.where(wet, b + drytol, eta)
    q1 = torch.where(wet, zero, hu)
    q2 = torch.where(wet, zero, hv)
    qh = qh + dh
    q1 = q1 + d1
    q2 = q2 + d2
    dry = qh < b + drytol
    qh = torch.where(dry, torch.where(b < 0.0, b, zero), qh)
    q1 = torch.where(dry, zero, q1)
    q2 = torch.where(dry, zero, q2)
    return qh, q1, q2, cand
