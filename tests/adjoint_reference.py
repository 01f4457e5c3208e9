"""Differentiable restatement of one SSP-RK3 step in torch (CPU, float64) -- test
infrastructure for the discrete adjoint.

It performs the operations of the reference stage in the reference's order, so its
forward values equal the oracle's bit for bit (checked by the CPU suite), and every
branch (limiter cases, MOOD resets, face and conserved floors, thermal resets, Riemann
regions and catches) is a ``torch.where`` on values identical to the forward pass.
Autograd through it therefore gives the exact Jacobian of the branch taken -- also at
MOOD decisions that sit exactly on a bound (the MC limiter puts face values there),
where finite differences cannot be used.  Ties of max/min pass the derivative of the
first argument, as the GPU adjoint does.

Reference: src/main.f90:65-303 (step), src/main.f90:346-403 (stage fluxes),
src/mhd_derivatives.f90:50-293 (limiter, reconstruction + MOOD, thermal check),
src/mhd_flux.f90:20-779 (update, CT, Rusanov / HLLE / HLLD),
src/mhd_change_states.f90:61-166 (primitives, floors), src/mhd_field_ops.f90 (curl,
face -> cell B), src/mhd_bc.f90:36-141 (ghost cells: wrap or copy).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

P_FLOOR = 1.0e-12
RHO_FLOOR = 1.0e-12
E_FLOOR_FRAC = 1.0e-10
V_MAX = 50.0


class _Sqrt(torch.autograd.Function):
    """Correctly rounded square root (IEEE, as the oracle and the GPU compute it);
    torch's CPU sqrt kernel is not correctly rounded in every lane."""

    @staticmethod
    def forward(ctx, x):
        y = torch.from_numpy(np.sqrt(x.detach().numpy()))
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return g * (0.5 / y)


def sqrt(x):
    return _Sqrt.apply(x)


def _where(c, a, b):
    if not torch.is_tensor(a):
        a = torch.full_like(b, float(a))
    if not torch.is_tensor(b):
        b = torch.full_like(a, float(b))
    return torch.where(c, a, b)


def dmax(a, b):
    """max passing the first argument on ties (value identical to fmax)."""
    if not torch.is_tensor(b):
        return _where(b > a, b, a)
    return torch.where(b > a, b, a)


def dmin(a, b):
    if not torch.is_tensor(b):
        return _where(b < a, b, a)
    return torch.where(b < a, b, a)


@dataclass
class Geometry:
    nx: int
    ny: int
    dx: float
    gamma: float
    px: bool
    py: bool
    solver: int
    limiter: int
    use_lim: bool = True
    mood: bool = True

    @classmethod
    def of(cls, prob):
        bc = prob.bc
        s = prob.scheme
        return cls(prob.nx, prob.ny, prob.dx, prob.gamma, bc[0] == 1, bc[2] == 1, s[0], s[1],
                   bool(s[2]), bool(s[3]))


def _nbr(n, per, off):
    i = torch.arange(n) + off
    return i % n if per else i.clamp(0, n - 1)


def _sx(f, idx):
    return f[:, idx]


def _sy(f, idx):
    return f[idx, :]


# ---- point physics -------------------------------------------------------------

def get_primitive(M, Mx, My, E, Bx, By, g, vol):
    """src/mhd_change_states.f90:91-110"""
    a = M / vol
    r = _where(RHO_FLOOR > a, RHO_FLOOR, a)
    u = Mx / (r * vol)
    w = My / (r * vol)
    halfB2 = 0.5 * (Bx * Bx + By * By)
    ke = 0.5 * r * (u * u + w * w)
    p = (E / vol - ke - halfB2) * (g - 1.0) + halfB2
    p = _where(P_FLOOR > p, P_FLOOR, p)
    q = halfB2 + E_FLOOR_FRAC * (ke + halfB2)
    p = torch.where(q > p, q, p)
    return r, u, w, p


def conserved_floors(M, Mx, My, E, Bx, By, vol, g):
    """src/mhd_change_states.f90:143-164 (B: the stage's cell B)."""
    halfB2 = 0.5 * (Bx * Bx + By * By)
    low = M < RHO_FLOOR * vol
    Mx = _where(low, 0.0, Mx)
    My = _where(low, 0.0, My)
    M = _where(low, RHO_FLOOR * vol, M)
    Ef = (P_FLOOR / (g - 1.0) + halfB2) * vol
    E = torch.where(low & (Ef > E), Ef, E)
    with torch.no_grad():
        mom = sqrt(Mx * Mx + My * My)
        cap = mom / M > V_MAX
    mom_s = sqrt(_where(cap, Mx * Mx + My * My, 1.0))
    scale = V_MAX * M / mom_s
    Mx = torch.where(cap, Mx * scale, Mx)
    My = torch.where(cap, My * scale, My)
    return M, Mx, My, E


def limited_slope(fm, f0, fp, inv_dx, limiter):
    """src/mhd_derivatives.f90:82-97"""
    dfL = (f0 - fm) * inv_dx
    dfR = (fp - f0) * inv_dx
    prod = dfL * dfR
    den = dfL + dfR
    zero = (prod <= 0.0) | (den.abs() < 1e-12)
    if limiter == 1:
        s = (2.0 * prod) / _where(zero, 1.0, den)
    else:
        m1, m2, m3 = 2.0 * dfL.abs(), 2.0 * dfR.abs(), 0.5 * den.abs()
        m12 = torch.where(m2 < m1, m2, m1)
        m = torch.where(m3 < m12, m3, m12)
        s = torch.copysign(torch.ones_like(dfL), dfL) * m
    return _where(zero, 0.0, s)


# ---- Riemann solvers (face arrays; L, R = (rho, u, w, P, Bn, Bt)) --------------

def _energy(P, rho, u, w, Bn, Bt, g):
    halfB2 = 0.5 * (Bn * Bn + Bt * Bt)
    pth = dmax(P - halfB2, P_FLOOR)
    en = pth / (g - 1.0) + 0.5 * rho * (u * u + w * w) + halfB2
    return halfB2, pth, en


def rusanov(L, R, g):
    """src/mhd_flux.f90:197-288"""
    rL, uL, wL, PL, BL, CL = L
    rR, uR, wR, PR, BR, CR = R
    hL, pthL, enL = _energy(PL, rL, uL, wL, BL, CL, g)
    hR, pthR, enR = _energy(PR, rR, uR, wR, BR, CR, g)
    inv = 1.0 / (0.5 * (rL + rR))
    mx = 0.5 * (rL * uL + rR * uR)
    my = 0.5 * (rL * wL + rR * wR)
    en = 0.5 * (enL + enR)
    bx = 0.5 * (BL + BR)
    by = 0.5 * (CL + CR)
    hb = 0.5 * (bx * bx + by * by)
    Pavg = (g - 1.0) * (en - 0.5 * (mx * mx + my * my) * inv - hb) + hb
    cL2 = (g * pthL + 2.0 * hL) / rL
    cR2 = (g * pthR + 2.0 * hR) / rR
    sL = sqrt(0.5 * (cL2 + cL2.abs())) + uL.abs()
    sR = sqrt(0.5 * (cR2 + cR2.abs())) + uR.abs()
    C = 0.5 * dmax(sL, sR)
    return [mx - C * (rR - rL),
            (mx * mx * inv + Pavg - bx * bx) - C * (rR * uR - rL * uL),
            (mx * my * inv - bx * by) - C * (rR * wR - rL * wL),
            ((en + Pavg) * mx * inv - bx * (bx * mx + by * my) * inv) - C * (enR - enL),
            ((by * mx - bx * my) * inv) - C * (CR - CL)]


def _phys_flux(r, u, w, P, Bn, Bt, en, hlld_form):
    m = r * u
    mom_n = r * (u * u) + P - Bn * Bn if hlld_form else m * u + P - Bn * Bn
    return [m, mom_n, m * w - Bn * Bt, (en + P) * u - Bn * (Bn * u + Bt * w), Bt * u - Bn * w]


def _hll(FL, FR, SL, SR, L, R, enL, enR, floor_den):
    inv = 1.0 / (dmax(SR - SL, 1.0e-12) if floor_den else (SR - SL))
    dU = [R[0] - L[0], R[0] * R[1] - L[0] * L[1], R[0] * R[2] - L[0] * L[2], enR - enL, R[5] - L[5]]
    sLR = SL * SR
    return [(SR * FL[c] - SL * FR[c] + sLR * dU[c]) * inv for c in range(5)]


def hlle(L, R, g):
    """src/mhd_flux.f90:291-405"""
    rL, uL, wL, PL, BL, CL = L
    rR, uR, wR, PR, BR, CR = R
    hL, pthL, enL = _energy(PL, rL, uL, wL, BL, CL, g)
    hR, pthR, enR = _energy(PR, rR, uR, wR, BR, CR, g)
    xL = (g * pthL + 2.0 * hL) / rL
    xR = (g * pthR + 2.0 * hR) / rR
    cfL = dmin(sqrt(0.5 * (xL + xL.abs())), 1.0e2)
    cfR = dmin(sqrt(0.5 * (xR + xR.abs())), 1.0e2)
    SL = dmin(uL - cfL, uR - cfR)
    SR = dmax(uL + cfL, uR + cfR)
    FL = _phys_flux(rL, uL, wL, PL, BL, CL, enL, False)
    FR = _phys_flux(rR, uR, wR, PR, BR, CR, enR, False)
    H = _hll(FL, FR, SL, SR, L, R, enL, enR, True)
    out = []
    for c in range(5):
        v = torch.where(SL >= 0.0, FL[c], H[c])
        out.append(torch.where(SR <= 0.0, FR[c], v))
    return out


def hlld(L, R, g):
    """src/mhd_flux.f90:408-779.  Star-region quantities are evaluated on benign
    stand-in states where that region is not used, so that no unused lane produces a
    non-finite value (which would turn zero gradients into NaN)."""
    rL, uL, wL, pTL, BL, CL = L
    rR, uR, wR, pTR, BR, CR = R
    hL, pthL, enL = _energy(pTL, rL, uL, wL, BL, CL, g)
    hR, pthR, enR = _energy(pTR, rR, uR, wR, BR, CR, g)
    FL = _phys_flux(rL, uL, wL, pTL, BL, CL, enL, True)
    FR = _phys_flux(rR, uR, wR, pTR, BR, CR, enR, True)

    def speeds(rL, uL, pthL, BL, CL, rR, uR, pthR, BR, CR):
        BL2, BR2 = BL * BL, BR * BR
        caL, caR = BL2 / rL, BR2 / rR
        csL2, csR2 = g * pthL / rL, g * pthR / rR
        b2L, b2R = (BL2 + CL * CL) / rL, (BR2 + CR * CR) / rR
        tL, tR = csL2 + b2L, csR2 + b2R
        cfL = sqrt(0.5 * (tL + sqrt(dmax(tL * tL - 4.0 * csL2 * caL, 0.0))))
        cfR = sqrt(0.5 * (tR + sqrt(dmax(tR * tR - 4.0 * csR2 * caR, 0.0))))
        return dmin(cfL, 50.0), dmin(cfR, 50.0)

    cfL, cfR = speeds(rL, uL, pthL, BL, CL, rR, uR, pthR, BR, CR)
    SL = dmin(uL - cfL, uR - cfR)
    SR = dmax(uL + cfL, uR + cfR)
    catch1 = (cfL >= 50.0) | (cfR >= 50.0)
    edgeL = ~catch1 & (SL >= 0.0)
    edgeR = ~catch1 & ~edgeL & (SR <= 0.0)
    rest = ~(catch1 | edgeL | edgeR)

    # star region on the lanes that reach it; benign states elsewhere
    one = torch.ones_like(rL)
    zero = torch.zeros_like(rL)
    sub = lambda a, b: torch.where(rest, a, b)  # noqa: E731
    rL_, uL_, wL_, pTL_, BL_, CL_ = sub(rL, one), sub(uL, -one), sub(wL, zero), sub(pTL, 2 * one), sub(BL, one), sub(CL, zero)
    rR_, uR_, wR_, pTR_, BR_, CR_ = sub(rR, one), sub(uR, one), sub(wR, zero), sub(pTR, 2 * one), sub(BR, one), sub(CR, zero)
    hL_, pthL_, enL_ = _energy(pTL_, rL_, uL_, wL_, BL_, CL_, g)
    hR_, pthR_, enR_ = _energy(pTR_, rR_, uR_, wR_, BR_, CR_, g)
    cfL_, cfR_ = speeds(rL_, uL_, pthL_, BL_, CL_, rR_, uR_, pthR_, BR_, CR_)
    SL_ = dmin(uL_ - cfL_, uR_ - cfR_)
    SR_ = dmax(uL_ + cfL_, uR_ + cfR_)
    FL_ = _phys_flux(rL_, uL_, wL_, pTL_, BL_, CL_, enL_, True)
    FR_ = _phys_flux(rR_, uR_, wR_, pTR_, BR_, CR_, enR_, True)
    aL = rL_ * (SL_ - uL_)
    aR = rR_ * (SR_ - uR_)
    denMD = aR - aL
    with torch.no_grad():
        tiny = denMD.abs() < 1.0e-15
        sgn_md = torch.copysign(torch.full_like(denMD, 1.0e-15), denMD)
    denMD = torch.where(tiny, sgn_md, denMD)
    SM = (rR_ * uR_ * (SR_ - uR_) - rL_ * uL_ * (SL_ - uL_) - (pTR_ - pTL_)) / denMD
    pTs = pTL_ + aL * (SM - uL_)
    catch2 = rest & ((SM <= SL_ + 1.0e-6 * (SR_ - SL_)) | (SM >= SR_ - 1.0e-6 * (SR_ - SL_)))
    star = rest & ~catch2
    pTs = dmax(pTs, P_FLOOR)
    dSL = torch.where(star, SL_ - SM, one)
    dSR = torch.where(star, SR_ - SM, one)
    BL2, BR2 = BL_ * BL_, BR_ * BR_

    def side(a, dS, r, u, w, B, B2, C, S, en, pT, vdotB):
        rs = a / dS
        k = a * dS
        den = k - B2
        with torch.no_grad():
            use = den.abs() > 1.0e-3 * torch.maximum(torch.maximum(k.abs(), B2),
                                                     torch.full_like(k, 1.0e-30))
        den_s = torch.where(use, den, one)
        t = S - u
        vs = torch.where(use, w - B * C * (SM - u) / den_s, w)
        Bs = torch.where(use, C * (r * (t * t) - B2) / den_s, C)
        vdotBs = SM * B + vs * Bs
        ens = (en * (S - u) - pT * u + pTs * SM + B * (vdotB - vdotBs)) / dS
        return rs, vs, Bs, vdotBs, ens

    vdotBL_ = uL_ * BL_ + wL_ * CL_
    vdotBR_ = uR_ * BR_ + wR_ * CR_
    rLs, vLs, BLs, vdotBLs, enLs = side(aL, dSL, rL_, uL_, wL_, BL_, BL2, CL_, SL_, enL_, pTL_, vdotBL_)
    rRs, vRs, BRs, vdotBRs, enRs = side(aR, dSR, rR_, uR_, wR_, BR_, BR2, CR_, SR_, enR_, pTR_, vdotBR_)
    sqL = sqrt(dmax(rLs, RHO_FLOOR))
    sqR = sqrt(dmax(rRs, RHO_FLOOR))
    Bavg = 0.5 * (BL_ + BR_)
    big = Bavg.abs() > 1.0e-10
    with torch.no_grad():
        sgn = torch.copysign(torch.ones_like(Bavg), Bavg)
    SLs = torch.where(big, SM - Bavg.abs() / sqL, SM)
    SRs = torch.where(big, SM + Bavg.abs() / sqR, SM)
    dss = sqL + sqR
    vss = torch.where(big, (sqL * vLs + sqR * vRs + sgn * (BRs - BLs)) / dss, 0.5 * (vLs + vRs))
    Bss = torch.where(big, (sqL * BRs + sqR * BLs + sgn * sqL * sqR * (vRs - vLs)) / dss,
                      0.5 * (BLs + BRs))
    vdotBss = SM * Bavg + vss * Bss
    enLss = torch.where(big, enLs - sqL * sgn * (vdotBLs - vdotBss), enLs)
    enRss = torch.where(big, enRs + sqR * sgn * (vdotBRs - vdotBss), enRs)

    left = SM >= 0.0
    fl = [FL_[0] + SL_ * (rLs - rL_), FL_[1] + SL_ * (rLs * SM - rL_ * uL_),
          FL_[2] + SL_ * (rLs * vLs - rL_ * wL_), FL_[3] + SL_ * (enLs - enL_), FL_[4] + SL_ * (BLs - CL_)]
    addL = ~(SLs >= 0.0)
    fl[2] = torch.where(addL, fl[2] + SLs * rLs * (vss - vLs), fl[2])
    fl[3] = torch.where(addL, fl[3] + SLs * (enLss - enLs), fl[3])
    fl[4] = torch.where(addL, fl[4] + SLs * (Bss - BLs), fl[4])
    fr = [FR_[0] + SR_ * (rRs - rR_), FR_[1] + SR_ * (rRs * SM - rR_ * uR_),
          FR_[2] + SR_ * (rRs * vRs - rR_ * wR_), FR_[3] + SR_ * (enRs - enR_), FR_[4] + SR_ * (BRs - CR_)]
    addR = ~(SRs <= 0.0)
    fr[2] = torch.where(addR, fr[2] + SRs * rRs * (vss - vRs), fr[2])
    fr[3] = torch.where(addR, fr[3] + SRs * (enRss - enRs), fr[3])
    fr[4] = torch.where(addR, fr[4] + SRs * (Bss - BRs), fr[4])
    fs = [torch.where(left, fl[c], fr[c]) for c in range(5)]

    Smax = dmax(SL_.abs(), SR_.abs())
    bnd = [dmax(FL_[0].abs(), FR_[0].abs()) + Smax * (rL_ + rR_),
           dmax(FL_[1].abs(), FR_[1].abs()) + Smax * ((rL_ * uL_).abs() + (rR_ * uR_).abs()),
           dmax(FL_[2].abs(), FR_[2].abs()) + Smax * ((rL_ * wL_).abs() + (rR_ * wR_).abs()),
           dmax(FL_[3].abs(), FR_[3].abs()) + Smax * (enL_.abs() + enR_.abs()),
           dmax(FL_[4].abs(), FR_[4].abs()) + Smax * (CL_.abs() + CR_.abs())]
    with torch.no_grad():
        bad = torch.zeros_like(star)
        for c in range(5):
            bad |= ~torch.isfinite(fs[c]) | (fs[c].abs() > 8.0 * bnd[c])
    hll_lane = rest & (catch2 | bad)
    H = _hll(FL_, FR_, SL_, SR_, (rL_, uL_, wL_, pTL_, BL_, CL_), (rR_, uR_, wR_, pTR_, BR_, CR_),
             enL_, enR_, False)
    pick_FL = catch1 & (cfL <= cfR)
    out = []
    for c in range(5):
        v = torch.where(hll_lane, H[c], fs[c])
        v = torch.where(edgeR, FR[c], v)
        v = torch.where(edgeL, FL[c], v)
        v = torch.where(catch1, torch.where(pick_FL, FL[c], FR[c]), v)
        out.append(v)
    return out


SOLVERS = {1: rusanov, 2: hlle, 3: hlld}


# ---- stage ------------------------------------------------------------------------

def stage(U, U0, k, dt, G: Geometry):
    """One stage k = 1, 2, 3 of src/main.f90:86-160 on state U (list of 6 (ny, nx)
    tensors: Mass, Momx, Momy, Energy, b_x, b_y); U0 is the step's input."""
    nx, ny, dx, g = G.nx, G.ny, G.dx, G.gamma
    vol = dx * dx
    imx, ipx = _nbr(nx, G.px, -1), _nbr(nx, G.px, 1)
    imy, ipy = _nbr(ny, G.py, -1), _nbr(ny, G.py, 1)
    M, Mx, My, E, bx, by = U
    Bx = 0.5 * (bx + _sx(bx, imx))
    By = 0.5 * (by + _sy(by, imy))
    r, u, w, P = get_primitive(M, Mx, My, E, Bx, By, g, vol)
    prims = [r, u, w, P, Bx, By]

    inv_dx, inv_2dx, hdx = 1.0 / dx, 1.0 / (2.0 * dx), 0.5 * dx
    faces = []   # per field: XL, XR, YL, YR
    for f, a in enumerate(prims):
        if G.use_lim:
            sx = limited_slope(_sx(a, imx), a, _sx(a, ipx), inv_dx, G.limiter)
            sy = limited_slope(_sy(a, imy), a, _sy(a, ipy), inv_dx, G.limiter)
        else:
            sx = (_sx(a, ipx) - _sx(a, imx)) * inv_2dx
            sy = (_sy(a, ipy) - _sy(a, imy)) * inv_2dx
        XL = a + hdx * sx
        YL = a + hdx * sy
        tx = a - hdx * sx
        ty = a - hdx * sy
        if G.px:
            XR = _sx(tx, ipx)
        else:
            XR = torch.cat([tx[:, 1:], a[:, -1:]], dim=1)
        if G.py:
            YR = _sy(ty, ipy)
        else:
            YR = torch.cat([ty[1:, :], a[-1:, :]], dim=0)
        rp = f in (0, 3)
        F4 = [XL, XR, YL, YR]
        if G.mood:
            with torch.no_grad():
                c = a
                nb = [_sx(a, ipx), _sx(a, imx), _sy(a, ipy), _sy(a, imy)]
                lo, hi = c, c
                for q in nb:
                    lo = torch.where(q < lo, q, lo)
                    hi = torch.where(q > hi, q, hi)
                if rp:
                    lo = torch.where(lo > 1e-9, lo, torch.full_like(lo, 1e-9))
            F4 = [torch.where((v < lo) | (v > hi), a, v) for v in F4]
        if rp:
            F4 = [_where(v < 1.0e-8, 1.0e-8, v) for v in F4]
        faces.append(F4)
    for s in range(4):
        p_th = faces[3][s] - 0.5 * (faces[4][s] * faces[4][s] + faces[5][s] * faces[5][s])
        low = p_th < P_FLOOR
        for f in range(6):
            faces[f][s] = torch.where(low, prims[f], faces[f][s])

    solver = SOLVERS[G.solver]
    fX = solver([faces[q][0] for q in (0, 1, 2, 3, 4, 5)], [faces[q][1] for q in (0, 1, 2, 3, 4, 5)], g)
    fY = solver([faces[q][2] for q in (0, 2, 1, 3, 5, 4)], [faces[q][3] for q in (0, 2, 1, 3, 5, 4)], g)
    flx = {0: (fX[0], fY[0]), 1: (fX[1], fY[2]), 2: (fX[2], fY[1]), 3: (fX[3], fY[3])}
    fByX, fBxY = fX[4], fY[4]

    dtdx = dt * dx
    V = []
    for q, Fq in enumerate((M, Mx, My, E)):
        fx, fy = flx[q]
        lx = _sx(fx, imx)
        v = Fq - dtdx * fx + dtdx * lx
        if not G.px:
            v = torch.cat([(Fq[:, :1] - dtdx * fx[:, :1]), v[:, 1:]], dim=1)
        ly = _sy(fy, imy)
        v2 = v - dtdx * fy + dtdx * ly
        if not G.py:
            v2 = torch.cat([(v[:1, :] - dtdx * fy[:1, :]), v2[1:, :]], dim=0)
        V.append(v2)
    V = list(conserved_floors(*V, Bx, By, vol, g))

    # CT (src/mhd_flux.f90:79-136) with the curl of -Ez (src/mhd_field_ops.f90:15-57)
    up = _sy(fByX, ipy)
    right = _sx(fBxY, ipx)
    negEz = -(0.25 * (-fByX - up + fBxY + right))
    dbx = (negEz - _sy(negEz, imy)) * inv_dx
    if not G.py:
        dbx = torch.cat([dbx[1:2, :], dbx[1:, :]], dim=0)
    dby = -((negEz - _sx(negEz, imx)) * inv_dx)
    if not G.px:
        dby = torch.cat([dby[:, 1:2], dby[:, 1:]], dim=1)
    bxn = bx + dt * dbx
    byn = by + dt * dby

    # outflow correction on the lo column / row (src/main.f90:116-148)
    if not G.px or not G.py:
        h = 0.5 * (Bx * Bx + By * By)
        en = (P - h) / (g - 1.0) + 0.5 * r * (u * u + w * w) + h
        if not G.px:
            add = [dtdx * r * u, dtdx * (r * (u * u) + P - Bx * Bx), dtdx * (r * u * w - Bx * By),
                   dtdx * ((en + P) * u - Bx * (Bx * u + By * w))]
            V = [torch.cat([V[q][:, :1] + add[q][:, :1], V[q][:, 1:]], dim=1) for q in range(4)]
        if not G.py:
            add = [dtdx * r * w, dtdx * (r * w * u - By * Bx), dtdx * (r * (w * w) + P - By * By),
                   dtdx * ((en + P) * w - By * (By * w + Bx * u))]
            V = [torch.cat([V[q][:1, :] + add[q][:1, :], V[q][1:, :]], dim=0) for q in range(4)]
    out = V + [bxn, byn]
    if k > 1:
        c0 = 0.75 if k == 2 else 1.0 / 3.0
        c1 = 0.25 if k == 2 else 2.0 / 3.0
        out = [c0 * U0[q] + c1 * out[q] for q in range(6)]
        out[:4] = list(conserved_floors(*out[:4], Bx, By, vol, g))
    return out


def step(U, dt, G: Geometry):
    """One SSP-RK3 step with prescribed dt; U is a (6, ny, nx) float64 tensor."""
    U0 = [U[q] for q in range(6)]
    X = stage(U0, U0, 1, dt, G)
    X = stage(X, U0, 2, dt, G)
    X = stage(X, U0, 3, dt, G)
    return torch.stack(X)
