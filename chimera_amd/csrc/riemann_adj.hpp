// Reverse-mode (adjoint) Rusanov and HLLE fluxes, written by hand from the forward
// solvers of mhd_phys.hpp (src/mhd_flux.f90:197-405): one backward sweep per face
// instead of one forward-mode pass per input.  Each branch (pressure floor, wave-speed
// cap, min/max of the signal speeds, upwind / HLL region, denominator floor) is the
// one the forward pass took, with ties resolved as in the FP64 solvers (fmax / fmin
// keep their first argument).
//
// Inputs and outputs are ordered (rho, u, w, P, Bn, Bt); lamF is the adjoint of
// (mass, mom_n, mom_t, energy, bt); lamL / lamR are written (not accumulated).
#pragma once

#include "mhd_phys.hpp"
#include "riemann_dual.hpp"

namespace chimera {

struct SideRev {
    double halfB2, pth, en;
    bool pth_floor;
};

__device__ __forceinline__ SideRev side_energy(const FaceState &s, double g) {
    SideRev r;
    r.halfB2 = 0.5 * (s.Bn * s.Bn + s.Bt * s.Bt);
    double a = s.P - r.halfB2;
    r.pth_floor = P_FLOOR > a;
    r.pth = r.pth_floor ? P_FLOOR : a;
    r.en = r.pth / (g - 1.0) + 0.5 * s.rho * (s.u * s.u + s.w * s.w) + r.halfB2;
    return r;
}

// lam of (en, pth, halfB2) of one side -> lam of its six inputs (accumulated into l).
__device__ __forceinline__ void side_energy_rev(const FaceState &s, const SideRev &r, double g,
                                                double len, double lpth, double lhalf,
                                                double l[6]) {
    lpth += len * (1.0 / (g - 1.0));
    l[0] += len * (0.5 * (s.u * s.u + s.w * s.w));
    l[1] += len * (s.rho * s.u);
    l[2] += len * (s.rho * s.w);
    lhalf += len;
    if (!r.pth_floor) {
        l[3] += lpth;
        lhalf -= lpth;
    }
    l[4] += lhalf * s.Bn;
    l[5] += lhalf * s.Bt;
}

// lam of the physical flux (the HLLE form: (rho*u)*u) of one side -> lam of its inputs
// and of its energy.
__device__ __forceinline__ void phys_flux_rev(const FaceState &s, double en, const double lF[5],
                                              double l[6], double &len) {
    const double r = s.rho, u = s.u, w = s.w, Bn = s.Bn, Bt = s.Bt;
    // F0 = r*u
    l[0] += lF[0] * u;
    l[1] += lF[0] * r;
    // F1 = (r*u)*u + P - Bn*Bn
    l[0] += lF[1] * (u * u);
    l[1] += lF[1] * (2.0 * r * u);
    l[3] += lF[1];
    l[4] -= lF[1] * (2.0 * Bn);
    // F2 = (r*u)*w - Bn*Bt
    l[0] += lF[2] * (u * w);
    l[1] += lF[2] * (r * w);
    l[2] += lF[2] * (r * u);
    l[4] -= lF[2] * Bt;
    l[5] -= lF[2] * Bn;
    // F3 = (en + P)*u - Bn*(Bn*u + Bt*w)
    len += lF[3] * u;
    l[3] += lF[3] * u;
    l[1] += lF[3] * ((en + s.P) - Bn * Bn);
    l[4] -= lF[3] * ((Bn * u + Bt * w) + Bn * u);
    l[5] -= lF[3] * (Bn * w);
    l[2] -= lF[3] * (Bn * Bt);
    // F4 = Bt*u - Bn*w
    l[5] += lF[4] * u;
    l[1] += lF[4] * Bt;
    l[4] -= lF[4] * w;
    l[2] -= lF[4] * Bn;
}

__device__ __forceinline__ void hlle_vjp(const FaceState &L, const FaceState &R, double g,
                                         const double lamF[5], double lamL[6], double lamR[6]) {
    const double cf_max = 1.0e2;
    SideRev eL = side_energy(L, g), eR = side_energy(R, g);
    double xL = (g * eL.pth + 2.0 * eL.halfB2) / L.rho;
    double xR = (g * eR.pth + 2.0 * eR.halfB2) / R.rho;
    double sqL = sqrt(0.5 * (xL + fabs(xL)));
    double sqR = sqrt(0.5 * (xR + fabs(xR)));
    bool capL = cf_max < sqL, capR = cf_max < sqR;
    double cfL = capL ? cf_max : sqL;
    double cfR = capR ? cf_max : sqR;
    double aL = L.u - cfL, bR = R.u - cfR;
    bool sl_right = bR < aL;                  // SL = fmin(aL, bR)
    double SL = sl_right ? bR : aL;
    double cL = L.u + cfL, dR = R.u + cfR;
    bool sr_right = dR > cL;                  // SR = fmax(cL, dR)
    double SR = sr_right ? dR : cL;

    for (int q = 0; q < 6; ++q) {
        lamL[q] = 0.0;
        lamR[q] = 0.0;
    }
    double lenL = 0.0, lenR = 0.0, lpthL = 0.0, lpthR = 0.0, lhL = 0.0, lhR = 0.0;
    double lSL = 0.0, lSR = 0.0;
    if (SR <= 0.0) {
        phys_flux_rev(R, eR.en, lamF, lamR, lenR);
    } else if (SL >= 0.0) {
        phys_flux_rev(L, eL.en, lamF, lamL, lenL);
    } else {
        double FL[5] = {L.rho * L.u, L.rho * L.u * L.u + L.P - L.Bn * L.Bn,
                        L.rho * L.u * L.w - L.Bn * L.Bt,
                        (eL.en + L.P) * L.u - L.Bn * (L.Bn * L.u + L.Bt * L.w),
                        L.Bt * L.u - L.Bn * L.w};
        double FR[5] = {R.rho * R.u, R.rho * R.u * R.u + R.P - R.Bn * R.Bn,
                        R.rho * R.u * R.w - R.Bn * R.Bt,
                        (eR.en + R.P) * R.u - R.Bn * (R.Bn * R.u + R.Bt * R.w),
                        R.Bt * R.u - R.Bn * R.w};
        double dU[5] = {R.rho - L.rho, R.rho * R.u - L.rho * L.u, R.rho * R.w - L.rho * L.w,
                        eR.en - eL.en, R.Bt - L.Bt};
        double span = SR - SL;
        bool span_floor = 1.0e-12 > span;
        double inv = 1.0 / (span_floor ? 1.0e-12 : span);
        double sLR = SL * SR;
        double lFL[5], lFR[5], linv = 0.0;
        for (int c = 0; c < 5; ++c) {
            double num = SR * FL[c] - SL * FR[c] + sLR * dU[c];
            double lc = lamF[c];
            linv += lc * num;
            lFL[c] = lc * SR * inv;
            lFR[c] = -lc * SL * inv;
            double ld = lc * sLR * inv;
            lSR += lc * inv * (FL[c] + SL * dU[c]);
            lSL += lc * inv * (SR * dU[c] - FR[c]);
            // dU back
            if (c == 0) {
                lamR[0] += ld;
                lamL[0] -= ld;
            } else if (c == 1) {
                lamR[0] += ld * R.u;
                lamR[1] += ld * R.rho;
                lamL[0] -= ld * L.u;
                lamL[1] -= ld * L.rho;
            } else if (c == 2) {
                lamR[0] += ld * R.w;
                lamR[2] += ld * R.rho;
                lamL[0] -= ld * L.w;
                lamL[2] -= ld * L.rho;
            } else if (c == 3) {
                lenR += ld;
                lenL -= ld;
            } else {
                lamR[5] += ld;
                lamL[5] -= ld;
            }
        }
        if (!span_floor) {
            double lspan = -linv * inv * inv;
            lSR += lspan;
            lSL -= lspan;
        }
        phys_flux_rev(L, eL.en, lFL, lamL, lenL);
        phys_flux_rev(R, eR.en, lFR, lamR, lenR);
    }
    // SL = fmin(uL - cfL, uR - cfR), SR = fmax(uL + cfL, uR + cfR)
    double lcfL = 0.0, lcfR = 0.0;
    if (sl_right) {
        lamR[1] += lSL;
        lcfR -= lSL;
    } else {
        lamL[1] += lSL;
        lcfL -= lSL;
    }
    if (sr_right) {
        lamR[1] += lSR;
        lcfR += lSR;
    } else {
        lamL[1] += lSR;
        lcfL += lSR;
    }
    // cf = fmin(sqrt(0.5 (x + |x|)), cap); x = (g pth + 2 halfB2) / rho
    // (lam of x) / rho with one division per side
    if (!capL && xL > 0.0 && sqL > 0.0) {
        const double lxr = lcfL * (0.5 / (sqL * L.rho));
        lpthL += lxr * g;
        lhL += lxr * 2.0;
        lamL[0] -= lxr * xL;
    }
    if (!capR && xR > 0.0 && sqR > 0.0) {
        const double lxr = lcfR * (0.5 / (sqR * R.rho));
        lpthR += lxr * g;
        lhR += lxr * 2.0;
        lamR[0] -= lxr * xR;
    }
    side_energy_rev(L, eL, g, lenL, lpthL, lhL, lamL);
    side_energy_rev(R, eR, g, lenR, lpthR, lhR, lamR);
}

__device__ __forceinline__ void rusanov_vjp(const FaceState &L, const FaceState &R, double g,
                                            const double lamF[5], double lamL[6], double lamR[6]) {
    SideRev eL = side_energy(L, g), eR = side_energy(R, g);
    const double gm1 = g - 1.0;
    double rho_avg = 0.5 * (L.rho + R.rho);
    double inv = 1.0 / rho_avg;
    double mLn = L.rho * L.u, mRn = R.rho * R.u;
    double mLt = L.rho * L.w, mRt = R.rho * R.w;
    double mx = 0.5 * (mLn + mRn);
    double my = 0.5 * (mLt + mRt);
    double en = 0.5 * (eL.en + eR.en);
    double bx = 0.5 * (L.Bn + R.Bn);
    double by = 0.5 * (L.Bt + R.Bt);
    double hb = 0.5 * (bx * bx + by * by);
    double Pavg = gm1 * (en - 0.5 * (mx * mx + my * my) * inv - hb) + hb;
    double cL2 = (g * eL.pth + 2.0 * eL.halfB2) / L.rho;
    double cR2 = (g * eR.pth + 2.0 * eR.halfB2) / R.rho;
    double sqL = sqrt(0.5 * (cL2 + fabs(cL2)));
    double sqR = sqrt(0.5 * (cR2 + fabs(cR2)));
    double sL = sqL + fabs(L.u);
    double sR = sqR + fabs(R.u);
    bool c_right = sR > sL;                   // C = 0.5 fmax(sL, sR)
    double C = 0.5 * (c_right ? sR : sL);

    const double l0 = lamF[0], l1 = lamF[1], l2 = lamF[2], l3 = lamF[3], l4 = lamF[4];
    for (int q = 0; q < 6; ++q) {
        lamL[q] = 0.0;
        lamR[q] = 0.0;
    }
    double lmx = 0.0, lmy = 0.0, linv = 0.0, len = 0.0, lbx = 0.0, lby = 0.0, lP = 0.0;
    double lmLn = 0.0, lmRn = 0.0, lmLt = 0.0, lmRt = 0.0, lenL = 0.0, lenR = 0.0;
    // dissipation C * jumps
    double lC = -(l0 * (R.rho - L.rho) + l1 * (mRn - mLn) + l2 * (mRt - mLt) +
                  l3 * (eR.en - eL.en) + l4 * (R.Bt - L.Bt));
    lamR[0] -= C * l0;
    lamL[0] += C * l0;
    lmRn -= C * l1;
    lmLn += C * l1;
    lmRt -= C * l2;
    lmLt += C * l2;
    lenR -= C * l3;
    lenL += C * l3;
    lamR[5] -= C * l4;
    lamL[5] += C * l4;
    // mass = mx
    lmx += l0;
    // mom_n = mx*mx*inv + Pavg - bx*bx
    lmx += l1 * (2.0 * mx * inv);
    linv += l1 * (mx * mx);
    lP += l1;
    lbx -= l1 * (2.0 * bx);
    // mom_t = mx*my*inv - bx*by
    lmx += l2 * (my * inv);
    lmy += l2 * (mx * inv);
    linv += l2 * (mx * my);
    lbx -= l2 * by;
    lby -= l2 * bx;
    // energy = (en + Pavg)*mx*inv - bx*(bx*mx + by*my)*inv
    {
        double bm = bx * mx + by * my;
        len += l3 * (mx * inv);
        lP += l3 * (mx * inv);
        lmx += l3 * ((en + Pavg) * inv);
        linv += l3 * ((en + Pavg) * mx);
        lbx -= l3 * (bm * inv + bx * mx * inv);
        lmx -= l3 * (bx * bx * inv);
        lby -= l3 * (bx * my * inv);
        lmy -= l3 * (bx * by * inv);
        linv -= l3 * (bx * bm);
    }
    // bt = (by*mx - bx*my)*inv
    lby += l4 * (mx * inv);
    lmx += l4 * (by * inv);
    lbx -= l4 * (my * inv);
    lmy -= l4 * (bx * inv);
    linv += l4 * (by * mx - bx * my);
    // Pavg = gm1*(en - 0.5*(mx^2+my^2)*inv - hb) + hb
    len += lP * gm1;
    lmx -= lP * gm1 * (mx * inv);
    lmy -= lP * gm1 * (my * inv);
    linv -= lP * gm1 * (0.5 * (mx * mx + my * my));
    double lhb = lP * (1.0 - gm1);
    lbx += lhb * bx;
    lby += lhb * by;
    // C = 0.5 fmax(sL, sR); s = sqrt(0.5 (c2 + |c2|)) + |u|
    double lsL = c_right ? 0.0 : 0.5 * lC;
    double lsR = c_right ? 0.5 * lC : 0.0;
    double lpthL = 0.0, lpthR = 0.0, lhL = 0.0, lhR = 0.0;
    if (L.u != 0.0) lamL[1] += lsL * copysign(1.0, L.u);
    if (R.u != 0.0) lamR[1] += lsR * copysign(1.0, R.u);
    // (lam of c2) / rho with one division per side
    if (cL2 > 0.0 && sqL > 0.0) {
        const double lcr = lsL * (0.5 / (sqL * L.rho));
        lpthL += lcr * g;
        lhL += lcr * 2.0;
        lamL[0] -= lcr * cL2;
    }
    if (cR2 > 0.0 && sqR > 0.0) {
        const double lcr = lsR * (0.5 / (sqR * R.rho));
        lpthR += lcr * g;
        lhR += lcr * 2.0;
        lamR[0] -= lcr * cR2;
    }
    // averages
    double lra = -linv * inv * inv;
    lamL[0] += 0.5 * lra;
    lamR[0] += 0.5 * lra;
    lmLn += 0.5 * lmx;
    lmRn += 0.5 * lmx;
    lmLt += 0.5 * lmy;
    lmRt += 0.5 * lmy;
    lenL += 0.5 * len;
    lenR += 0.5 * len;
    lamL[4] += 0.5 * lbx;
    lamR[4] += 0.5 * lbx;
    lamL[5] += 0.5 * lby;
    lamR[5] += 0.5 * lby;
    // momenta
    lamL[0] += lmLn * L.u + lmLt * L.w;
    lamL[1] += lmLn * L.rho;
    lamL[2] += lmLt * L.rho;
    lamR[0] += lmRn * R.u + lmRt * R.w;
    lamR[1] += lmRn * R.rho;
    lamR[2] += lmRt * R.rho;
    side_energy_rev(L, eL, g, lenL, lpthL, lhL, lamL);
    side_energy_rev(R, eR, g, lenR, lpthR, lhR, lamR);
}

// ---- HLLD (src/mhd_flux.f90:490-777) ------------------------------------------------
//
// The forward recompute keeps every intermediate and branch of hlld() (mhd_phys.hpp);
// the reverse sweep then runs through the branch that produced the flux: c_f cap ->
// upwind side, supersonic upwind, fan-edge catch and the non-finite / out-of-bound
// catch-all -> HLL average, else the star / double-star states with the p_T* floor,
// the degenerate-denominator and |B_n| -> 0 branches.  Conventions match the
// dual-number solver: fmax / fmin take the argument the FP64 pass returned (the first
// on ties), sqrt and |x| have derivative 0 at 0, the SM denominator floor is a
// constant.

struct HlldSide {
    double r, u, w, pT, B, C;   // rho, u_n, u_t, total pressure, B_n, B_t
    SideRev e;                  // halfB2, thermal pressure (floor), energy
    double cf;
    __device__ __forceinline__ double B2() const { return B * B; }
    __device__ __forceinline__ double vdotB() const { return u * B + w * C; }
};

// the magnetosonic intermediates of c_f, recomputed with the forward's expressions
struct HlldCf {
    double ca, cs2, b2, tmp, s1;
    bool disc_floor;
};

__device__ __forceinline__ HlldCf hlld_cf_parts(const HlldSide &h, double g) {
    HlldCf c;
    const double B2 = h.B2();
    c.ca = B2 / h.r;
    c.cs2 = g * h.e.pth / h.r;
    c.b2 = (B2 + h.C * h.C) / h.r;
    c.tmp = c.cs2 + c.b2;
    const double d = c.tmp * c.tmp - 4.0 * c.cs2 * c.ca;
    c.disc_floor = 0.0 > d;
    c.s1 = sqrt(c.disc_floor ? 0.0 : d);
    return c;
}

__device__ __forceinline__ HlldSide hlld_side(const FaceState &s, double g) {
    HlldSide h;
    h.r = s.rho;
    h.u = s.u;
    h.w = s.w;
    h.pT = s.P;
    h.B = s.Bn;
    h.C = s.Bt;
    h.e = side_energy(s, g);
    const HlldCf c = hlld_cf_parts(h, g);
    h.cf = sqrt(0.5 * (c.tmp + c.s1));
    return h;
}

// physical flux of one side in the HLLD form
__device__ __forceinline__ void hlld_phys(const HlldSide &h, double F[5]) {
    F[0] = h.r * h.u;
    F[1] = h.r * (h.u * h.u) + h.pT - h.B * h.B;
    F[2] = h.r * h.u * h.w - h.B * h.C;
    F[3] = (h.e.en + h.pT) * h.u - h.B * h.vdotB();
    F[4] = h.C * h.u - h.B * h.w;
}

// adjoints gathered per side before they reach its six inputs
struct HlldAdj {
    double l[6];
    double len, lpth, lhalf, lB2, lvdotB;
    double cF;   // lam of the side's physical flux = cF * lamF
};

__device__ __forceinline__ void hlld_adj_zero(HlldAdj &a) {
#pragma unroll
    for (int q = 0; q < 6; ++q) a.l[q] = 0.0;
    a.cF = 0.0;
    a.len = a.lpth = a.lhalf = a.lB2 = a.lvdotB = 0.0;
}

// c_f = sqrt(0.5 (tmp + sqrt(max(tmp^2 - 4 cs2 ca, 0)))) and its inputs
__device__ __forceinline__ void hlld_cf_rev(const HlldSide &h, double g, double lcf, HlldAdj &a) {
    const HlldCf c = hlld_cf_parts(h, g);
    const double lx = h.cf > 0.0 ? lcf * (0.5 / h.cf) : 0.0;
    double ltmp = 0.5 * lx;
    const double ls1 = 0.5 * lx;
    const double ld = (!c.disc_floor && c.s1 > 0.0) ? ls1 * (0.5 / c.s1) : 0.0;
    ltmp += ld * (2.0 * c.tmp);
    double lcs2 = -ld * (4.0 * c.ca);
    const double lca = -ld * (4.0 * c.cs2);
    lcs2 += ltmp;
    const double lb2 = ltmp;
    const double ir = 1.0 / h.r;   // one reciprocal for the seven quotients below
    a.lB2 += lb2 * ir;
    a.l[5] += lb2 * (2.0 * h.C) * ir;
    a.l[0] -= lb2 * c.b2 * ir;
    a.lpth += lcs2 * g * ir;
    a.l[0] -= lcs2 * c.cs2 * ir;
    a.lB2 += lca * ir;
    a.l[0] -= lca * c.ca * ir;
}

// phys flux, v.B, B^2 and energy of one side -> its inputs (written into lam)
__device__ __forceinline__ void hlld_side_finish(const FaceState &s, const HlldSide &h, double g,
                                                 const double lamF[5], HlldAdj &a, double lam[6]) {
    if (a.cF != 0.0) {
        double lF[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) lF[c] = a.cF * lamF[c];
        phys_flux_rev(s, h.e.en, lF, a.l, a.len);
    }
    a.l[1] += a.lvdotB * h.B;
    a.l[4] += a.lvdotB * h.u;
    a.l[2] += a.lvdotB * h.C;
    a.l[5] += a.lvdotB * h.w;
    a.l[4] += a.lB2 * (2.0 * h.B);
    side_energy_rev(s, h.e, g, a.len, a.lpth, a.lhalf, a.l);
#pragma unroll
    for (int q = 0; q < 6; ++q) lam[q] = a.l[q];
}

// one side's star state (Miyoshi-Kusano): the forward values the rest of the face
// needs; the cheap intermediates are rebuilt by hlld_star_parts
struct HlldStar {
    double rs, vs, Bs, vdotBs, ens, sq;
    bool sq_floor;
};

struct HlldStarParts {
    double a, dS, k, den, t, nv;
    bool cond;
};

__device__ __forceinline__ HlldStarParts hlld_star_parts(const HlldSide &h, double S, double SM) {
    HlldStarParts p;
    p.a = h.r * (S - h.u);
    p.dS = S - SM;
    p.k = p.a * p.dS;
    p.den = p.k - h.B2();
    p.cond = fabs(p.den) > 1.0e-3 * fmax(fmax(fabs(p.k), h.B2()), 1.0e-30);
    p.t = S - h.u;
    p.nv = p.cond ? h.B * h.C * (SM - h.u) : 0.0;
    return p;
}

__device__ __forceinline__ HlldStar hlld_star(const HlldSide &h, double S, double SM, double pTs) {
    const HlldStarParts p = hlld_star_parts(h, S, SM);
    HlldStar st;
    st.rs = p.a / p.dS;
    st.vs = h.w;
    st.Bs = h.C;
    if (p.cond) {
        st.vs = h.w - p.nv / p.den;
        st.Bs = h.C * (h.r * (p.t * p.t) - h.B2()) / p.den;
    }
    st.vdotBs = SM * h.B + st.vs * st.Bs;
    st.ens = (h.e.en * (S - h.u) - h.pT * h.u + pTs * SM + h.B * (h.vdotB() - st.vdotBs)) / p.dS;
    st.sq_floor = RHO_FLOOR > st.rs;
    st.sq = sqrt(st.sq_floor ? RHO_FLOOR : st.rs);
    return st;
}

// reverse of hlld_star given the adjoints of rs, vs, Bs, vdotBs, ens and sq; adds to the
// side's adjoints and to those of S, SM, p_T* and of a (returned through la)
__device__ __forceinline__ void hlld_star_rev(const HlldSide &h, const HlldStar &st, double S,
                                              double SM, double pTs, double lrs, double lvs,
                                              double lBs, double lvdotBs, double lens, double lsq,
                                              HlldAdj &a, double &lS, double &lSM, double &lpTs,
                                              double &la) {
    const HlldStarParts p = hlld_star_parts(h, S, SM);
    if (!st.sq_floor && st.sq > 0.0) lrs += lsq * (0.5 / st.sq);
    double ldS = 0.0;
    const double idS = 1.0 / p.dS;   // reverse only: one reciprocal per denominator
    // ens = numE / dS
    const double lnumE = lens * idS;
    ldS -= lens * st.ens * idS;
    a.len += lnumE * (S - h.u);
    lS += lnumE * h.e.en;
    a.l[1] -= lnumE * h.e.en;
    a.l[3] -= lnumE * h.u;
    a.l[1] -= lnumE * h.pT;
    lpTs += lnumE * SM;
    lSM += lnumE * pTs;
    a.l[4] += lnumE * (h.vdotB() - st.vdotBs);
    a.lvdotB += lnumE * h.B;
    lvdotBs -= lnumE * h.B;
    // vdotBs = SM B + vs Bs
    lSM += lvdotBs * h.B;
    a.l[4] += lvdotBs * SM;
    lvs += lvdotBs * st.Bs;
    lBs += lvdotBs * st.vs;
    double lk = 0.0;
    if (p.cond) {
        double lden = 0.0;
        const double iden = 1.0 / p.den;
        // Bs = C (r t^2 - B^2) / den
        const double lnb = lBs * iden;
        lden -= lBs * st.Bs * iden;
        a.l[5] += lnb * (h.r * (p.t * p.t) - h.B2());
        const double lin = lnb * h.C;
        a.l[0] += lin * (p.t * p.t);
        const double lt = lin * h.r * (2.0 * p.t);
        a.lB2 -= lin;
        lS += lt;
        a.l[1] -= lt;
        // vs = w - nv / den, nv = B C (SM - u)
        a.l[2] += lvs;
        const double lnv = -lvs * iden;
        lden += lvs * (p.nv * iden) * iden;
        a.l[4] += lnv * h.C * (SM - h.u);
        a.l[5] += lnv * h.B * (SM - h.u);
        lSM += lnv * (h.B * h.C);
        a.l[1] -= lnv * (h.B * h.C);
        // den = k - B^2
        lk += lden;
        a.lB2 -= lden;
    } else {
        a.l[2] += lvs;
        a.l[5] += lBs;
    }
    // k = a dS, rs = a / dS, dS = S - SM
    la += lk * p.dS;
    ldS += lk * p.a;
    la += lrs * idS;
    ldS -= lrs * st.rs * idS;
    lS += ldS;
    lSM -= ldS;
}

// HLL average of the two physical fluxes (no span floor) -> adjoints of FL, FR, SL,
// SR and of the conserved jumps
__device__ __forceinline__ void hll_avg_rev(const HlldSide &hL, const HlldSide &hR,
                                            const double FL[5], const double FR[5], double SL,
                                            double SR, const double lamF[5], HlldAdj &aL,
                                            HlldAdj &aR, double &lSL, double &lSR) {
    const double dU[5] = {hR.r - hL.r, hR.r * hR.u - hL.r * hL.u, hR.r * hR.w - hL.r * hL.w,
                          hR.e.en - hL.e.en, hR.C - hL.C};
    const double inv = 1.0 / (SR - SL);
    const double sLR = SL * SR;
    double linv = 0.0;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const double num = SR * FL[c] - SL * FR[c] + sLR * dU[c];
        const double lc = lamF[c];
        linv += lc * num;
        const double ld = lc * sLR * inv;
        lSR += lc * inv * (FL[c] + SL * dU[c]);
        lSL += lc * inv * (SR * dU[c] - FR[c]);
        if (c == 0) {
            aR.l[0] += ld;
            aL.l[0] -= ld;
        } else if (c == 1) {
            aR.l[0] += ld * hR.u;
            aR.l[1] += ld * hR.r;
            aL.l[0] -= ld * hL.u;
            aL.l[1] -= ld * hL.r;
        } else if (c == 2) {
            aR.l[0] += ld * hR.w;
            aR.l[2] += ld * hR.r;
            aL.l[0] -= ld * hL.w;
            aL.l[2] -= ld * hL.r;
        } else if (c == 3) {
            aR.len += ld;
            aL.len -= ld;
        } else {
            aR.l[5] += ld;
            aL.l[5] -= ld;
        }
    }
    aL.cF += SR * inv;
    aR.cF -= SL * inv;
    const double lspan = -linv * inv * inv;
    lSR += lspan;
    lSL -= lspan;
}

__device__ __forceinline__ void hlld_vjp(const FaceState &L, const FaceState &R, double g,
                                         const double lamF[5], double lamL[6], double lamR[6]) {
    const double EPS_FAN = 1.0e-6, cf_max = 50.0;
    const HlldSide hL = hlld_side(L, g), hR = hlld_side(R, g);
    HlldAdj aL, aR;
    hlld_adj_zero(aL);
    hlld_adj_zero(aR);

    const bool capL = cf_max < hL.cf, capR = cf_max < hR.cf;
    const double cfL = capL ? cf_max : hL.cf, cfR = capR ? cf_max : hR.cf;
    const double xa = hL.u - cfL, xb = hR.u - cfR;
    const bool sl_right = xb < xa;
    const double SL = sl_right ? xb : xa;
    const double ya = hL.u + cfL, yb = hR.u + cfR;
    const bool sr_right = yb > ya;
    const double SR = sr_right ? yb : ya;

    double lSL = 0.0, lSR = 0.0;
    bool speeds = true;   // SL / SR (hence c_f) enter the flux
    if (cfL >= cf_max || cfR >= cf_max) {
        HlldAdj &a = (cfL <= cfR) ? aL : aR;
        a.cF += 1.0;
        speeds = false;
    } else if (SL >= 0.0) {
        aL.cF += 1.0;
        speeds = false;
    } else if (SR <= 0.0) {
        aR.cF += 1.0;
        speeds = false;
    } else {
        const double aLv = hL.r * (SL - hL.u), aRv = hR.r * (SR - hR.u);
        double denMD = aRv - aLv;
        const bool md_floor = fabs(denMD) < 1.0e-15;
        if (md_floor) denMD = copysign(1.0e-15, denMD);
        const double num = hR.r * hR.u * (SR - hR.u) - hL.r * hL.u * (SL - hL.u) - (hR.pT - hL.pT);
        const double SM = num / denMD;
        const double pTs0 = hL.pT + aLv * (SM - hL.u);
        bool hll = SM <= SL + EPS_FAN * (SR - SL) || SM >= SR - EPS_FAN * (SR - SL);
        if (!hll) {
            const bool pts_floor = P_FLOOR > pTs0;
            const double pTs = pts_floor ? P_FLOOR : pTs0;
            const HlldStar sL = hlld_star(hL, SL, SM, pTs), sR = hlld_star(hR, SR, SM, pTs);
            const double Bavg = 0.5 * (hL.B + hR.B);
            const double absB = fabs(Bavg);
            const bool big = absB > 1.0e-10;
            double SLs = SM - absB / sL.sq, SRs = SM + absB / sR.sq;
            double vss, Bss, enLss, enRss, vdotBss = 0.0, sgn = 1.0, dss = 1.0;
            if (big) {
                sgn = copysign(1.0, Bavg);
                dss = sL.sq + sR.sq;
                vss = (sL.sq * sL.vs + sR.sq * sR.vs + sgn * (sR.Bs - sL.Bs)) / dss;
                Bss = (sL.sq * sR.Bs + sR.sq * sL.Bs + sgn * sL.sq * sR.sq * (sR.vs - sL.vs)) / dss;
                vdotBss = SM * Bavg + vss * Bss;
                enLss = sL.ens - sL.sq * sgn * (sL.vdotBs - vdotBss);
                enRss = sR.ens + sR.sq * sgn * (sR.vdotBs - vdotBss);
            } else {
                vss = 0.5 * (sL.vs + sR.vs);
                Bss = 0.5 * (sL.Bs + sR.Bs);
                enLss = sL.ens;
                enRss = sR.ens;
                SLs = SM;
                SRs = SM;
            }
            // the forward flux, for the catch-all check
            const bool left = SM >= 0.0;
            const HlldSide &hX = left ? hL : hR;
            const HlldStar &sX = left ? sL : sR;
            const double S = left ? SL : SR, Ss = left ? SLs : SRs;
            const double enXss = left ? enLss : enRss;
            double FX[5];
            hlld_phys(left ? hL : hR, FX);
            const bool star2 = left ? !(SLs >= 0.0) : !(SRs <= 0.0);
            double f[5];
            f[0] = FX[0] + S * (sX.rs - hX.r);
            f[1] = FX[1] + S * (sX.rs * SM - hX.r * hX.u);
            f[2] = FX[2] + S * (sX.rs * sX.vs - hX.r * hX.w);
            f[3] = FX[3] + S * (sX.ens - hX.e.en);
            f[4] = FX[4] + S * (sX.Bs - hX.C);
            if (star2) {
                f[2] = f[2] + Ss * sX.rs * (vss - sX.vs);
                f[3] = f[3] + Ss * (enXss - sX.ens);
                f[4] = f[4] + Ss * (Bss - sX.Bs);
            }
            const double Smax = fmax(fabs(SL), fabs(SR));
            double FL[5], FR[5];
            hlld_phys(hL, FL);
            hlld_phys(hR, FR);
            const bool ok =
                finite_in(f[0], fmax(fabs(FL[0]), fabs(FR[0])) + Smax * (hL.r + hR.r)) &&
                finite_in(f[1], fmax(fabs(FL[1]), fabs(FR[1])) +
                                    Smax * (fabs(hL.r * hL.u) + fabs(hR.r * hR.u))) &&
                finite_in(f[2], fmax(fabs(FL[2]), fabs(FR[2])) +
                                    Smax * (fabs(hL.r * hL.w) + fabs(hR.r * hR.w))) &&
                finite_in(f[3], fmax(fabs(FL[3]), fabs(FR[3])) +
                                    Smax * (fabs(hL.e.en) + fabs(hR.e.en))) &&
                finite_in(f[4], fmax(fabs(FL[4]), fabs(FR[4])) + Smax * (fabs(hL.C) + fabs(hR.C)));
            if (!ok) {
                hll = true;
            } else {
                HlldAdj &aX = left ? aL : aR;
                const double l0 = lamF[0], l1 = lamF[1], l2 = lamF[2], l3 = lamF[3], l4 = lamF[4];
                double lS = 0.0, lSM = 0.0, lrs = 0.0, lvs = 0.0, lBs = 0.0, lens = 0.0;
                double lSs = 0.0, lvss = 0.0, lBss = 0.0, lenss = 0.0;
                aX.cF += 1.0;
                lS += l0 * (sX.rs - hX.r);
                lrs += l0 * S;
                aX.l[0] -= l0 * S;
                lS += l1 * (sX.rs * SM - hX.r * hX.u);
                lrs += l1 * S * SM;
                lSM += l1 * S * sX.rs;
                aX.l[0] -= l1 * S * hX.u;
                aX.l[1] -= l1 * S * hX.r;
                lS += l2 * (sX.rs * sX.vs - hX.r * hX.w);
                lrs += l2 * S * sX.vs;
                lvs += l2 * S * sX.rs;
                aX.l[0] -= l2 * S * hX.w;
                aX.l[2] -= l2 * S * hX.r;
                lS += l3 * (sX.ens - hX.e.en);
                lens += l3 * S;
                aX.len -= l3 * S;
                lS += l4 * (sX.Bs - hX.C);
                lBs += l4 * S;
                aX.l[5] -= l4 * S;
                if (star2) {
                    lSs += l2 * sX.rs * (vss - sX.vs);
                    lrs += l2 * Ss * (vss - sX.vs);
                    lvss += l2 * (Ss * sX.rs);
                    lvs -= l2 * (Ss * sX.rs);
                    lSs += l3 * (enXss - sX.ens);
                    lenss += l3 * Ss;
                    lens -= l3 * Ss;
                    lSs += l4 * (Bss - sX.Bs);
                    lBss += l4 * Ss;
                    lBs -= l4 * Ss;
                }
                if (left) lSL += lS;
                else lSR += lS;
                // per-side star adjoints
                double lrsL = left ? lrs : 0.0, lrsR = left ? 0.0 : lrs;
                double lvsL = left ? lvs : 0.0, lvsR = left ? 0.0 : lvs;
                double lBsL = left ? lBs : 0.0, lBsR = left ? 0.0 : lBs;
                double lensL = left ? lens : 0.0, lensR = left ? 0.0 : lens;
                double lvdotBsL = 0.0, lvdotBsR = 0.0, lsqL = 0.0, lsqR = 0.0;
                double lSLs = left ? lSs : 0.0, lSRs = left ? 0.0 : lSs;
                double lBavg = 0.0;
                if (big) {
                    double lvdotBss = 0.0;
                    if (left) {   // enLss = enLs - sqL sgn (vdotBLs - vdotBss)
                        lensL += lenss;
                        lsqL -= lenss * sgn * (sL.vdotBs - vdotBss);
                        lvdotBsL -= lenss * (sL.sq * sgn);
                        lvdotBss += lenss * (sL.sq * sgn);
                    } else {      // enRss = enRs + sqR sgn (vdotBRs - vdotBss)
                        lensR += lenss;
                        lsqR += lenss * sgn * (sR.vdotBs - vdotBss);
                        lvdotBsR += lenss * (sR.sq * sgn);
                        lvdotBss -= lenss * (sR.sq * sgn);
                    }
                    lSM += lvdotBss * Bavg;
                    lBavg += lvdotBss * SM;
                    lvss += lvdotBss * Bss;
                    lBss += lvdotBss * vss;
                    double ldss = 0.0;
                    const double idss = 1.0 / dss;   // reverse only
                    const double lnB = lBss * idss;
                    ldss -= lBss * Bss * idss;
                    const double dv = sR.vs - sL.vs;
                    lsqL += lnB * sR.Bs;
                    lBsR += lnB * sL.sq;
                    lsqR += lnB * sL.Bs;
                    lBsL += lnB * sR.sq;
                    lsqL += lnB * sgn * sR.sq * dv;
                    lsqR += lnB * sgn * sL.sq * dv;
                    lvsR += lnB * (sgn * sL.sq * sR.sq);
                    lvsL -= lnB * (sgn * sL.sq * sR.sq);
                    const double lnv = lvss * idss;
                    ldss -= lvss * vss * idss;
                    lsqL += lnv * sL.vs;
                    lvsL += lnv * sL.sq;
                    lsqR += lnv * sR.vs;
                    lvsR += lnv * sR.sq;
                    lBsR += lnv * sgn;
                    lBsL -= lnv * sgn;
                    lsqL += ldss;
                    lsqR += ldss;
                    // SLs = SM - |Bavg| / sqL, SRs = SM + |Bavg| / sqR
                    double labs = 0.0;
                    const double isqL = 1.0 / sL.sq, isqR = 1.0 / sR.sq;
                    lSM += lSLs;
                    labs -= lSLs * isqL;
                    lsqL += lSLs * (absB * isqL) * isqL;
                    lSM += lSRs;
                    labs += lSRs * isqR;
                    lsqR -= lSRs * (absB * isqR) * isqR;
                    lBavg += labs * sgn;
                } else {
                    lvsL += 0.5 * lvss;
                    lvsR += 0.5 * lvss;
                    lBsL += 0.5 * lBss;
                    lBsR += 0.5 * lBss;
                    if (left) lensL += lenss;
                    else lensR += lenss;
                    lSM += lSLs;
                    lSM += lSRs;
                }
                aL.l[4] += 0.5 * lBavg;
                aR.l[4] += 0.5 * lBavg;
                double lpTs = 0.0, laL = 0.0, laR = 0.0;
                hlld_star_rev(hL, sL, SL, SM, pTs, lrsL, lvsL, lBsL, lvdotBsL, lensL, lsqL, aL,
                              lSL, lSM, lpTs, laL);
                hlld_star_rev(hR, sR, SR, SM, pTs, lrsR, lvsR, lBsR, lvdotBsR, lensR, lsqR, aR,
                              lSR, lSM, lpTs, laR);
                // pTs = max(pTL + aL (SM - uL), P_floor)
                const double lpTs0 = pts_floor ? 0.0 : lpTs;
                aL.l[3] += lpTs0;
                laL += lpTs0 * (SM - hL.u);
                lSM += lpTs0 * aLv;
                aL.l[1] -= lpTs0 * aLv;
                // SM = num / denMD
                const double lnum = lSM / denMD;
                const double ldenMD = md_floor ? 0.0 : -lSM * SM / denMD;
                aR.l[0] += lnum * hR.u * (SR - hR.u);
                aR.l[1] += lnum * hR.r * (SR - hR.u);
                lSR += lnum * (hR.r * hR.u);
                aR.l[1] -= lnum * (hR.r * hR.u);
                aL.l[0] -= lnum * hL.u * (SL - hL.u);
                aL.l[1] -= lnum * hL.r * (SL - hL.u);
                lSL -= lnum * (hL.r * hL.u);
                aL.l[1] += lnum * (hL.r * hL.u);
                aR.l[3] -= lnum;
                aL.l[3] += lnum;
                laR += ldenMD;
                laL -= ldenMD;
                // a = r (S - u)
                aL.l[0] += laL * (SL - hL.u);
                lSL += laL * hL.r;
                aL.l[1] -= laL * hL.r;
                aR.l[0] += laR * (SR - hR.u);
                lSR += laR * hR.r;
                aR.l[1] -= laR * hR.r;
            }
        }
        if (hll) {
            double FL[5], FR[5];
            hlld_phys(hL, FL);
            hlld_phys(hR, FR);
            hll_avg_rev(hL, hR, FL, FR, SL, SR, lamF, aL, aR, lSL, lSR);
        }
    }
    if (speeds) {
        // SL = min(uL - cfL, uR - cfR), SR = max(uL + cfL, uR + cfR); no cap on this path
        double lcfL = 0.0, lcfR = 0.0;
        if (sl_right) {
            aR.l[1] += lSL;
            lcfR -= lSL;
        } else {
            aL.l[1] += lSL;
            lcfL -= lSL;
        }
        if (sr_right) {
            aR.l[1] += lSR;
            lcfR += lSR;
        } else {
            aL.l[1] += lSR;
            lcfL += lSR;
        }
        if (!capL) hlld_cf_rev(hL, g, lcfL, aL);
        if (!capR) hlld_cf_rev(hR, g, lcfR, aR);
    }
    hlld_side_finish(L, hL, g, lamF, aL, lamL);
    hlld_side_finish(R, hR, g, lamF, aR, lamR);
}

// Face VJP of the solver SOLVER, reverse mode (by hand) for all three solvers; the
// forward-mode dual-number VJP (riemann_vjp, one pass per input) is the cross-check.
template <int SOLVER>
__device__ __forceinline__ void riemann_rev(const FaceState &L, const FaceState &R, double g,
                                            const double lamF[5], double lamL[6], double lamR[6]) {
    if (SOLVER == RUSANOV) rusanov_vjp(L, R, g, lamF, lamL, lamR);
    else if (SOLVER == HLLE) hlle_vjp(L, R, g, lamF, lamL, lamR);
    else hlld_vjp(L, R, g, lamF, lamL, lamR);
}

}  // namespace chimera
