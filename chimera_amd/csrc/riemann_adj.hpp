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
    lpth += len / (g - 1.0);
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
    if (!capL && xL > 0.0 && sqL > 0.0) {
        double lx = lcfL * (0.5 / sqL);
        lpthL += lx * g / L.rho;
        lhL += lx * 2.0 / L.rho;
        lamL[0] -= lx * xL / L.rho;
    }
    if (!capR && xR > 0.0 && sqR > 0.0) {
        double lx = lcfR * (0.5 / sqR);
        lpthR += lx * g / R.rho;
        lhR += lx * 2.0 / R.rho;
        lamR[0] -= lx * xR / R.rho;
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
    if (cL2 > 0.0 && sqL > 0.0) {
        double lc = lsL * (0.5 / sqL);
        lpthL += lc * g / L.rho;
        lhL += lc * 2.0 / L.rho;
        lamL[0] -= lc * cL2 / L.rho;
    }
    if (cR2 > 0.0 && sqR > 0.0) {
        double lc = lsR * (0.5 / sqR);
        lpthR += lc * g / R.rho;
        lhR += lc * 2.0 / R.rho;
        lamR[0] -= lc * cR2 / R.rho;
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

// Directions per forward-mode pass of the HLLD face VJP (12 inputs: 12 / N passes, each
// carrying the primal and N tangents).  Every N gives the same bits.
#ifndef CHIMERA_HLLD_DUAL_N
#define CHIMERA_HLLD_DUAL_N 1
#endif

// Face VJP of the solver SOLVER: reverse mode for Rusanov and HLLE, forward-mode dual
// numbers for HLLD.
template <int SOLVER>
__device__ __forceinline__ void riemann_rev(const FaceState &L, const FaceState &R, double g,
                                            const double lamF[5], double lamL[6], double lamR[6]) {
    if (SOLVER == RUSANOV) rusanov_vjp(L, R, g, lamF, lamL, lamR);
    else if (SOLVER == HLLE) hlle_vjp(L, R, g, lamF, lamL, lamR);
    else riemann_vjp<HLLD, CHIMERA_HLLD_DUAL_N>(L, R, g, lamF, lamL, lamR);
}

}  // namespace chimera
