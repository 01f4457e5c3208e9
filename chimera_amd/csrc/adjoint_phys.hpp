// Hand-written vector-Jacobian products of the point operations of the stage
// (adjoint of mhd_phys.hpp), each the reverse of the forward function it names and
// differentiating the branch the forward pass took.
#pragma once

#include "mhd_phys.hpp"

namespace chimera {

// reverse of cons_to_prim (src/mhd_change_states.f90:91-110).
// in: conserved + cell B, lam of (rho, vx, vy, P); out: lam of (M, Mx, My, E, Bx, By)
// (accumulated into lU[4], lBx, lBy).
__device__ __forceinline__ void cons_to_prim_vjp(double M, double Mx, double My, double E,
                                                 double Bx, double By, double g, double vol,
                                                 const double lp[4], double lU[4], double &lBx,
                                                 double &lBy) {
    double a = M / vol;
    bool r_is_a = !(RHO_FLOOR > a);          // fmax(a, floor) keeps a unless floor > a
    double r = r_is_a ? a : RHO_FLOOR;
    double u = Mx / (r * vol);
    double w = My / (r * vol);
    double halfB2 = 0.5 * (Bx * Bx + By * By);
    double ke = 0.5 * r * (u * u + w * w);
    double p0 = (E / vol - ke - halfB2) * (g - 1.0) + halfB2;
    bool p1_is_p0 = !(P_FLOOR > p0);
    double p1 = p1_is_p0 ? p0 : P_FLOOR;
    double q = halfB2 + E_FLOOR_FRAC * (ke + halfB2);
    bool p_is_q = q > p1;

    double lr = lp[0], lu = lp[1], lw = lp[2];
    double lP = lp[3];
    double lq = p_is_q ? lP : 0.0;
    double lp1 = p_is_q ? 0.0 : lP;
    double lhalf = lq * (1.0 + E_FLOOR_FRAC);
    double lke = lq * E_FLOOR_FRAC;
    double lp0 = p1_is_p0 ? lp1 : 0.0;
    double gm1 = g - 1.0;
    const double ivol = 1.0 / vol, ir = 1.0 / r, irv = ir * ivol;   // reverse only
    lU[3] += lp0 * gm1 * ivol;
    lke += -lp0 * gm1;
    lhalf += lp0 * (1.0 - gm1);
    lr += lke * 0.5 * (u * u + w * w);
    lu += lke * r * u;
    lw += lke * r * w;
    lU[1] += lu * irv;
    lU[2] += lw * irv;
    lr += -(lu * u + lw * w) * ir;
    if (r_is_a) lU[0] += lr * ivol;
    lBx += lhalf * Bx;
    lBy += lhalf * By;
}

// reverse of conserved_floors (src/mhd_change_states.f90:143-164), B stale cell B.
// U: inputs of the floor; lout: lam of the outputs; lin: lam of the inputs (set).
__device__ __forceinline__ void floors_vjp(const double U[4], double Bx, double By, double vol,
                                           double g, const double lout[4], double lin[4],
                                           double &lBx, double &lBy) {
    const double V_MAX = 50.0;
    double halfB2 = 0.5 * (Bx * Bx + By * By);
    double M = U[0], Mx = U[1], My = U[2], E = U[3];
    bool mass_floor = M < RHO_FLOOR * vol;
    bool e_floor = false;
    if (mass_floor) {
        Mx = 0.0;
        My = 0.0;
        M = RHO_FLOOR * vol;
        double Ef = (P_FLOOR / (g - 1.0) + halfB2) * vol;
        e_floor = Ef > E;
    }
    double mom = sqrt(Mx * Mx + My * My);
    bool cap = mom / M > V_MAX;
    double lM = lout[0], lMx = lout[1], lMy = lout[2], lE = lout[3];
    if (cap) {
        double s = V_MAX * M / mom;
        double ls = lMx * Mx + lMy * My;
        double lmom = -ls * s / mom;
        lM += ls * V_MAX / mom;
        lMx = lMx * s + lmom * Mx / mom;
        lMy = lMy * s + lmom * My / mom;
    }
    if (mass_floor) {
        lin[0] = 0.0;
        lin[1] = 0.0;
        lin[2] = 0.0;
        if (e_floor) {
            lin[3] = 0.0;
            double lh = lE * vol;
            lBx += lh * Bx;
            lBy += lh * By;
        } else {
            lin[3] = lE;
        }
    } else {
        lin[0] = lM;
        lin[1] = lMx;
        lin[2] = lMy;
        lin[3] = lE;
    }
}

// Partial derivatives of limited_slope (src/mhd_derivatives.f90:82-97) with respect
// to (fm, f0, fp) for the branch taken.
template <int LIMITER>
__device__ __forceinline__ void limited_slope_grad(double fm, double f0, double fp, double inv_dx,
                                                   double &dm, double &d0, double &dp) {
    double dfL = (f0 - fm) * inv_dx;
    double dfR = (fp - f0) * inv_dx;
    double prod = dfL * dfR;
    double den = dfL + dfR;
    dm = d0 = dp = 0.0;
    if (prod <= 0.0 || fabs(den) < 1e-12) return;
    double sL, sR;  // d slope / d dfL, d slope / d dfR
    if (LIMITER == LIM_VAN_LEER) {
        const double r = 1.0 / den;
        const double a = dfR * r, b = dfL * r;
        sL = 2.0 * a * a;
        sR = 2.0 * b * b;
    } else {
        double m1 = 2.0 * fabs(dfL), m2 = 2.0 * fabs(dfR), m3 = 0.5 * fabs(den);
        double m12 = (m2 < m1) ? m2 : m1;  // fmin keeps the first on ties
        bool pick3 = m3 < m12;
        if (pick3) {
            sL = 0.5;
            sR = 0.5;
        } else if (m2 < m1) {
            sL = 0.0;
            sR = 2.0;
        } else {
            sL = 2.0;
            sR = 0.0;
        }
    }
    dm = -sL * inv_dx;
    d0 = (sL - sR) * inv_dx;
    dp = sR * inv_dx;
}

// reverse of the outflow correction (src/main.f90:116-148) on one boundary cell:
// lam of the corrected conserved -> lam of the cell's primitives (accumulated).
// dir 0: x (xlo), 1: y (ylo).
__device__ __forceinline__ void outflow_vjp(int dir, const double P6[6], double g, double dtdx,
                                            const double lU[4], double lP6[6]) {
    double rho = P6[0], P = P6[3];
    // normal / tangential roles
    double un = dir == 0 ? P6[1] : P6[2], ut = dir == 0 ? P6[2] : P6[1];
    double Bn = dir == 0 ? P6[4] : P6[5], Bt = dir == 0 ? P6[5] : P6[4];
    double lM = lU[0], lMn = dir == 0 ? lU[1] : lU[2], lMt = dir == 0 ? lU[2] : lU[1];
    double lE = lU[3];
    double h = 0.5 * (Bn * Bn + Bt * Bt);
    double en = (P - h) / (g - 1.0) + 0.5 * rho * (un * un + ut * ut) + h;
    double t = dtdx;
    double len = t * lE * un;
    double lrho = t * lM * un + t * lMn * un * un + t * lMt * un * ut + len * 0.5 * (un * un + ut * ut);
    double lun = t * lM * rho + t * lMn * 2.0 * rho * un + t * lMt * rho * ut +
                 t * lE * ((en + P) - Bn * Bn) + len * rho * un;
    double lut = t * lMt * rho * un - t * lE * Bn * Bt + len * rho * ut;
    double lPP = t * lMn + t * lE * un + len / (g - 1.0);
    double lh = len * (1.0 - 1.0 / (g - 1.0));
    double lBn = -t * lMn * 2.0 * Bn - t * lMt * Bt - t * lE * (2.0 * Bn * un + Bt * ut) + lh * Bn;
    double lBt = -t * lMt * Bn - t * lE * Bn * ut + lh * Bt;
    lP6[0] += lrho;
    lP6[3] += lPP;
    if (dir == 0) {
        lP6[1] += lun;
        lP6[2] += lut;
        lP6[4] += lBn;
        lP6[5] += lBt;
    } else {
        lP6[2] += lun;
        lP6[1] += lut;
        lP6[5] += lBn;
        lP6[4] += lBt;
    }
}

}  // namespace chimera
