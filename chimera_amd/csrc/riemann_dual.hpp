// The three Riemann solvers of mhd_phys.hpp over dual numbers, operation for
// operation, for the adjoint's flux Jacobians (src/mhd_flux.f90:197-779).  Branch
// conditions compare values computed exactly as in the FP64 solvers, so each
// derivative is that of the branch the forward pass took.
#pragma once

#include "dual.hpp"
#include "mhd_phys.hpp"

namespace chimera {

template <int N>
struct FaceD {
    Dual<N> rho, u, w, P, Bn, Bt;
};
template <int N>
struct FluxD {
    Dual<N> mass, mom_n, mom_t, energy, bt;
};

template <int N>
__device__ __forceinline__ FluxD<N> rusanov_d(const FaceD<N> &L, const FaceD<N> &R, double g) {
    using D = Dual<N>;
    D halfBL2 = 0.5 * (L.Bn * L.Bn + L.Bt * L.Bt);
    D halfBR2 = 0.5 * (R.Bn * R.Bn + R.Bt * R.Bt);
    D pthL = dmax(L.P - halfBL2, P_FLOOR);
    D pthR = dmax(R.P - halfBR2, P_FLOOR);
    double gm1 = g - 1.0;
    D enL = pthL / gm1 + 0.5 * L.rho * (L.u * L.u + L.w * L.w) + halfBL2;
    D enR = pthR / gm1 + 0.5 * R.rho * (R.u * R.u + R.w * R.w) + halfBR2;
    D rho_avg = 0.5 * (L.rho + R.rho);
    D inv = 1.0 / rho_avg;
    D mLn = L.rho * L.u, mRn = R.rho * R.u;
    D mLt = L.rho * L.w, mRt = R.rho * R.w;
    D mx = 0.5 * (mLn + mRn);
    D my = 0.5 * (mLt + mRt);
    D en = 0.5 * (enL + enR);
    D bx = 0.5 * (L.Bn + R.Bn);
    D by = 0.5 * (L.Bt + R.Bt);
    D hb = 0.5 * (bx * bx + by * by);
    D Pavg = gm1 * (en - 0.5 * (mx * mx + my * my) * inv - hb) + hb;
    D cL2 = (g * pthL + 2.0 * halfBL2) / L.rho;
    D cR2 = (g * pthR + 2.0 * halfBR2) / R.rho;
    D sL = dsqrt(0.5 * (cL2 + dabs(cL2))) + dabs(L.u);
    D sR = dsqrt(0.5 * (cR2 + dabs(cR2))) + dabs(R.u);
    D C = 0.5 * dmax(sL, sR);
    FluxD<N> f;
    f.mass = mx - C * (R.rho - L.rho);
    f.mom_n = (mx * mx * inv + Pavg - bx * bx) - C * (mRn - mLn);
    f.mom_t = (mx * my * inv - bx * by) - C * (mRt - mLt);
    f.energy = ((en + Pavg) * mx * inv - bx * (bx * mx + by * my) * inv) - C * (enR - enL);
    f.bt = ((by * mx - bx * my) * inv) - C * (R.Bt - L.Bt);
    return f;
}

template <int N>
__device__ __forceinline__ FluxD<N> hll_avg_d(const FluxD<N> &FL, const FluxD<N> &FR,
                                              const Dual<N> &SL, const Dual<N> &SR,
                                              const FaceD<N> &L, const FaceD<N> &R,
                                              const Dual<N> &enL, const Dual<N> &enR,
                                              bool floor_den) {
    using D = Dual<N>;
    D inv = floor_den ? 1.0 / dmax(SR - SL, 1.0e-12) : 1.0 / (SR - SL);
    D sLR = SL * SR;
    FluxD<N> f;
    f.mass = (SR * FL.mass - SL * FR.mass + sLR * (R.rho - L.rho)) * inv;
    f.mom_n = (SR * FL.mom_n - SL * FR.mom_n + sLR * (R.rho * R.u - L.rho * L.u)) * inv;
    f.mom_t = (SR * FL.mom_t - SL * FR.mom_t + sLR * (R.rho * R.w - L.rho * L.w)) * inv;
    f.energy = (SR * FL.energy - SL * FR.energy + sLR * (enR - enL)) * inv;
    f.bt = (SR * FL.bt - SL * FR.bt + sLR * (R.Bt - L.Bt)) * inv;
    return f;
}

template <int N>
__device__ __forceinline__ FluxD<N> hlle_d(const FaceD<N> &L, const FaceD<N> &R, double g) {
    using D = Dual<N>;
    const double cf_max = 1.0e2;
    D halfBL2 = 0.5 * (L.Bn * L.Bn + L.Bt * L.Bt);
    D halfBR2 = 0.5 * (R.Bn * R.Bn + R.Bt * R.Bt);
    D pthL = dmax(L.P - halfBL2, P_FLOOR);
    D pthR = dmax(R.P - halfBR2, P_FLOOR);
    double gm1 = g - 1.0;
    D enL = pthL / gm1 + 0.5 * L.rho * (L.u * L.u + L.w * L.w) + halfBL2;
    D enR = pthR / gm1 + 0.5 * R.rho * (R.u * R.u + R.w * R.w) + halfBR2;
    D xL = (g * pthL + 2.0 * halfBL2) / L.rho;
    D xR = (g * pthR + 2.0 * halfBR2) / R.rho;
    D cfL = dmin(dsqrt(0.5 * (xL + dabs(xL))), cf_max);
    D cfR = dmin(dsqrt(0.5 * (xR + dabs(xR))), cf_max);
    D SL = dmin(L.u - cfL, R.u - cfR);
    D SR = dmax(L.u + cfL, R.u + cfR);
    D mLn = L.rho * L.u, mRn = R.rho * R.u;
    FluxD<N> FL, FR;
    FL.mass = mLn;
    FL.mom_n = mLn * L.u + L.P - L.Bn * L.Bn;
    FL.mom_t = mLn * L.w - L.Bn * L.Bt;
    FL.energy = (enL + L.P) * L.u - L.Bn * (L.Bn * L.u + L.Bt * L.w);
    FL.bt = L.Bt * L.u - L.Bn * L.w;
    FR.mass = mRn;
    FR.mom_n = mRn * R.u + R.P - R.Bn * R.Bn;
    FR.mom_t = mRn * R.w - R.Bn * R.Bt;
    FR.energy = (enR + R.P) * R.u - R.Bn * (R.Bn * R.u + R.Bt * R.w);
    FR.bt = R.Bt * R.u - R.Bn * R.w;
    if (SR <= 0.0) return FR;
    if (SL >= 0.0) return FL;
    return hll_avg_d(FL, FR, SL, SR, L, R, enL, enR, true);
}

template <int N>
__device__ __forceinline__ bool finite_in_d(const Dual<N> &f, const Dual<N> &bound) {
    return dfinite(f) && !(fabs(f.v) > 8.0 * bound.v);
}

template <int N>
__device__ __forceinline__ FluxD<N> hlld_d(const FaceD<N> &L, const FaceD<N> &R, double g) {
    using D = Dual<N>;
    const double EPS_FAN = 1.0e-6, cf_max = 50.0;
    D rL = L.rho, rR = R.rho, uL = L.u, uR = R.u, wL = L.w, wR = R.w;
    D pTL = L.P, pTR = R.P, BL = L.Bn, BR = R.Bn, CL = L.Bt, CR = R.Bt;
    D halfBL2 = 0.5 * (BL * BL + CL * CL);
    D halfBR2 = 0.5 * (BR * BR + CR * CR);
    D pthL = dmax(pTL - halfBL2, P_FLOOR);
    D pthR = dmax(pTR - halfBR2, P_FLOOR);
    double gm1 = g - 1.0;
    D enL = pthL / gm1 + 0.5 * rL * (uL * uL + wL * wL) + halfBL2;
    D enR = pthR / gm1 + 0.5 * rR * (uR * uR + wR * wR) + halfBR2;
    D vdotBL = uL * BL + wL * CL;
    D vdotBR = uR * BR + wR * CR;
    FluxD<N> FL, FR;
    FL.mass = rL * uL;
    FL.mom_n = rL * (uL * uL) + pTL - BL * BL;
    FL.mom_t = rL * uL * wL - BL * CL;
    FL.energy = (enL + pTL) * uL - BL * vdotBL;
    FL.bt = CL * uL - BL * wL;
    FR.mass = rR * uR;
    FR.mom_n = rR * (uR * uR) + pTR - BR * BR;
    FR.mom_t = rR * uR * wR - BR * CR;
    FR.energy = (enR + pTR) * uR - BR * vdotBR;
    FR.bt = CR * uR - BR * wR;
    D BL2 = BL * BL, BR2 = BR * BR;
    D caL = BL2 / rL, caR = BR2 / rR;
    D csL2 = g * pthL / rL, csR2 = g * pthR / rR;
    D b2L = (BL2 + CL * CL) / rL, b2R = (BR2 + CR * CR) / rR;
    D tmpL = csL2 + b2L, tmpR = csR2 + b2R;
    D cfL = dsqrt(0.5 * (tmpL + dsqrt(dmax(tmpL * tmpL - 4.0 * csL2 * caL, 0.0))));
    D cfR = dsqrt(0.5 * (tmpR + dsqrt(dmax(tmpR * tmpR - 4.0 * csR2 * caR, 0.0))));
    cfL = dmin(cfL, cf_max);
    cfR = dmin(cfR, cf_max);
    D SL = dmin(uL - cfL, uR - cfR);
    D SR = dmax(uL + cfL, uR + cfR);
    if (cfL >= cf_max || cfR >= cf_max) return (cfL <= cfR) ? FL : FR;
    if (SL >= 0.0) return FL;
    if (SR <= 0.0) return FR;
    D aL = rL * (SL - uL);
    D aR = rR * (SR - uR);
    D denMD = aR - aL;
    if (fabs(denMD.v) < 1.0e-15) denMD = dconst<N>(copysign(1.0e-15, denMD.v));
    D SM = (rR * uR * (SR - uR) - rL * uL * (SL - uL) - (pTR - pTL)) / denMD;
    D pTs = pTL + aL * (SM - uL);
    if (SM <= SL + EPS_FAN * (SR - SL) || SM >= SR - EPS_FAN * (SR - SL))
        return hll_avg_d(FL, FR, SL, SR, L, R, enL, enR, false);
    pTs = dmax(pTs, P_FLOOR);
    D dSL = SL - SM, dSR = SR - SM;
    D rLs = aL / dSL;
    D kL = aL * dSL;
    D denL = kL - BL2;
    D vLs = wL, BLs = CL;
    if (fabs(denL.v) > 1.0e-3 * fmax(fmax(fabs(kL.v), BL2.v), 1.0e-30)) {
        vLs = wL - BL * CL * (SM - uL) / denL;
        D t = SL - uL;
        BLs = CL * (rL * (t * t) - BL2) / denL;
    }
    D vdotBLs = SM * BL + vLs * BLs;
    D enLs = (enL * (SL - uL) - pTL * uL + pTs * SM + BL * (vdotBL - vdotBLs)) / dSL;
    D rRs = aR / dSR;
    D kR = aR * dSR;
    D denR = kR - BR2;
    D vRs = wR, BRs = CR;
    if (fabs(denR.v) > 1.0e-3 * fmax(fmax(fabs(kR.v), BR2.v), 1.0e-30)) {
        vRs = wR - BR * CR * (SM - uR) / denR;
        D t = SR - uR;
        BRs = CR * (rR * (t * t) - BR2) / denR;
    }
    D vdotBRs = SM * BR + vRs * BRs;
    D enRs = (enR * (SR - uR) - pTR * uR + pTs * SM + BR * (vdotBR - vdotBRs)) / dSR;
    D sqL = dsqrt(dmax(rLs, RHO_FLOOR));
    D sqR = dsqrt(dmax(rRs, RHO_FLOOR));
    D Bavg = 0.5 * (BL + BR);
    D SLs = SM - dabs(Bavg) / sqL;
    D SRs = SM + dabs(Bavg) / sqR;
    D vss, Bss, enLss, enRss;
    if (fabs(Bavg.v) > 1.0e-10) {
        double sgn = copysign(1.0, Bavg.v);
        D dss = sqL + sqR;
        vss = (sqL * vLs + sqR * vRs + sgn * (BRs - BLs)) / dss;
        Bss = (sqL * BRs + sqR * BLs + sgn * sqL * sqR * (vRs - vLs)) / dss;
        D vdotBss = SM * Bavg + vss * Bss;
        enLss = enLs - sqL * sgn * (vdotBLs - vdotBss);
        enRss = enRs + sqR * sgn * (vdotBRs - vdotBss);
    } else {
        vss = 0.5 * (vLs + vRs);
        Bss = 0.5 * (BLs + BRs);
        enLss = enLs;
        enRss = enRs;
        SLs = SM;
        SRs = SM;
    }
    FluxD<N> f;
    if (SM >= 0.0) {
        f.mass = FL.mass + SL * (rLs - rL);
        f.mom_n = FL.mom_n + SL * (rLs * SM - rL * uL);
        f.mom_t = FL.mom_t + SL * (rLs * vLs - rL * wL);
        f.energy = FL.energy + SL * (enLs - enL);
        f.bt = FL.bt + SL * (BLs - CL);
        if (!(SLs >= 0.0)) {
            f.mom_t = f.mom_t + SLs * rLs * (vss - vLs);
            f.energy = f.energy + SLs * (enLss - enLs);
            f.bt = f.bt + SLs * (Bss - BLs);
        }
    } else {
        f.mass = FR.mass + SR * (rRs - rR);
        f.mom_n = FR.mom_n + SR * (rRs * SM - rR * uR);
        f.mom_t = FR.mom_t + SR * (rRs * vRs - rR * wR);
        f.energy = FR.energy + SR * (enRs - enR);
        f.bt = FR.bt + SR * (BRs - CR);
        if (!(SRs <= 0.0)) {
            f.mom_t = f.mom_t + SRs * rRs * (vss - vRs);
            f.energy = f.energy + SRs * (enRss - enRs);
            f.bt = f.bt + SRs * (Bss - BRs);
        }
    }
    D Smax = dmax(dabs(SL), dabs(SR));
    bool ok = finite_in_d(f.mass, dmax(dabs(FL.mass), dabs(FR.mass)) + Smax * (rL + rR)) &&
              finite_in_d(f.mom_n, dmax(dabs(FL.mom_n), dabs(FR.mom_n)) +
                                       Smax * (dabs(rL * uL) + dabs(rR * uR))) &&
              finite_in_d(f.mom_t, dmax(dabs(FL.mom_t), dabs(FR.mom_t)) +
                                       Smax * (dabs(rL * wL) + dabs(rR * wR))) &&
              finite_in_d(f.energy, dmax(dabs(FL.energy), dabs(FR.energy)) +
                                        Smax * (dabs(enL) + dabs(enR))) &&
              finite_in_d(f.bt, dmax(dabs(FL.bt), dabs(FR.bt)) + Smax * (dabs(CL) + dabs(CR)));
    if (!ok) return hll_avg_d(FL, FR, SL, SR, L, R, enL, enR, false);
    return f;
}

template <int SOLVER, int N>
__device__ __forceinline__ FluxD<N> riemann_d(const FaceD<N> &L, const FaceD<N> &R, double g) {
    if (SOLVER == RUSANOV) return rusanov_d<N>(L, R, g);
    if (SOLVER == HLLE) return hlle_d<N>(L, R, g);
    return hlld_d<N>(L, R, g);
}

// lam_L / lam_R (6 each: rho, u, w, P, Bn, Bt) = J^T lam_F for one face, J from
// 12 / N forward-mode passes with N seeded inputs each.
template <int SOLVER, int N>
__device__ __forceinline__ void riemann_vjp(const FaceState &L, const FaceState &R, double g,
                                            const double lamF[5], double lamL[6], double lamR[6]) {
    const double in[12] = {L.rho, L.u, L.w, L.P, L.Bn, L.Bt, R.rho, R.u, R.w, R.P, R.Bn, R.Bt};
    double out[12];
    for (int base = 0; base < 12; base += N) {
        Dual<N> x[12];
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            x[q] = dconst<N>(in[q]);
#pragma unroll
            for (int k = 0; k < N; ++k) x[q].d[k] = (q == base + k) ? 1.0 : 0.0;
        }
        FaceD<N> a{x[0], x[1], x[2], x[3], x[4], x[5]};
        FaceD<N> b{x[6], x[7], x[8], x[9], x[10], x[11]};
        FluxD<N> f = riemann_d<SOLVER, N>(a, b, g);
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (base + k < 12)
                out[base + k] = lamF[0] * f.mass.d[k] + lamF[1] * f.mom_n.d[k] +
                                lamF[2] * f.mom_t.d[k] + lamF[3] * f.energy.d[k] +
                                lamF[4] * f.bt.d[k];
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        lamL[q] = out[q];
        lamR[q] = out[6 + q];
    }
}

}  // namespace chimera
