// Device-side point physics of the CHIMERA stage update (gfx950, FP64).
//
// Every function keeps the reference's floating-point evaluation order (Fortran
// evaluates a*b*c as (a*b)*c and x**2 as x*x) and the library is compiled with
// -ffp-contract=off, so results match the reference Fortran bit for bit.  The
// reference line ranges are cited per function.
#pragma once

#include <hip/hip_runtime.h>

namespace chimera {

// src/mhd_config.f90:40-61
constexpr int BC_PERIODIC = 1;
constexpr int LIM_VAN_LEER = 1;
constexpr int LIM_MC = 2;
constexpr int RUSANOV = 1;
constexpr int HLLE = 2;
constexpr int HLLD = 3;
constexpr double P_FLOOR = 1.0e-12;
constexpr double RHO_FLOOR = 1.0e-12;
constexpr double E_FLOOR_FRAC = 1.0e-10;
constexpr double COURANT = 0.3;
constexpr double T_OUT = 0.01;

struct Prim {
    double rho, vx, vy, P, Bx, By;
};

// get_primitive, src/mhd_change_states.f90:91-110 (P is total pressure).
__device__ __forceinline__ void cons_to_prim(double M, double Mx, double My, double E,
                                             double Bx, double By, double g, double vol,
                                             double &rho, double &vx, double &vy, double &P) {
    double r = fmax(M / vol, RHO_FLOOR);
    double u = Mx / (r * vol);
    double w = My / (r * vol);
    double halfB2 = 0.5 * (Bx * Bx + By * By);
    double ke = 0.5 * r * (u * u + w * w);
    double p = (E / vol - ke - halfB2) * (g - 1.0) + halfB2;
    p = fmax(p, P_FLOOR);
    p = fmax(p, halfB2 + E_FLOOR_FRAC * (ke + halfB2));
    rho = r;
    vx = u;
    vy = w;
    P = p;
}

// get_conserved, src/mhd_change_states.f90:45-56.
__device__ __forceinline__ void prim_to_cons(double rho, double vx, double vy, double P,
                                             double Bx, double By, double g, double vol,
                                             double &M, double &Mx, double &My, double &E) {
    double r = fmax(rho, RHO_FLOOR);
    double p = fmax(P, P_FLOOR);
    double m = r * vol;
    double halfB2 = 0.5 * (Bx * Bx + By * By);
    M = m;
    Mx = m * vx;
    My = m * vy;
    E = ((p - halfB2) / (g - 1.0) + 0.5 * r * (vx * vx + vy * vy) + halfB2) * vol;
}

// apply_conserved_floors, src/mhd_change_states.f90:143-164 (B is the stage's
// cell-centred field, not yet advanced by CT).
__device__ __forceinline__ void conserved_floors(double &M, double &Mx, double &My, double &E,
                                                 double Bx, double By, double vol, double g) {
    const double V_MAX = 50.0;
    double halfB2 = 0.5 * (Bx * Bx + By * By);
    if (M < RHO_FLOOR * vol) {
        Mx = 0.0;
        My = 0.0;
        M = RHO_FLOOR * vol;
        E = fmax(E, (P_FLOOR / (g - 1.0) + halfB2) * vol);
    }
    double mom = sqrt(Mx * Mx + My * My);
    if (mom / M > V_MAX) {
        double scale = V_MAX * M / mom;
        Mx = Mx * scale;
        My = My * scale;
    }
}

// Signal speed cf + |v| of the CFL condition, src/main.f90:72-79.
__device__ __forceinline__ double cfl_speed(double rho, double vx, double vy, double P,
                                            double Bx, double By, double g) {
    double inv_rho = 1.0 / rho;
    double B2 = Bx * Bx + By * By;
    double c0_sq = g * (P - 0.5 * B2) * inv_rho;
    double ca_sq = B2 * inv_rho;
    double cf = sqrt(c0_sq + ca_sq);
    if (rho < 1.0e-3) cf = fmin(cf, 50.0);
    return cf + sqrt(vx * vx + vy * vy);
}

// ---- FP64 division and square root for operands in a known range ---------------
// hipcc lowers a / b to d = div_scale(b), n = div_scale(a), y = rcp(d), two Newton
// steps y += y (1 - d y), q = n y, r = n - d q, q' = div_fmas(r, y, q) and
// div_fixup(q', b, a).  div_scale rescales only an operand that is zero, infinite, NaN
// or denormal, a numerator below 2^-969, a reciprocal or quotient in the denormal
// range or exponents more than 767 apart; otherwise both are the identity, div_fmas
// is a plain fma and div_fixup returns q' (with the sign of a / b, which q' already
// has).  sqrt(x) is rsq(x) and the coupled Newton steps below, with a 2^256 pre-scale
// for x < 2^-767 and a select for zero and infinity.  For finite nonzero a and b with
// |a|, |b| and |a / b| in [2^-900, 2^700], and for x in [2^-700, 2^700] -- the callers
// establish this before they use the short forms -- the core steps alone are the
// whole computation, so these return the bits of a / b and sqrt(x).
struct Recip {
    double b, y;   // the divisor and its Newton-refined reciprocal
};
__device__ __forceinline__ Recip recip_nr(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    return {b, y};
}
__device__ __forceinline__ double div_nr(double a, Recip r) {
    const double q = a * r.y;
    const double e = fma(-r.b, q, a);
    return fma(e, r.y, q);
}
__device__ __forceinline__ double sqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = y * 0.5;
    const double e = fma(-h, s, 0.5);
    s = fma(s, e, s);
    h = fma(h, e, h);
    double d = fma(-s, s, x);
    s = fma(d, h, s);
    d = fma(-s, s, x);
    return fma(d, h, s);
}
// a / b and sqrt(x): the short forms when the operands are in range (SHORT).
template <bool SHORT>
__device__ __forceinline__ double qdiv(double a, double b) {
    return SHORT ? div_nr(a, recip_nr(b)) : a / b;
}
template <bool SHORT>
__device__ __forceinline__ double qsqrt(double x) {
    return SHORT ? sqrt_nr(x) : sqrt(x);
}
// The reciprocal of a wave-uniform divisor, held in scalar registers.
__device__ __forceinline__ Recip recip_uniform(double b) {
    const Recip r = recip_nr(b);
    const unsigned long long u = (unsigned long long)__double_as_longlong(r.y);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return {b, __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo))};
}

// The slope limiter of src/mhd_derivatives.f90:82-97 (MC / Van Leer, zero at extrema
// and flat spots) from the one-sided differences dfL = (f0 - fm) / dx and
// dfR = (fp - f0) / dx; a cell's dfR is its right neighbour's dfL, bit for bit.
template <int LIMITER>
__device__ __forceinline__ double limiter_of(double dfL, double dfR) {
    double prod = dfL * dfR;
    double den = dfL + dfR;
    if (prod <= 0.0 || fabs(den) < 1e-12) return 0.0;
    if (LIMITER == LIM_VAN_LEER) return (2.0 * prod) / den;
    double m = fmin(fmin(2.0 * fabs(dfL), 2.0 * fabs(dfR)), 0.5 * fabs(den));
    return copysign(1.0, dfL) * m;
}

// One limited slope from three consecutive values.
template <int LIMITER>
__device__ __forceinline__ double limited_slope(double fm, double f0, double fp, double inv_dx) {
    return limiter_of<LIMITER>((f0 - fm) * inv_dx, (fp - f0) * inv_dx);
}

// Central slope of compute_gradients, src/mhd_derivatives.f90:37-45 (used only
// when slope limiting is switched off).
__device__ __forceinline__ double central_slope(double fm, double fp, double inv_2dx) {
    return (fp - fm) * inv_2dx;
}

// A face state as the Riemann solvers see it: normal (u, Bn) / tangential (w, Bt).
struct FaceState {
    double rho, u, w, P, Bn, Bt;
};

struct Flux5 {
    double mass, mom_n, mom_t, energy, bt;
};

// rusanov_flux, src/mhd_flux.f90:238-286: the physical flux of the averaged state
// minus C/2-scaled jumps, C = max of cf + |u| on the two sides.
//
// SHORT (the short division and square root, same bits): for face states with rho, P
// in [1e-8, 2^66], |u|, |w|, |Bn|, |Bt| <= 2^66 and g - 1 = rg.b in [2^-100, 2^100]
// (rg its reciprocal), P_th / (g - 1) is in [2^-140, 2^233], 1 / rho_avg in
// [2^-66, 2^27] and c^2 = (g P_th + B^2) / rho, the square root's argument, in
// [2^-106, 2^261].
template <bool SHORT = false>
__device__ __forceinline__ Flux5 rusanov(const FaceState &L, const FaceState &R, double g,
                                         const Recip &rg = Recip{}) {
    double halfBL2 = 0.5 * (L.Bn * L.Bn + L.Bt * L.Bt);
    double halfBR2 = 0.5 * (R.Bn * R.Bn + R.Bt * R.Bt);
    double pthL = fmax(L.P - halfBL2, P_FLOOR);
    double pthR = fmax(R.P - halfBR2, P_FLOOR);
    double gm1 = g - 1.0;
    double enL = (SHORT ? div_nr(pthL, rg) : pthL / gm1) +
                 0.5 * L.rho * (L.u * L.u + L.w * L.w) + halfBL2;
    double enR = (SHORT ? div_nr(pthR, rg) : pthR / gm1) +
                 0.5 * R.rho * (R.u * R.u + R.w * R.w) + halfBR2;
    double rho_avg = 0.5 * (L.rho + R.rho);
    double inv = qdiv<SHORT>(1.0, rho_avg);
    double mLn = L.rho * L.u, mRn = R.rho * R.u;
    double mLt = L.rho * L.w, mRt = R.rho * R.w;
    double mx = 0.5 * (mLn + mRn);
    double my = 0.5 * (mLt + mRt);
    double en = 0.5 * (enL + enR);
    double bx = 0.5 * (L.Bn + R.Bn);
    double by = 0.5 * (L.Bt + R.Bt);
    double hb = 0.5 * (bx * bx + by * by);
    double Pavg = gm1 * (en - 0.5 * (mx * mx + my * my) * inv - hb) + hb;
    double cL2 = qdiv<SHORT>(g * pthL + 2.0 * halfBL2, L.rho);
    double cR2 = qdiv<SHORT>(g * pthR + 2.0 * halfBR2, R.rho);
    double sL = qsqrt<SHORT>(0.5 * (cL2 + fabs(cL2))) + fabs(L.u);
    double sR = qsqrt<SHORT>(0.5 * (cR2 + fabs(cR2))) + fabs(R.u);
    double C = 0.5 * fmax(sL, sR);
    Flux5 f;
    f.mass = mx - C * (R.rho - L.rho);
    f.mom_n = (mx * mx * inv + Pavg - bx * bx) - C * (mRn - mLn);
    f.mom_t = (mx * my * inv - bx * by) - C * (mRt - mLt);
    f.energy = ((en + Pavg) * mx * inv - bx * (bx * mx + by * my) * inv) - C * (enR - enL);
    f.bt = ((by * mx - bx * my) * inv) - C * (R.Bt - L.Bt);
    return f;
}

// hlle_flux, src/mhd_flux.f90:343-403: Davis speeds with c_f capped at 100, HLL
// average, exact upwind where the fan does not straddle the face (S_R <= 0 wins).
// SHORT as for rusanov; 1 / max(S_R - S_L, 1e-12) then has a divisor in
// [1e-12, 2^68].
template <bool SHORT = false>
__device__ __forceinline__ Flux5 hlle(const FaceState &L, const FaceState &R, double g,
                                      const Recip &rg = Recip{}) {
    const double cf_max = 1.0e2;
    double halfBL2 = 0.5 * (L.Bn * L.Bn + L.Bt * L.Bt);
    double halfBR2 = 0.5 * (R.Bn * R.Bn + R.Bt * R.Bt);
    double pthL = fmax(L.P - halfBL2, P_FLOOR);
    double pthR = fmax(R.P - halfBR2, P_FLOOR);
    double gm1 = g - 1.0;
    double enL = (SHORT ? div_nr(pthL, rg) : pthL / gm1) +
                 0.5 * L.rho * (L.u * L.u + L.w * L.w) + halfBL2;
    double enR = (SHORT ? div_nr(pthR, rg) : pthR / gm1) +
                 0.5 * R.rho * (R.u * R.u + R.w * R.w) + halfBR2;
    double xL = qdiv<SHORT>(g * pthL + 2.0 * halfBL2, L.rho);
    double xR = qdiv<SHORT>(g * pthR + 2.0 * halfBR2, R.rho);
    double cfL = fmin(qsqrt<SHORT>(0.5 * (xL + fabs(xL))), cf_max);
    double cfR = fmin(qsqrt<SHORT>(0.5 * (xR + fabs(xR))), cf_max);
    double SL = fmin(L.u - cfL, R.u - cfR);
    double SR = fmax(L.u + cfL, R.u + cfR);
    double mLn = L.rho * L.u, mRn = R.rho * R.u;
    Flux5 FL, FR;
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
    double inv = qdiv<SHORT>(1.0, fmax(SR - SL, 1.0e-12));
    double sLR = SL * SR;
    Flux5 f;
    f.mass = (SR * FL.mass - SL * FR.mass + sLR * (R.rho - L.rho)) * inv;
    f.mom_n = (SR * FL.mom_n - SL * FR.mom_n + sLR * (mRn - mLn)) * inv;
    f.mom_t = (SR * FL.mom_t - SL * FR.mom_t + sLR * (R.rho * R.w - L.rho * L.w)) * inv;
    f.energy = (SR * FL.energy - SL * FR.energy + sLR * (enR - enL)) * inv;
    f.bt = (SR * FL.bt - SL * FR.bt + sLR * (R.Bt - L.Bt)) * inv;
    return f;
}

__device__ __forceinline__ Flux5 hll_average(const Flux5 &FL, const Flux5 &FR, double SL,
                                             double SR, const FaceState &L, const FaceState &R,
                                             double enL, double enR) {
    double inv = 1.0 / (SR - SL);
    double sLR = SL * SR;
    Flux5 f;
    f.mass = (SR * FL.mass - SL * FR.mass + sLR * (R.rho - L.rho)) * inv;
    f.mom_n = (SR * FL.mom_n - SL * FR.mom_n + sLR * (R.rho * R.u - L.rho * L.u)) * inv;
    f.mom_t = (SR * FL.mom_t - SL * FR.mom_t + sLR * (R.rho * R.w - L.rho * L.w)) * inv;
    f.energy = (SR * FL.energy - SL * FR.energy + sLR * (enR - enL)) * inv;
    f.bt = (SR * FL.bt - SL * FR.bt + sLR * (R.Bt - L.Bt)) * inv;
    return f;
}

__device__ __forceinline__ bool finite_in(double f, double bound) {
    // !ieee_is_finite(f) .or. |f| > 8*bound  <=>  !(|f| <= 8*bound) for finite bound
    return __builtin_isfinite(f) && !(fabs(f) > 8.0 * bound);
}

// hlld_flux, src/mhd_flux.f90:490-777 (Miyoshi & Kusano 2005): c_f cap 50 (catch 1),
// supersonic upwind, SM near the fan edges -> HLL (catch 2), p_T* floor (catch 3),
// degenerate fast/Alfven merge and |Bx| -> 0 branches, non-finite or out-of-bound
// catch-all -> HLL.
__device__ __forceinline__ Flux5 hlld(const FaceState &L, const FaceState &R, double g) {
    const double EPS_FAN = 1.0e-6, cf_max = 50.0;
    double rL = L.rho, rR = R.rho, uL = L.u, uR = R.u, wL = L.w, wR = R.w;
    double pTL = L.P, pTR = R.P, BL = L.Bn, BR = R.Bn, CL = L.Bt, CR = R.Bt;
    double halfBL2 = 0.5 * (BL * BL + CL * CL);
    double halfBR2 = 0.5 * (BR * BR + CR * CR);
    double pthL = fmax(pTL - halfBL2, P_FLOOR);
    double pthR = fmax(pTR - halfBR2, P_FLOOR);
    double gm1 = g - 1.0;
    double enL = pthL / gm1 + 0.5 * rL * (uL * uL + wL * wL) + halfBL2;
    double enR = pthR / gm1 + 0.5 * rR * (uR * uR + wR * wR) + halfBR2;

    double vdotBL = uL * BL + wL * CL;
    double vdotBR = uR * BR + wR * CR;
    Flux5 FL, FR;
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

    double BL2 = BL * BL, BR2 = BR * BR;
    double caL = BL2 / rL, caR = BR2 / rR;
    double csL2 = g * pthL / rL, csR2 = g * pthR / rR;
    double b2L = (BL2 + CL * CL) / rL, b2R = (BR2 + CR * CR) / rR;
    double tmpL = csL2 + b2L, tmpR = csR2 + b2R;
    double cfL = sqrt(0.5 * (tmpL + sqrt(fmax(tmpL * tmpL - 4.0 * csL2 * caL, 0.0))));
    double cfR = sqrt(0.5 * (tmpR + sqrt(fmax(tmpR * tmpR - 4.0 * csR2 * caR, 0.0))));
    cfL = fmin(cfL, cf_max);
    cfR = fmin(cfR, cf_max);
    double SL = fmin(uL - cfL, uR - cfR);
    double SR = fmax(uL + cfL, uR + cfR);

    if (cfL >= cf_max || cfR >= cf_max) return (cfL <= cfR) ? FL : FR;
    if (SL >= 0.0) return FL;
    if (SR <= 0.0) return FR;

    double aL = rL * (SL - uL);  // rho_L (S_L - u_L)
    double aR = rR * (SR - uR);
    double denMD = aR - aL;
    if (fabs(denMD) < 1.0e-15) denMD = copysign(1.0e-15, denMD);
    double SM = (rR * uR * (SR - uR) - rL * uL * (SL - uL) - (pTR - pTL)) / denMD;
    double pTs = pTL + aL * (SM - uL);

    if (SM <= SL + EPS_FAN * (SR - SL) || SM >= SR - EPS_FAN * (SR - SL))
        return hll_average(FL, FR, SL, SR, L, R, enL, enR);
    pTs = fmax(pTs, P_FLOOR);

    double dSL = SL - SM, dSR = SR - SM;
    double rLs = aL / dSL;
    double kL = aL * dSL;
    double denL = kL - BL2;
    double vLs = wL, BLs = CL;
    if (fabs(denL) > 1.0e-3 * fmax(fmax(fabs(kL), BL2), 1.0e-30)) {
        vLs = wL - BL * CL * (SM - uL) / denL;
        double t = SL - uL;
        BLs = CL * (rL * (t * t) - BL2) / denL;
    }
    double vdotBLs = SM * BL + vLs * BLs;
    double enLs = (enL * (SL - uL) - pTL * uL + pTs * SM + BL * (vdotBL - vdotBLs)) / dSL;

    double rRs = aR / dSR;
    double kR = aR * dSR;
    double denR = kR - BR2;
    double vRs = wR, BRs = CR;
    if (fabs(denR) > 1.0e-3 * fmax(fmax(fabs(kR), BR2), 1.0e-30)) {
        vRs = wR - BR * CR * (SM - uR) / denR;
        double t = SR - uR;
        BRs = CR * (rR * (t * t) - BR2) / denR;
    }
    double vdotBRs = SM * BR + vRs * BRs;
    double enRs = (enR * (SR - uR) - pTR * uR + pTs * SM + BR * (vdotBR - vdotBRs)) / dSR;

    double sqL = sqrt(fmax(rLs, RHO_FLOOR));
    double sqR = sqrt(fmax(rRs, RHO_FLOOR));
    double Bavg = 0.5 * (BL + BR);
    double SLs = SM - fabs(Bavg) / sqL;
    double SRs = SM + fabs(Bavg) / sqR;
    double vss, Bss, enLss, enRss;
    if (fabs(Bavg) > 1.0e-10) {
        double sgn = copysign(1.0, Bavg);
        double dss = sqL + sqR;
        vss = (sqL * vLs + sqR * vRs + sgn * (BRs - BLs)) / dss;
        Bss = (sqL * BRs + sqR * BLs + sgn * sqL * sqR * (vRs - vLs)) / dss;
        double vdotBss = SM * Bavg + vss * Bss;
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

    Flux5 f;
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

    double Smax = fmax(fabs(SL), fabs(SR));
    bool ok = finite_in(f.mass, fmax(fabs(FL.mass), fabs(FR.mass)) + Smax * (rL + rR)) &&
              finite_in(f.mom_n, fmax(fabs(FL.mom_n), fabs(FR.mom_n)) +
                                     Smax * (fabs(rL * uL) + fabs(rR * uR))) &&
              finite_in(f.mom_t, fmax(fabs(FL.mom_t), fabs(FR.mom_t)) +
                                     Smax * (fabs(rL * wL) + fabs(rR * wR))) &&
              finite_in(f.energy, fmax(fabs(FL.energy), fabs(FR.energy)) +
                                      Smax * (fabs(enL) + fabs(enR))) &&
              finite_in(f.bt, fmax(fabs(FL.bt), fabs(FR.bt)) + Smax * (fabs(CL) + fabs(CR)));
    if (!ok) return hll_average(FL, FR, SL, SR, L, R, enL, enR);
    return f;
}

// SHORT: Rusanov and HLLE with the short division and square root (the operand ranges
// stated at rusanov hold); HLLD always takes the full forms.
template <int SOLVER, bool SHORT = false>
__device__ __forceinline__ Flux5 riemann(const FaceState &L, const FaceState &R, double g,
                                         const Recip &rg = Recip{}) {
    if (SOLVER == RUSANOV) return rusanov<SHORT>(L, R, g, rg);
    if (SOLVER == HLLE) return hlle<SHORT>(L, R, g, rg);
    return hlld(L, R, g);
}

// Order-preserving encoding of non-negative doubles for atomicMax on uint64.
__device__ __forceinline__ unsigned long long encode_nonneg(double x) {
    return (unsigned long long)__double_as_longlong(x);
}

}  // namespace chimera
