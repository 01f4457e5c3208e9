/*
 * CPU oracle for the CHIMERA SSP-RK3 stage update -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference Fortran (tjnb0/CHIMERA src/).  It is the
 * checker for the HIP product path and the reported CPU baseline; the product
 * never calls it.  Each routine names the reference lines it restates and keeps
 * the reference's floating-point evaluation order (Fortran evaluates a*b*c as
 * (a*b)*c, x**2 as x*x); build with -ffp-contract=off so no FMA is formed.
 *
 * Pinning (tests/test_oracle_*.py): bit-for-bit against the amdflang build of
 * the reference modules in oracle/_ref (every routine, every solver/limiter,
 * 1- and multi-step runs), against the reference's own Fortran unit-test cases
 * (tests/fortran/test_*.f90), and against closed forms / invariants (flux
 * consistency F(U,U) = F_phys(U), rotation symmetry, div(curl) = 0, mass
 * conservation on periodic grids, limiter tables).
 */
#include "chimera_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/mhd_config.f90:52-61 */
static const double T_OUT = 0.01;
static const double COURANT = 0.3;
static const double P_FLOOR = 1.0e-12;
static const double RHO_FLOOR = 1.0e-12;
static const double E_FLOOR_FRAC = 1.0e-10;

#define IX(i, j) ((size_t)(i) + (size_t)nx * (size_t)(j))
#define PX(i, j) ((size_t)(i) + (size_t)(nx + 2) * (size_t)(j))

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline int periodic(int flag) { return flag == ORC_BC_PERIODIC; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- mhd_bc */

/* corner_val, src/mhd_bc.f90:144-170 (opposite corner index uses nx / ny). */
static double corner_val(int nx, int ny, const double *f, int ix, int iy, int bcx, int bcy) {
    int ixw = nx - 1 - ix, iyw = ny - 1 - iy;
    if (periodic(bcx) && periodic(bcy)) return f[IX(ixw, iyw)];
    if (periodic(bcx)) return f[IX(ixw, iy)];
    if (periodic(bcy)) return f[IX(ix, iyw)];
    return f[IX(ix, iy)];
}

/* fill_ghost_cells + fill_corners, src/mhd_bc.f90:36-141.  The reference sizes
 * everything by the module N (square grids only); the same mapping is applied
 * per axis here so nx != ny also works. */
void orc_fill_ghost_cells(int nx, int ny, const int *bc, const double *f, double *f_pad,
                          const double *f_amb) {
    int xlo = bc[0], xhi = bc[1], ylo = bc[2], yhi = bc[3];
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) f_pad[PX(i + 1, j + 1)] = f[IX(i, j)];
    for (int j = 0; j < ny; ++j) {
        double l, r;
        if (periodic(xlo)) l = f[IX(nx - 1, j)];
        else if ((xlo == ORC_BC_FIXED || xlo == ORC_BC_INFLOW) && f_amb) l = f_amb[IX(0, j)];
        else l = f[IX(0, j)];
        if (periodic(xhi)) r = f[IX(0, j)];
        else if ((xhi == ORC_BC_FIXED || xhi == ORC_BC_INFLOW) && f_amb) r = f_amb[IX(nx - 1, j)];
        else r = f[IX(nx - 1, j)];
        f_pad[PX(0, j + 1)] = l;
        f_pad[PX(nx + 1, j + 1)] = r;
    }
    for (int i = 0; i < nx; ++i) {
        double b, t;
        if (periodic(ylo)) b = f[IX(i, ny - 1)];
        else if ((ylo == ORC_BC_FIXED || ylo == ORC_BC_INFLOW) && f_amb) b = f_amb[IX(i, 0)];
        else b = f[IX(i, 0)];
        if (periodic(yhi)) t = f[IX(i, 0)];
        else if ((yhi == ORC_BC_FIXED || yhi == ORC_BC_INFLOW) && f_amb) t = f_amb[IX(i, ny - 1)];
        else t = f[IX(i, ny - 1)];
        f_pad[PX(i + 1, 0)] = b;
        f_pad[PX(i + 1, ny + 1)] = t;
    }
    f_pad[PX(0, 0)] = corner_val(nx, ny, f, 0, 0, xlo, ylo);
    f_pad[PX(nx + 1, 0)] = corner_val(nx, ny, f, nx - 1, 0, xhi, ylo);
    f_pad[PX(0, ny + 1)] = corner_val(nx, ny, f, 0, ny - 1, xlo, yhi);
    f_pad[PX(nx + 1, ny + 1)] = corner_val(nx, ny, f, nx - 1, ny - 1, xhi, yhi);
}

/* --------------------------------------------------------- mhd_derivatives */

/* compute_gradients, src/mhd_derivatives.f90:15-47 */
void orc_compute_gradients(int nx, int ny, const double *f_pad, double dx, double *f_dx,
                           double *f_dy) {
    double inv_2dx = 1.0 / (2.0 * dx);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            f_dx[IX(i, j)] = (f_pad[PX(i + 2, j + 1)] - f_pad[PX(i, j + 1)]) * inv_2dx;
            f_dy[IX(i, j)] = (f_pad[PX(i + 1, j + 2)] - f_pad[PX(i + 1, j)]) * inv_2dx;
        }
}

/* one limited slope, src/mhd_derivatives.f90:82-97 (x) and 108-121 (y) */
static inline double limited_slope(int limiter, double fm, double f0, double fp, double inv_dx) {
    double dfL = (f0 - fm) * inv_dx;
    double dfR = (fp - f0) * inv_dx;
    double dfRdfL = dfL * dfR;
    double den = dfL + dfR;
    if (dfRdfL <= 0.0 || fabs(den) < 1e-12) return 0.0;
    if (limiter == ORC_LIM_VAN_LEER) return (2.0 * dfRdfL) / den;
    /* MC: sign(1, dfL) * min(2|dfL|, 2|dfR|, |den|/2) */
    double m = dmin(dmin(2.0 * fabs(dfL), 2.0 * fabs(dfR)), 0.5 * fabs(den));
    return copysign(1.0, dfL) * m;
}

/* apply_slope_limiter, src/mhd_derivatives.f90:50-127 */
void orc_apply_slope_limiter(int nx, int ny, int limiter, const double *f_pad, double dx,
                             double *f_dx, double *f_dy) {
    double inv_dx = 1.0 / dx;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            f_dx[IX(i, j)] = limited_slope(limiter, f_pad[PX(i, j + 1)], f_pad[PX(i + 1, j + 1)],
                                           f_pad[PX(i + 2, j + 1)], inv_dx);
            f_dy[IX(i, j)] = limited_slope(limiter, f_pad[PX(i + 1, j)], f_pad[PX(i + 1, j + 1)],
                                           f_pad[PX(i + 1, j + 2)], inv_dx);
        }
}

/* reconstruction, src/mhd_derivatives.f90:130-229.  MOOD checks every face of
 * cell (i,j) -- including the right states pulled from cell i+1 / j+1 -- against
 * the 5-point stencil of (i,j) and resets to f(i,j) (reference behaviour). */
void orc_reconstruction(int nx, int ny, const int *bc, int mood, const double *f,
                        const double *f_dx, const double *f_dy, double dx, double *XL,
                        double *XR, double *YL, double *YR, int rho_or_p, unsigned *mask,
                        int field) {
    size_t n = (size_t)nx * ny;
    double hdx = 0.5 * dx;
    double *tx = (double *)malloc(n * sizeof(double));
    double *ty = (double *)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; ++k) {
        XL[k] = f[k] + hdx * f_dx[k];
        YL[k] = f[k] + hdx * f_dy[k];
        tx[k] = f[k] - hdx * f_dx[k];
        ty[k] = f[k] - hdx * f_dy[k];
    }
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx - 1; ++i) XR[IX(i, j)] = tx[IX(i + 1, j)];
        XR[IX(nx - 1, j)] = periodic(bc[1]) ? tx[IX(0, j)] : f[IX(nx - 1, j)];
    }
    for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx; ++i) YR[IX(i, j)] = ty[IX(i, j + 1)];
    for (int i = 0; i < nx; ++i)
        YR[IX(i, ny - 1)] = periodic(bc[3]) ? ty[IX(i, 0)] : f[IX(i, ny - 1)];
    free(tx);
    free(ty);

    if (mood) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int ip1 = periodic(bc[1]) ? (i + 1) % nx : (i + 1 < nx - 1 ? i + 1 : nx - 1);
                int im1 = periodic(bc[0]) ? (i - 1 + nx) % nx : (i - 1 > 0 ? i - 1 : 0);
                int jp1 = periodic(bc[3]) ? (j + 1) % ny : (j + 1 < ny - 1 ? j + 1 : ny - 1);
                int jm1 = periodic(bc[2]) ? (j - 1 + ny) % ny : (j - 1 > 0 ? j - 1 : 0);
                double c = f[IX(i, j)];
                double a1 = f[IX(ip1, j)], a2 = f[IX(im1, j)], a3 = f[IX(i, jp1)], a4 = f[IX(i, jm1)];
                double lo = dmin(dmin(dmin(dmin(c, a1), a2), a3), a4);
                double hi = dmax(dmax(dmax(dmax(c, a1), a2), a3), a4);
                if (rho_or_p) lo = dmax(1e-9, lo);
                size_t k = IX(i, j);
                unsigned bits = 0;
                if (XL[k] < lo || XL[k] > hi) { XL[k] = c; bits |= 1u; }
                if (XR[k] < lo || XR[k] > hi) { XR[k] = c; bits |= 2u; }
                if (YL[k] < lo || YL[k] > hi) { YL[k] = c; bits |= 4u; }
                if (YR[k] < lo || YR[k] > hi) { YR[k] = c; bits |= 8u; }
                if (mask) mask[k] |= bits << (4 * field);
            }
    }
    if (rho_or_p) {
        const double eps = 1.0e-8;
        for (size_t k = 0; k < n; ++k) {
            XL[k] = dmax(XL[k], eps);
            XR[k] = dmax(XR[k], eps);
            YL[k] = dmax(YL[k], eps);
            YR[k] = dmax(YR[k], eps);
        }
    }
}

/* thermal_pressure_check, src/mhd_derivatives.f90:232-293.  A face whose implied
 * thermal pressure is below P_floor gets all six fields of cell (i,j). */
void orc_thermal_pressure_check(int n, double *const *faces, const double *const *cells,
                                unsigned *mask) {
    /* faces: P[0..3], Bx[4..7], By[8..11], rho[12..15], vx[16..19], vy[20..23] */
    static const int cell_of_group[6] = {3, 4, 5, 0, 1, 2}; /* P,Bx,By,rho,vx,vy */
    for (int side = 0; side < 4; ++side) {
        double *P = faces[side], *Bx = faces[4 + side], *By = faces[8 + side];
        for (int k = 0; k < n; ++k) {
            double p_th = P[k] - 0.5 * (Bx[k] * Bx[k] + By[k] * By[k]);
            if (p_th < P_FLOOR) {
                for (int g = 0; g < 6; ++g) faces[4 * g + side][k] = cells[cell_of_group[g]][k];
                if (mask) mask[k] |= 1u << (24 + side);
            }
        }
    }
}

/* ---------------------------------------------------------------- mhd_flux */

/* rusanov_flux, src/mhd_flux.f90:197-288: F(average state) - C/2 jump, no cap */
static void rusanov(double g, double rL, double rR, double uL, double uR, double wL,
                    double wR, double PL, double PR, double BL, double BR, double CL,
                    double CR, double *out) {
    double halfBL2 = 0.5 * (BL * BL + CL * CL);
    double halfBR2 = 0.5 * (BR * BR + CR * CR);
    double pthL = dmax(PL - halfBL2, P_FLOOR);
    double pthR = dmax(PR - halfBR2, P_FLOOR);
    double enL = pthL / (g - 1.0) + 0.5 * rL * (uL * uL + wL * wL) + halfBL2;
    double enR = pthR / (g - 1.0) + 0.5 * rR * (uR * uR + wR * wR) + halfBR2;
    double rho_avg = 0.5 * (rL + rR);
    double inv = 1.0 / rho_avg;
    double mx = 0.5 * (rL * uL + rR * uR);
    double my = 0.5 * (rL * wL + rR * wR);
    double en = 0.5 * (enL + enR);
    double bx = 0.5 * (BL + BR);
    double by = 0.5 * (CL + CR);
    double Pavg = (g - 1.0) * (en - 0.5 * (mx * mx + my * my) * inv - 0.5 * (bx * bx + by * by))
                  + 0.5 * (bx * bx + by * by);
    double fM = mx;
    double fMx = mx * mx * inv + Pavg - bx * bx;
    double fMy = mx * my * inv - bx * by;
    double fE = (en + Pavg) * mx * inv - bx * (bx * mx + by * my) * inv;
    double fB = (by * mx - bx * my) * inv;
    double cL2 = (g * pthL + 2.0 * halfBL2) / rL;
    double cR2 = (g * pthR + 2.0 * halfBR2) / rR;
    double sL = sqrt(0.5 * (cL2 + fabs(cL2))) + fabs(uL);
    double sR = sqrt(0.5 * (cR2 + fabs(cR2))) + fabs(uR);
    double C = 0.5 * dmax(sL, sR);
    out[0] = fM - C * (rR - rL);
    out[1] = fMx - C * (rR * uR - rL * uL);
    out[2] = fMy - C * (rR * wR - rL * wL);
    out[3] = fE - C * (enR - enL);
    out[4] = fB - C * (CR - CL);
}

/* hlle_flux, src/mhd_flux.f90:291-405 (c_f capped at 100) */
static void hlle(double g, double rL, double rR, double uL, double uR, double wL, double wR,
                 double PL, double PR, double BL, double BR, double CL, double CR,
                 double *out) {
    const double cf_max = 1.0e2;
    double halfBL2 = 0.5 * (BL * BL + CL * CL);
    double halfBR2 = 0.5 * (BR * BR + CR * CR);
    double pthL = dmax(PL - halfBL2, P_FLOOR);
    double pthR = dmax(PR - halfBR2, P_FLOOR);
    double enL = pthL / (g - 1.0) + 0.5 * rL * (uL * uL + wL * wL) + halfBL2;
    double enR = pthR / (g - 1.0) + 0.5 * rR * (uR * uR + wR * wR) + halfBR2;
    double xL = (g * pthL + 2.0 * halfBL2) / rL;
    double xR = (g * pthR + 2.0 * halfBR2) / rR;
    double cfL = dmin(sqrt(0.5 * (xL + fabs(xL))), cf_max);
    double cfR = dmin(sqrt(0.5 * (xR + fabs(xR))), cf_max);
    double SL = dmin(uL - cfL, uR - cfR);
    double SR = dmax(uL + cfL, uR + cfR);
    double FL[5], FR[5];
    FL[0] = rL * uL;
    FL[1] = rL * uL * uL + PL - BL * BL;
    FL[2] = rL * uL * wL - BL * CL;
    FL[3] = (enL + PL) * uL - BL * (BL * uL + CL * wL);
    FL[4] = CL * uL - BL * wL;
    FR[0] = rR * uR;
    FR[1] = rR * uR * uR + PR - BR * BR;
    FR[2] = rR * uR * wR - BR * CR;
    FR[3] = (enR + PR) * uR - BR * (BR * uR + CR * wR);
    FR[4] = CR * uR - BR * wR;
    double dU[5] = {rR - rL, rR * uR - rL * uL, rR * wR - rL * wL, enR - enL, CR - CL};
    double inv = 1.0 / dmax(SR - SL, 1.0e-12);
    for (int c = 0; c < 5; ++c) out[c] = (SR * FL[c] - SL * FR[c] + SL * SR * dU[c]) * inv;
    if (SL >= 0.0)
        for (int c = 0; c < 5; ++c) out[c] = FL[c];
    if (SR <= 0.0)
        for (int c = 0; c < 5; ++c) out[c] = FR[c];
}

/* hlld_flux, src/mhd_flux.f90:408-779 (Miyoshi & Kusano 2005 with the
 * reference's cap 50, catches 1-3, degenerate branches and catch-all) */
static void hlld(double g, double rL, double rR, double uL, double uR, double wL, double wR,
                 double pTL, double pTR, double BL, double BR, double CL, double CR,
                 double *out) {
    const double EPS_FAN = 1.0e-6, FAC_BND = 8.0, cf_max = 50.0;
    double halfBL2 = 0.5 * (BL * BL + CL * CL);
    double halfBR2 = 0.5 * (BR * BR + CR * CR);
    double pthL = dmax(pTL - halfBL2, P_FLOOR);
    double pthR = dmax(pTR - halfBR2, P_FLOOR);
    double enL = pthL / (g - 1.0) + 0.5 * rL * (uL * uL + wL * wL) + halfBL2;
    double enR = pthR / (g - 1.0) + 0.5 * rR * (uR * uR + wR * wR) + halfBR2;

    double vdotBL = uL * BL + wL * CL;
    double FL[5], FR[5];
    FL[0] = rL * uL;
    FL[1] = rL * (uL * uL) + pTL - BL * BL;
    FL[2] = rL * uL * wL - BL * CL;
    FL[3] = (enL + pTL) * uL - BL * vdotBL;
    FL[4] = CL * uL - BL * wL;
    double vdotBR = uR * BR + wR * CR;
    FR[0] = rR * uR;
    FR[1] = rR * (uR * uR) + pTR - BR * BR;
    FR[2] = rR * uR * wR - BR * CR;
    FR[3] = (enR + pTR) * uR - BR * vdotBR;
    FR[4] = CR * uR - BR * wR;

    double caL = BL * BL / rL, caR = BR * BR / rR;
    double csL2 = g * pthL / rL, csR2 = g * pthR / rR;
    double b2L = (BL * BL + CL * CL) / rL, b2R = (BR * BR + CR * CR) / rR;
    double tmpL = csL2 + b2L, tmpR = csR2 + b2R;
    double cfL = sqrt(0.5 * (tmpL + sqrt(dmax(tmpL * tmpL - 4.0 * csL2 * caL, 0.0))));
    double cfR = sqrt(0.5 * (tmpR + sqrt(dmax(tmpR * tmpR - 4.0 * csR2 * caR, 0.0))));
    cfL = dmin(cfL, cf_max);
    cfR = dmin(cfR, cf_max);
    double SL = dmin(uL - cfL, uR - cfR);
    double SR = dmax(uL + cfL, uR + cfR);

    if (cfL >= cf_max || cfR >= cf_max) { /* catch 1 */
        const double *F = (cfL <= cfR) ? FL : FR;
        for (int c = 0; c < 5; ++c) out[c] = F[c];
        return;
    }
    if (SL >= 0.0) { for (int c = 0; c < 5; ++c) out[c] = FL[c]; return; }
    if (SR <= 0.0) { for (int c = 0; c < 5; ++c) out[c] = FR[c]; return; }

    double denMD = rR * (SR - uR) - rL * (SL - uL);
    if (fabs(denMD) < 1.0e-15) denMD = copysign(1.0e-15, denMD);
    double SM = (rR * uR * (SR - uR) - rL * uL * (SL - uL) - (pTR - pTL)) / denMD;
    double pTs = pTL + rL * (SL - uL) * (SM - uL);

    double dU[5] = {rR - rL, rR * uR - rL * uL, rR * wR - rL * wL, enR - enL, CR - CL};
    if (SM <= SL + EPS_FAN * (SR - SL) || SM >= SR - EPS_FAN * (SR - SL)) { /* catch 2 */
        double inv = 1.0 / (SR - SL);
        for (int c = 0; c < 5; ++c) out[c] = (SR * FL[c] - SL * FR[c] + SL * SR * dU[c]) * inv;
        return;
    }
    pTs = dmax(pTs, P_FLOOR); /* catch 3 */

    double rLs = rL * (SL - uL) / (SL - SM);
    double denL = rL * (SL - uL) * (SL - SM) - BL * BL;
    double vLs, BLs;
    if (fabs(denL) > 1.0e-3 * dmax(dmax(fabs(rL * (SL - uL) * (SL - SM)), BL * BL), 1.0e-30)) {
        vLs = wL - BL * CL * (SM - uL) / denL;
        double t = SL - uL;
        BLs = CL * (rL * (t * t) - BL * BL) / denL;
    } else {
        vLs = wL;
        BLs = CL;
    }
    double vdotBLs = SM * BL + vLs * BLs;
    double enLs = (enL * (SL - uL) - pTL * uL + pTs * SM + BL * (vdotBL - vdotBLs)) / (SL - SM);

    double rRs = rR * (SR - uR) / (SR - SM);
    double denR = rR * (SR - uR) * (SR - SM) - BR * BR;
    double vRs, BRs;
    if (fabs(denR) > 1.0e-3 * dmax(dmax(fabs(rR * (SR - uR) * (SR - SM)), BR * BR), 1.0e-30)) {
        vRs = wR - BR * CR * (SM - uR) / denR;
        double t = SR - uR;
        BRs = CR * (rR * (t * t) - BR * BR) / denR;
    } else {
        vRs = wR;
        BRs = CR;
    }
    double vdotBRs = SM * BR + vRs * BRs;
    double enRs = (enR * (SR - uR) - pTR * uR + pTs * SM + BR * (vdotBR - vdotBRs)) / (SR - SM);

    double sqL = sqrt(dmax(rLs, RHO_FLOOR));
    double sqR = sqrt(dmax(rRs, RHO_FLOOR));
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

    double f[5];
    if (SM >= 0.0) {
        f[0] = FL[0] + SL * (rLs - rL);
        f[1] = FL[1] + SL * (rLs * SM - rL * uL);
        f[2] = FL[2] + SL * (rLs * vLs - rL * wL);
        f[3] = FL[3] + SL * (enLs - enL);
        f[4] = FL[4] + SL * (BLs - CL);
        if (!(SLs >= 0.0)) {
            f[2] = f[2] + SLs * rLs * (vss - vLs);
            f[3] = f[3] + SLs * (enLss - enLs);
            f[4] = f[4] + SLs * (Bss - BLs);
        }
    } else {
        f[0] = FR[0] + SR * (rRs - rR);
        f[1] = FR[1] + SR * (rRs * SM - rR * uR);
        f[2] = FR[2] + SR * (rRs * vRs - rR * wR);
        f[3] = FR[3] + SR * (enRs - enR);
        f[4] = FR[4] + SR * (BRs - CR);
        if (!(SRs <= 0.0)) {
            f[2] = f[2] + SRs * rRs * (vss - vRs);
            f[3] = f[3] + SRs * (enRss - enRs);
            f[4] = f[4] + SRs * (Bss - BRs);
        }
    }

    /* catch-all, src/mhd_flux.f90:747-768 */
    double Smax = dmax(fabs(SL), fabs(SR));
    double bnd[5];
    bnd[0] = dmax(fabs(FL[0]), fabs(FR[0])) + Smax * (rL + rR);
    bnd[1] = dmax(fabs(FL[1]), fabs(FR[1])) + Smax * (fabs(rL * uL) + fabs(rR * uR));
    bnd[2] = dmax(fabs(FL[2]), fabs(FR[2])) + Smax * (fabs(rL * wL) + fabs(rR * wR));
    bnd[3] = dmax(fabs(FL[3]), fabs(FR[3])) + Smax * (fabs(enL) + fabs(enR));
    bnd[4] = dmax(fabs(FL[4]), fabs(FR[4])) + Smax * (fabs(CL) + fabs(CR));
    int use_hll = 0;
    for (int c = 0; c < 5; ++c)
        if (!isfinite(f[c]) || fabs(f[c]) > FAC_BND * bnd[c]) use_hll = 1;
    if (use_hll) {
        double inv = 1.0 / (SR - SL);
        for (int c = 0; c < 5; ++c) f[c] = (SR * FL[c] - SL * FR[c] + SL * SR * dU[c]) * inv;
    }
    for (int c = 0; c < 5; ++c) out[c] = f[c];
}

/* compute_fluxes, src/mhd_flux.f90:139-191 */
void orc_compute_fluxes(int n, int solver, double gamma, const double *const *in,
                        double *const *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) {
        double o[5];
        double a[12];
        for (int q = 0; q < 12; ++q) a[q] = in[q][k];
        if (solver == ORC_RUSANOV)
            rusanov(gamma, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], o);
        else if (solver == ORC_HLLE)
            hlle(gamma, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], o);
        else
            hlld(gamma, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11], o);
        for (int c = 0; c < 5; ++c) out[c][k] = o[c];
    }
}

/* update_conserved, src/mhd_flux.f90:20-76 (conserved = cell integrals: dt*dx) */
void orc_update_conserved(int nx, int ny, const int *bc, double *F, const double *fX,
                          const double *fY, double dx, double dt) {
    double dtdx = dt * dx;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double v = F[IX(i, j)];
            if (i > 0) v = v - dtdx * fX[IX(i, j)] + dtdx * fX[IX(i - 1, j)];
            else if (periodic(bc[0])) v = v - dtdx * fX[IX(0, j)] + dtdx * fX[IX(nx - 1, j)];
            else v = v - dtdx * fX[IX(0, j)];
            F[IX(i, j)] = v;
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double v = F[IX(i, j)];
            if (j > 0) v = v - dtdx * fY[IX(i, j)] + dtdx * fY[IX(i, j - 1)];
            else if (periodic(bc[2])) v = v - dtdx * fY[IX(i, 0)] + dtdx * fY[IX(i, ny - 1)];
            else v = v - dtdx * fY[IX(i, 0)];
            F[IX(i, j)] = v;
        }
}

/* compute_curl_2d, src/mhd_field_ops.f90:15-57 (backward differences; a
 * non-periodic lo side copies the first interior row/column) */
void orc_compute_curl_2d(int nx, int ny, const int *bc, const double *Az, double dx,
                         double *bx, double *by) {
    double inv_dx = 1.0 / dx;
    for (int j = 1; j < ny; ++j)
        for (int i = 0; i < nx; ++i) bx[IX(i, j)] = (Az[IX(i, j)] - Az[IX(i, j - 1)]) * inv_dx;
    for (int i = 0; i < nx; ++i)
        bx[IX(i, 0)] = periodic(bc[2]) ? (Az[IX(i, 0)] - Az[IX(i, ny - 1)]) * inv_dx : bx[IX(i, 1)];
    for (int j = 0; j < ny; ++j)
        for (int i = 1; i < nx; ++i) by[IX(i, j)] = -((Az[IX(i, j)] - Az[IX(i - 1, j)]) * inv_dx);
    for (int j = 0; j < ny; ++j)
        by[IX(0, j)] = periodic(bc[0]) ? -((Az[IX(0, j)] - Az[IX(nx - 1, j)]) * inv_dx) : by[IX(1, j)];
}

/* constrained_transport, src/mhd_flux.f90:79-136 */
void orc_constrained_transport(int nx, int ny, const int *bc, double *bx, double *by,
                               const double *fByX, const double *fBxY, double dx, double dt) {
    size_t n = (size_t)nx * ny;
    double *negEz = (double *)malloc(n * sizeof(double));
    double *dbx = (double *)malloc(n * sizeof(double));
    double *dby = (double *)malloc(n * sizeof(double));
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double up = (j < ny - 1) ? fByX[IX(i, j + 1)]
                                     : (periodic(bc[3]) ? fByX[IX(i, 0)] : fByX[IX(i, ny - 1)]);
            double right = (i < nx - 1) ? fBxY[IX(i + 1, j)]
                                        : (periodic(bc[1]) ? fBxY[IX(0, j)] : fBxY[IX(nx - 1, j)]);
            double Ez = 0.25 * (-fByX[IX(i, j)] - up + fBxY[IX(i, j)] + right);
            negEz[IX(i, j)] = -Ez;
        }
    orc_compute_curl_2d(nx, ny, bc, negEz, dx, dbx, dby);
    for (size_t k = 0; k < n; ++k) {
        bx[k] = bx[k] + dt * dbx[k];
        by[k] = by[k] + dt * dby[k];
    }
    free(negEz);
    free(dbx);
    free(dby);
}

/* compute_divB, src/mhd_field_ops.f90:60-117 */
void orc_compute_divB(int nx, int ny, const int *bc, const double *bx, const double *by,
                      double dx, double *divB) {
    double inv_dx = 1.0 / dx;
    int px = periodic(bc[0]), py = periodic(bc[2]);
    for (int j = 1; j < ny; ++j)
        for (int i = 1; i < nx; ++i)
            divB[IX(i, j)] = (bx[IX(i, j)] - bx[IX(i - 1, j)] + by[IX(i, j)] - by[IX(i, j - 1)]) * inv_dx;
    for (int j = 1; j < ny; ++j)
        divB[IX(0, j)] = px ? (bx[IX(0, j)] - bx[IX(nx - 1, j)] + by[IX(0, j)] - by[IX(0, j - 1)]) * inv_dx
                            : (by[IX(0, j)] - by[IX(0, j - 1)]) * inv_dx;
    for (int i = 1; i < nx; ++i)
        divB[IX(i, 0)] = py ? (bx[IX(i, 0)] - bx[IX(i - 1, 0)] + by[IX(i, 0)] - by[IX(i, ny - 1)]) * inv_dx
                            : (bx[IX(i, 0)] - bx[IX(i - 1, 0)]) * inv_dx;
    if (px && py)
        divB[0] = (bx[0] - bx[IX(nx - 1, 0)] + by[0] - by[IX(0, ny - 1)]) * inv_dx;
    else if (px)
        divB[0] = (bx[0] - bx[IX(nx - 1, 0)]) * inv_dx;
    else if (py)
        divB[0] = (by[0] - by[IX(0, ny - 1)]) * inv_dx;
    else
        divB[0] = 0.0;
}

/* average_face_to_cell_B, src/mhd_field_ops.f90:120-156 */
void orc_average_face_to_cell_B(int nx, int ny, const int *bc, const double *b_x,
                                const double *b_y, double *Bx, double *By) {
    for (int j = 0; j < ny; ++j) {
        for (int i = 1; i < nx; ++i) Bx[IX(i, j)] = 0.5 * (b_x[IX(i, j)] + b_x[IX(i - 1, j)]);
        Bx[IX(0, j)] = periodic(bc[0]) ? 0.5 * (b_x[IX(0, j)] + b_x[IX(nx - 1, j)]) : b_x[IX(0, j)];
    }
    for (int j = 1; j < ny; ++j)
        for (int i = 0; i < nx; ++i) By[IX(i, j)] = 0.5 * (b_y[IX(i, j)] + b_y[IX(i, j - 1)]);
    for (int i = 0; i < nx; ++i)
        By[IX(i, 0)] = periodic(bc[2]) ? 0.5 * (b_y[IX(i, 0)] + b_y[IX(i, ny - 1)]) : b_y[IX(i, 0)];
}

/* ------------------------------------------------------- mhd_change_states */

/* get_conserved, src/mhd_change_states.f90:16-58 */
void orc_get_conserved(int n, const double *rho, const double *vx, const double *vy,
                       const double *P, const double *Bx, const double *By, double g,
                       double vol, double *Mass, double *Momx, double *Momy, double *Energy) {
    for (int k = 0; k < n; ++k) {
        double r = dmax(rho[k], RHO_FLOOR);
        double p = dmax(P[k], P_FLOOR);
        double m = r * vol;
        double halfB2 = 0.5 * (Bx[k] * Bx[k] + By[k] * By[k]);
        Mass[k] = m;
        Momx[k] = m * vx[k];
        Momy[k] = m * vy[k];
        Energy[k] = ((p - halfB2) / (g - 1.0) + 0.5 * r * (vx[k] * vx[k] + vy[k] * vy[k]) + halfB2) * vol;
    }
}

/* get_primitive, src/mhd_change_states.f90:61-113 (absolute then proportional
 * pressure floor; P is total pressure) */
void orc_get_primitive(int n, const double *Mass, const double *Momx, const double *Momy,
                       const double *Energy, const double *Bx, const double *By, double g,
                       double vol, double *rho, double *vx, double *vy, double *P) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) {
        double r = dmax(Mass[k] / vol, RHO_FLOOR);
        double u = Momx[k] / (r * vol);
        double w = Momy[k] / (r * vol);
        double halfB2 = 0.5 * (Bx[k] * Bx[k] + By[k] * By[k]);
        double p = (Energy[k] / vol - 0.5 * r * (u * u + w * w) - halfB2) * (g - 1.0) + halfB2;
        p = dmax(p, P_FLOOR);
        p = dmax(p, halfB2 + E_FLOOR_FRAC * (0.5 * r * (u * u + w * w) + halfB2));
        rho[k] = r;
        vx[k] = u;
        vy[k] = w;
        P[k] = p;
    }
}

/* apply_conserved_floors, src/mhd_change_states.f90:116-166 */
void orc_apply_conserved_floors(int n, double *Mass, double *Momx, double *Momy,
                                double *Energy, const double *Bx, const double *By, double vol,
                                double g) {
    const double V_MAX = 50.0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; ++k) {
        double halfB2 = 0.5 * (Bx[k] * Bx[k] + By[k] * By[k]);
        if (Mass[k] < RHO_FLOOR * vol) {
            Momx[k] = 0.0;
            Momy[k] = 0.0;
            Mass[k] = RHO_FLOOR * vol;
            Energy[k] = dmax(Energy[k], (P_FLOOR / (g - 1.0) + halfB2) * vol);
        }
        double v_mag = sqrt(Momx[k] * Momx[k] + Momy[k] * Momy[k]) / Mass[k];
        if (v_mag > V_MAX) {
            double scale = V_MAX * Mass[k] / sqrt(Momx[k] * Momx[k] + Momy[k] * Momy[k]);
            Momx[k] = Momx[k] * scale;
            Momy[k] = Momy[k] * scale;
        }
    }
}

/* signal speed cf + |v| of main.f90:72-79, max-reduced.  minval(dx/s) equals
 * dx/maxval(s) exactly because correctly rounded division is monotone. */
double orc_cfl_max_speed(int n, double g, const double *rho, const double *vx,
                         const double *vy, const double *P, const double *Bx, const double *By) {
    double smax = 0.0;
#pragma omp parallel for reduction(max : smax) schedule(static)
    for (int k = 0; k < n; ++k) {
        double inv_rho = 1.0 / rho[k];
        double B2 = Bx[k] * Bx[k] + By[k] * By[k];
        double c0_sq = g * (P[k] - 0.5 * B2) * inv_rho;
        double ca_sq = B2 * inv_rho;
        double cf = sqrt(c0_sq + ca_sq);
        if (rho[k] < 1.0e-3) cf = dmin(cf, 50.0);
        double s = cf + sqrt(vx[k] * vx[k] + vy[k] * vy[k]);
        if (s > smax) smax = s;
    }
    return smax;
}

/* ------------------------------------------------------------ main.f90 */

/* compute_stage_fluxes, src/main.f90:346-403 */
int orc_stage_fluxes(int nx, int ny, const int *bc, const int *scheme, double gamma,
                     double dx, const double *const *prims, double *const *flux,
                     unsigned *mask) {
    size_t n = (size_t)nx * ny;
    size_t np = (size_t)(nx + 2) * (ny + 2);
    double *pad = (double *)malloc(np * sizeof(double));
    double *fdx = (double *)malloc(n * sizeof(double));
    double *fdy = (double *)malloc(n * sizeof(double));
    double *faces = (double *)malloc(24 * n * sizeof(double));
    if (!pad || !fdx || !fdy || !faces) {
        free(pad); free(fdx); free(fdy); free(faces);
        return -1;
    }
    if (mask) memset(mask, 0, n * sizeof(unsigned));
    /* face storage per field f (rho, vx, vy, P, Bx, By): XL, XR, YL, YR */
    for (int f = 0; f < 6; ++f) {
        double *F4 = faces + (size_t)4 * f * n;
        orc_fill_ghost_cells(nx, ny, bc, prims[f], pad, NULL);
        if (scheme[2]) orc_apply_slope_limiter(nx, ny, scheme[1], pad, dx, fdx, fdy);
        else orc_compute_gradients(nx, ny, pad, dx, fdx, fdy);
        orc_reconstruction(nx, ny, bc, scheme[3], prims[f], fdx, fdy, dx, F4, F4 + n, F4 + 2 * n,
                           F4 + 3 * n, (f == 0 || f == 3), mask, f);
    }
    /* thermal check argument order: P, Bx, By, rho, vx, vy (4 faces each) */
    static const int grp[6] = {3, 4, 5, 0, 1, 2};
    double *fp[24];
    for (int g = 0; g < 6; ++g)
        for (int s = 0; s < 4; ++s) fp[4 * g + s] = faces + (size_t)(4 * grp[g] + s) * n;
    orc_thermal_pressure_check((int)n, fp, prims, mask);

#define FACE(f, s) (faces + (size_t)(4 * (f) + (s)) * n)
    /* x faces: normal (vx, Bx), tangential (vy, By) */
    const double *inx[12] = {FACE(0, 0), FACE(0, 1), FACE(1, 0), FACE(1, 1), FACE(2, 0), FACE(2, 1),
                             FACE(3, 0), FACE(3, 1), FACE(4, 0), FACE(4, 1), FACE(5, 0), FACE(5, 1)};
    double *outx[5] = {flux[0], flux[2], flux[4], flux[6], flux[8]};
    orc_compute_fluxes((int)n, scheme[0], gamma, inx, outx);
    /* y faces, rotated: normal (vy, By), tangential (vx, Bx) */
    const double *iny[12] = {FACE(0, 2), FACE(0, 3), FACE(2, 2), FACE(2, 3), FACE(1, 2), FACE(1, 3),
                             FACE(3, 2), FACE(3, 3), FACE(5, 2), FACE(5, 3), FACE(4, 2), FACE(4, 3)};
    double *outy[5] = {flux[1], flux[5], flux[3], flux[7], flux[9]};
    orc_compute_fluxes((int)n, scheme[0], gamma, iny, outy);
#undef FACE
    free(pad);
    free(fdx);
    free(fdy);
    free(faces);
    return 0;
}

/* outflow boundary flux correction on row/column 1, src/main.f90:116-148 */
static void outflow_correction(int nx, int ny, const int *bc, double gamma, double dt,
                               double dx, double *const *st, const double *const *pr) {
    double *Mass = st[0], *Momx = st[1], *Momy = st[2], *Energy = st[3];
    const double *rho = pr[0], *vx = pr[1], *vy = pr[2], *P = pr[3], *Bx = pr[4], *By = pr[5];
    if (!periodic(bc[0])) {
        for (int j = 0; j < ny; ++j) {
            size_t k = IX(0, j);
            double h = 0.5 * (Bx[k] * Bx[k] + By[k] * By[k]);
            double en = (P[k] - h) / (gamma - 1.0) + 0.5 * rho[k] * (vx[k] * vx[k] + vy[k] * vy[k]) + h;
            Mass[k] = Mass[k] + dt * dx * rho[k] * vx[k];
            Momx[k] = Momx[k] + dt * dx * (rho[k] * (vx[k] * vx[k]) + P[k] - Bx[k] * Bx[k]);
            Momy[k] = Momy[k] + dt * dx * (rho[k] * vx[k] * vy[k] - Bx[k] * By[k]);
            Energy[k] = Energy[k] + dt * dx * ((en + P[k]) * vx[k] - Bx[k] * (Bx[k] * vx[k] + By[k] * vy[k]));
        }
    }
    if (!periodic(bc[2])) {
        for (int i = 0; i < nx; ++i) {
            size_t k = IX(i, 0);
            double h = 0.5 * (Bx[k] * Bx[k] + By[k] * By[k]);
            double en = (P[k] - h) / (gamma - 1.0) + 0.5 * rho[k] * (vx[k] * vx[k] + vy[k] * vy[k]) + h;
            Mass[k] = Mass[k] + dt * dx * rho[k] * vy[k];
            Momy[k] = Momy[k] + dt * dx * (rho[k] * (vy[k] * vy[k]) + P[k] - By[k] * By[k]);
            Momx[k] = Momx[k] + dt * dx * (rho[k] * vy[k] * vx[k] - By[k] * Bx[k]);
            Energy[k] = Energy[k] + dt * dx * ((en + P[k]) * vy[k] - By[k] * (By[k] * vy[k] + Bx[k] * vx[k]));
        }
    }
}

/* main loop, src/main.f90:65-303 */
int orc_run(int nx, int ny, const int *bc, const int *scheme, double gamma, double dx,
            double *const *state, int nsteps, int dt_mode, const double *dt_in,
            double *dt_out, double *t_io, int *out_count_io, double t_end, double *divb_out,
            unsigned *mask_stage1) {
    if (nx < 2 || ny < 2 || nsteps < 0) return -1;
    size_t n = (size_t)nx * ny;
    double vol = dx * dx;
    double *buf = (double *)malloc(23 * n * sizeof(double));
    if (!buf) return -2;
    double *prim[6], *u0[6], *flux[10];
    for (int q = 0; q < 6; ++q) prim[q] = buf + q * n;
    for (int q = 0; q < 6; ++q) u0[q] = buf + (6 + q) * n;
    for (int q = 0; q < 10; ++q) flux[q] = buf + (12 + q) * n;
    double *divB = buf + 22 * n;
    double *rho = prim[0], *vx = prim[1], *vy = prim[2], *P = prim[3], *Bx = prim[4], *By = prim[5];
    double *Mass = state[0], *Momx = state[1], *Momy = state[2], *Energy = state[3];
    double *b_x = state[4], *b_y = state[5];
    double t = t_io ? *t_io : 0.0;
    int out_count = out_count_io ? *out_count_io : 1;
    int s;
    for (s = 0; s < nsteps; ++s) {
        if (t_end > 0.0 && !(t < t_end)) break;
        orc_average_face_to_cell_B(nx, ny, bc, b_x, b_y, Bx, By);
        orc_get_primitive((int)n, Mass, Momx, Momy, Energy, Bx, By, gamma, vol, rho, vx, vy, P);
        double dt;
        if (dt_mode == ORC_DT_PRESCRIBED) {
            dt = dt_in[s];
        } else {
            double smax = orc_cfl_max_speed((int)n, gamma, rho, vx, vy, P, Bx, By);
            dt = COURANT * (dx / smax);
            if (dt_mode == ORC_DT_REFERENCE && t + dt > out_count * T_OUT) dt = out_count * T_OUT - t;
            dt = dmax(dt, (double)1e-12f); /* main.f90:83, single-precision literal */
        }
        if (dt_out) dt_out[s] = dt;
        for (int q = 0; q < 6; ++q) memcpy(u0[q], state[q], n * sizeof(double));
        for (int k = 1; k <= 3; ++k) {
            if (k > 1) {
                orc_average_face_to_cell_B(nx, ny, bc, b_x, b_y, Bx, By);
                orc_get_primitive((int)n, Mass, Momx, Momy, Energy, Bx, By, gamma, vol, rho, vx, vy, P);
            }
            if (orc_stage_fluxes(nx, ny, bc, scheme, gamma, dx, (const double *const *)prim, flux,
                                 (k == 1 && s == 0) ? mask_stage1 : NULL) != 0) {
                free(buf);
                return -2;
            }
            orc_update_conserved(nx, ny, bc, Mass, flux[0], flux[1], dx, dt);
            orc_update_conserved(nx, ny, bc, Momx, flux[2], flux[3], dx, dt);
            orc_update_conserved(nx, ny, bc, Momy, flux[4], flux[5], dx, dt);
            orc_update_conserved(nx, ny, bc, Energy, flux[6], flux[7], dx, dt);
            orc_apply_conserved_floors((int)n, Mass, Momx, Momy, Energy, Bx, By, vol, gamma);
            orc_constrained_transport(nx, ny, bc, b_x, b_y, flux[8], flux[9], dx, dt);
            outflow_correction(nx, ny, bc, gamma, dt, dx, state, (const double *const *)prim);
            if (k > 1) {
                double c0 = (k == 2) ? 0.75 : 1.0 / 3.0;
                double c1 = (k == 2) ? 0.25 : 2.0 / 3.0;
                for (int q = 0; q < 6; ++q)
                    for (size_t m = 0; m < n; ++m) state[q][m] = c0 * u0[q][m] + c1 * state[q][m];
                orc_apply_conserved_floors((int)n, Mass, Momx, Momy, Energy, Bx, By, vol, gamma);
            }
        }
        t = t + dt;
        if (divb_out) {
            orc_compute_divB(nx, ny, bc, b_x, b_y, dx, divB);
            double mx = 0.0, sm = 0.0;
            for (size_t m = 0; m < n; ++m) {
                double a = fabs(divB[m]);
                if (a > mx) mx = a;
                sm += a;
            }
            divb_out[2 * s] = mx;
            divb_out[2 * s + 1] = sm;
        }
        if (t >= out_count * T_OUT) out_count = out_count + 1;
    }
    if (t_io) *t_io = t;
    if (out_count_io) *out_count_io = out_count;
    free(buf);
    return s;
}
