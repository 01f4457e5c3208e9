// Discrete adjoint of the SSP-RK3 step (frozen dt): hand-written reverse kernels for
// the reconstruction (limiter, MOOD, face floors, thermal check), the Riemann fluxes,
// the conservative update, constrained transport, the outflow correction, the conserved
// floors and the RK blends.
//
// Per stage, forward kernels af_* recompute the stage from its input exactly as the
// fused kernel does (same operations in the same order) and keep what the reverse
// needs -- primitives, face states after MOOD and the thermal check, branch bits,
// fluxes, the pre-floor and pre-blend states.  Reverse kernels ar_* are written in
// gather form (each thread owns one output location and collects the contributions
// of every forward operation that read it), so there are no atomics and the
// gradient is deterministic.  Branches (limiter cases, MOOD resets, face and
// conserved floors, thermal resets, Riemann-solver regions and catches) are the ones
// the forward pass took; Riemann Jacobians come from forward-mode dual numbers
// through riemann_dual.hpp.
#include <vector>

#include "adjoint_phys.hpp"
#include "common.hpp"
#include "riemann_dual.hpp"

namespace chimera {
namespace {

constexpr int BS = 256;

struct AdjGrid {
    int nx, ny, px, py;
    double dx, vol, gamma;
    int use_lim, mood;
};

// per-stage intermediates (device arrays of n doubles unless noted)
struct StageBufs {
    double *P[6];       // primitives rho, vx, vy, P, Bx, By
    double *face[24];   // field f, side s (XL, XR, YL, YR) at 4*f + s, after thermal check
    unsigned long long *bits;  // MOOD 4f+s | thermal 24+s | face floor 32 + 4k + s (k: rho, P)
    double *FX[5], *FY[5];     // Mass, Momx, Momy, Energy, B flux
    double *Up[4];      // after the flux update (before floors)
    double *U3[4];      // after floors, CT and the outflow correction (before the blend)
};

struct RevBufs {
    double *lUp[4], *lBp[2], *lBc[2], *lPd[6], *lnegEz;
    double *lXL[6], *lXR[6], *lYL[6], *lYR[6], *lPr[6];
    double *lSx[6], *lSy[6], *lP[6];
    double *lBcT[2];
};

__device__ __forceinline__ int wrap_or_clamp(int g, int n, int periodic) {
    if (periodic) return g < 0 ? g + n : (g >= n ? g - n : g);
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

#define CELL_LOOP(k, n) \
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (n); \
         k += (size_t)gridDim.x * blockDim.x)

// ---- forward recompute ------------------------------------------------------------

__global__ void af_prims(AdjGrid G, const double *const *U, StageBufs S) {
    const int nx = G.nx, ny = G.ny;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        int im = i > 0 ? i - 1 : (G.px ? nx - 1 : 0);
        int jm = j > 0 ? j - 1 : (G.py ? ny - 1 : 0);
        double Bx = 0.5 * (U[4][k] + U[4][(size_t)im + (size_t)nx * j]);
        double By = 0.5 * (U[5][k] + U[5][(size_t)i + (size_t)nx * jm]);
        double r, u, w, p;
        cons_to_prim(U[0][k], U[1][k], U[2][k], U[3][k], Bx, By, G.gamma, G.vol, r, u, w, p);
        S.P[0][k] = r;
        S.P[1][k] = u;
        S.P[2][k] = w;
        S.P[3][k] = p;
        S.P[4][k] = Bx;
        S.P[5][k] = By;
    }
}

template <int LIMITER>
__device__ __forceinline__ double slope_at(const double *f, int i, int j, int dir, const AdjGrid &G,
                                           double inv_dx, double inv_2dx) {
    const int nx = G.nx, ny = G.ny;
    double fm, f0, fp;
    f0 = f[(size_t)i + (size_t)nx * j];
    if (dir == 0) {
        fm = f[(size_t)wrap_or_clamp(i - 1, nx, G.px) + (size_t)nx * j];
        fp = f[(size_t)wrap_or_clamp(i + 1, nx, G.px) + (size_t)nx * j];
    } else {
        fm = f[(size_t)i + (size_t)nx * wrap_or_clamp(j - 1, ny, G.py)];
        fp = f[(size_t)i + (size_t)nx * wrap_or_clamp(j + 1, ny, G.py)];
    }
    return G.use_lim ? limited_slope<LIMITER>(fm, f0, fp, inv_dx) : central_slope(fm, fp, inv_2dx);
}

// faces of cell (i,j): XL, XR (from i+1), YL, YR (from j+1); MOOD against the stencil
// of (i,j); rho/P face floor; thermal check against cell (i,j).
template <int LIMITER>
__global__ void af_faces(AdjGrid G, StageBufs S) {
    const int nx = G.nx, ny = G.ny;
    const double inv_dx = 1.0 / G.dx, inv_2dx = 1.0 / (2.0 * G.dx), hdx = 0.5 * G.dx;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        int ip = wrap_or_clamp(i + 1, nx, G.px), im = wrap_or_clamp(i - 1, nx, G.px);
        int jp = wrap_or_clamp(j + 1, ny, G.py), jm = wrap_or_clamp(j - 1, ny, G.py);
        bool lastx = !G.px && i == nx - 1, lasty = !G.py && j == ny - 1;
        unsigned long long bits = 0;
        double F[4][6], C[6];
        for (int f = 0; f < 6; ++f) {
            const double *a = S.P[f];
            double c = a[k];
            C[f] = c;
            double sx = slope_at<LIMITER>(a, i, j, 0, G, inv_dx, inv_2dx);
            double sy = slope_at<LIMITER>(a, i, j, 1, G, inv_dx, inv_2dx);
            double xl = c + hdx * sx;
            double yl = c + hdx * sy;
            double xr = lastx ? c : a[(size_t)ip + (size_t)nx * j] - hdx * slope_at<LIMITER>(a, ip, j, 0, G, inv_dx, inv_2dx);
            double yr = lasty ? c : a[(size_t)i + (size_t)nx * jp] - hdx * slope_at<LIMITER>(a, i, jp, 1, G, inv_dx, inv_2dx);
            bool rp = (f == 0 || f == 3);
            if (G.mood) {
                double a1 = a[(size_t)ip + (size_t)nx * j], a2 = a[(size_t)im + (size_t)nx * j];
                double a3 = a[(size_t)i + (size_t)nx * jp], a4 = a[(size_t)i + (size_t)nx * jm];
                double lo = fmin(fmin(fmin(fmin(c, a1), a2), a3), a4);
                double hi = fmax(fmax(fmax(fmax(c, a1), a2), a3), a4);
                if (rp) lo = fmax(1e-9, lo);
                if (xl < lo || xl > hi) { xl = c; bits |= 1ull << (4 * f + 0); }
                if (xr < lo || xr > hi) { xr = c; bits |= 1ull << (4 * f + 1); }
                if (yl < lo || yl > hi) { yl = c; bits |= 1ull << (4 * f + 2); }
                if (yr < lo || yr > hi) { yr = c; bits |= 1ull << (4 * f + 3); }
            }
            if (rp) {
                const int kk = f == 0 ? 0 : 1;
                double v[4] = {xl, xr, yl, yr};
                for (int s = 0; s < 4; ++s)
                    if (1.0e-8 > v[s]) bits |= 1ull << (32 + 4 * kk + s);
                xl = fmax(xl, 1.0e-8);
                xr = fmax(xr, 1.0e-8);
                yl = fmax(yl, 1.0e-8);
                yr = fmax(yr, 1.0e-8);
            }
            F[0][f] = xl;
            F[1][f] = xr;
            F[2][f] = yl;
            F[3][f] = yr;
        }
        for (int s = 0; s < 4; ++s) {
            double p_th = F[s][3] - 0.5 * (F[s][4] * F[s][4] + F[s][5] * F[s][5]);
            if (p_th < P_FLOOR) {
                for (int f = 0; f < 6; ++f) F[s][f] = C[f];
                bits |= 1ull << (24 + s);
            }
        }
        for (int f = 0; f < 6; ++f)
            for (int s = 0; s < 4; ++s) S.face[4 * f + s][k] = F[s][f];
        S.bits[k] = bits;
    }
}

template <int SOLVER>
__global__ void af_fluxes(AdjGrid G, StageBufs S) {
    CELL_LOOP(k, (size_t)G.nx * G.ny) {
        FaceState a{S.face[0][k], S.face[4][k], S.face[8][k], S.face[12][k], S.face[16][k], S.face[20][k]};
        FaceState b{S.face[1][k], S.face[5][k], S.face[9][k], S.face[13][k], S.face[17][k], S.face[21][k]};
        Flux5 f = riemann<SOLVER>(a, b, G.gamma);
        S.FX[0][k] = f.mass;
        S.FX[1][k] = f.mom_n;
        S.FX[2][k] = f.mom_t;
        S.FX[3][k] = f.energy;
        S.FX[4][k] = f.bt;
        FaceState c{S.face[2][k], S.face[10][k], S.face[6][k], S.face[14][k], S.face[22][k], S.face[18][k]};
        FaceState d{S.face[3][k], S.face[11][k], S.face[7][k], S.face[15][k], S.face[23][k], S.face[19][k]};
        Flux5 h = riemann<SOLVER>(c, d, G.gamma);
        S.FY[0][k] = h.mass;
        S.FY[1][k] = h.mom_t;
        S.FY[2][k] = h.mom_n;
        S.FY[3][k] = h.energy;
        S.FY[4][k] = h.bt;
    }
}

// -Ez at corner (a, b) from the B fluxes (src/mhd_flux.f90:108-129)
__device__ __forceinline__ double neg_ez_at(const StageBufs &S, const AdjGrid &G, int a, int b) {
    const int nx = G.nx, ny = G.ny;
    size_t k = (size_t)a + (size_t)nx * b;
    double f0 = S.FX[4][k];
    double up = (b < ny - 1) ? S.FX[4][k + nx] : (G.py ? S.FX[4][a] : f0);
    double e0 = S.FY[4][k];
    double rt = (a < nx - 1) ? S.FY[4][k + 1] : (G.px ? S.FY[4][(size_t)nx * b] : e0);
    double Ez = 0.25 * (-f0 - up + e0 + rt);
    return -Ez;
}

__global__ void af_update(AdjGrid G, const double *const *U, const double *const *U0,
                          double *const *Uout, StageBufs S, int stage, double dt) {
    const int nx = G.nx, ny = G.ny;
    const double dtdx = dt * G.dx, inv_dx = 1.0 / G.dx, g = G.gamma, vol = G.vol;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        bool xlo = i == 0 && !G.px, ylo = j == 0 && !G.py;
        size_t kl = i > 0 ? k - 1 : k + nx - 1;
        size_t kb = j > 0 ? k - nx : k + (size_t)nx * (ny - 1);
        double V[4];
        for (int q = 0; q < 4; ++q) {
            double v = U[q][k];
            if (xlo) v = v - dtdx * S.FX[q][k];
            else v = v - dtdx * S.FX[q][k] + dtdx * S.FX[q][kl];
            if (ylo) v = v - dtdx * S.FY[q][k];
            else v = v - dtdx * S.FY[q][k] + dtdx * S.FY[q][kb];
            V[q] = v;
            S.Up[q][k] = v;
        }
        double Bxc = S.P[4][k], Byc = S.P[5][k];
        conserved_floors(V[0], V[1], V[2], V[3], Bxc, Byc, vol, g);
        double dbx, dby;
        if (!ylo) dbx = (neg_ez_at(S, G, i, j) - neg_ez_at(S, G, i, j > 0 ? j - 1 : ny - 1)) * inv_dx;
        else dbx = (neg_ez_at(S, G, i, 1) - neg_ez_at(S, G, i, 0)) * inv_dx;
        if (!xlo) dby = -((neg_ez_at(S, G, i, j) - neg_ez_at(S, G, i > 0 ? i - 1 : nx - 1, j)) * inv_dx);
        else dby = -((neg_ez_at(S, G, 1, j) - neg_ez_at(S, G, 0, j)) * inv_dx);
        double bxn = U[4][k] + dt * dbx;
        double byn = U[5][k] + dt * dby;
        if (xlo || ylo) {
            double rho = S.P[0][k], vx = S.P[1][k], vy = S.P[2][k], P = S.P[3][k];
            double h = 0.5 * (Bxc * Bxc + Byc * Byc);
            double en = (P - h) / (g - 1.0) + 0.5 * rho * (vx * vx + vy * vy) + h;
            if (xlo) {
                V[0] = V[0] + dtdx * rho * vx;
                V[1] = V[1] + dtdx * (rho * (vx * vx) + P - Bxc * Bxc);
                V[2] = V[2] + dtdx * (rho * vx * vy - Bxc * Byc);
                V[3] = V[3] + dtdx * ((en + P) * vx - Bxc * (Bxc * vx + Byc * vy));
            }
            if (ylo) {
                V[0] = V[0] + dtdx * rho * vy;
                V[2] = V[2] + dtdx * (rho * (vy * vy) + P - Byc * Byc);
                V[1] = V[1] + dtdx * (rho * vy * vx - Byc * Bxc);
                V[3] = V[3] + dtdx * ((en + P) * vy - Byc * (Byc * vy + Bxc * vx));
            }
        }
        for (int q = 0; q < 4; ++q) S.U3[q][k] = V[q];
        if (stage > 1) {
            const double c0 = (stage == 2) ? 0.75 : 1.0 / 3.0;
            const double c1 = (stage == 2) ? 0.25 : 2.0 / 3.0;
            for (int q = 0; q < 4; ++q) V[q] = c0 * U0[q][k] + c1 * V[q];
            bxn = c0 * U0[4][k] + c1 * bxn;
            byn = c0 * U0[5][k] + c1 * byn;
            conserved_floors(V[0], V[1], V[2], V[3], Bxc, Byc, vol, g);
        }
        for (int q = 0; q < 4; ++q) Uout[q][k] = V[q];
        Uout[4][k] = bxn;
        Uout[5][k] = byn;
    }
}

// ---- reverse -------------------------------------------------------------------

// blend + floors (stages 2, 3), outflow correction, floors after the update.
__global__ void ar_cell_tail(AdjGrid G, const double *const *U0, StageBufs S, RevBufs R,
                             const double *const *lout, double *const *lin, double *const *lU0,
                             int stage, double dt) {
    const int nx = G.nx, ny = G.ny;
    const double g = G.gamma, vol = G.vol;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        double Bx = S.P[4][k], By = S.P[5][k];
        double lBx = 0.0, lBy = 0.0;
        double l3[4], lb[2];
        if (stage > 1) {
            const double c0 = (stage == 2) ? 0.75 : 1.0 / 3.0;
            const double c1 = (stage == 2) ? 0.25 : 2.0 / 3.0;
            double V[4], lo[4], lV[4];
            for (int q = 0; q < 4; ++q) {
                V[q] = c0 * U0[q][k] + c1 * S.U3[q][k];
                lo[q] = lout[q][k];
            }
            floors_vjp(V, Bx, By, vol, g, lo, lV, lBx, lBy);
            for (int q = 0; q < 4; ++q) {
                lU0[q][k] += c0 * lV[q];
                l3[q] = c1 * lV[q];
            }
            for (int q = 0; q < 2; ++q) {
                lU0[4 + q][k] += c0 * lout[4 + q][k];
                lb[q] = c1 * lout[4 + q][k];
            }
        } else {
            for (int q = 0; q < 4; ++q) l3[q] = lout[q][k];
            for (int q = 0; q < 2; ++q) lb[q] = lout[4 + q][k];
        }
        double lPd[6] = {0, 0, 0, 0, 0, 0};
        bool xlo = i == 0 && !G.px, ylo = j == 0 && !G.py;
        if (xlo || ylo) {
            double P6[6];
            for (int f = 0; f < 6; ++f) P6[f] = S.P[f][k];
            if (xlo) outflow_vjp(0, P6, g, dt * G.dx, l3, lPd);
            if (ylo) outflow_vjp(1, P6, g, dt * G.dx, l3, lPd);
        }
        double Up[4], lUp[4];
        for (int q = 0; q < 4; ++q) Up[q] = S.Up[q][k];
        floors_vjp(Up, Bx, By, vol, g, l3, lUp, lBx, lBy);
        for (int q = 0; q < 4; ++q) {
            R.lUp[q][k] = lUp[q];
            lin[q][k] = lUp[q];   // the update adds the stage input with weight 1
        }
        for (int q = 0; q < 2; ++q) {
            R.lBp[q][k] = lb[q];
            lin[4 + q][k] = lb[q];  // CT adds dt*curl to the stage input's face B
        }
        R.lBc[0][k] = lBx;
        R.lBc[1][k] = lBy;
        for (int f = 0; f < 6; ++f) R.lPd[f][k] = lPd[f];
    }
}

// lam of -Ez at every corner from the lam of the CT-updated face B (reverse of the
// compute_curl_2d of -Ez, src/mhd_field_ops.f90:15-57, with dt).
__global__ void ar_corner(AdjGrid G, RevBufs R, double dt) {
    const int nx = G.nx, ny = G.ny;
    const double s = dt / G.dx;
    CELL_LOOP(k, (size_t)nx * ny) {
        int a = (int)(k % nx), b = (int)(k / nx);
        const double *lbx = R.lBp[0], *lby = R.lBp[1];
        double acc = 0.0;
        // dbx(i,j) = (negEz(i,j) - negEz(i,j-1)) / dx for j >= 1 or periodic;
        // a non-periodic ylo row uses dbx(i,1).
        if (b >= 1 || G.py) acc += s * lbx[k];
        {
            int jn = b + 1;
            if (jn < ny) acc -= s * lbx[k + nx];
            else if (G.py) acc -= s * lbx[(size_t)a];
        }
        if (!G.py) {
            if (b == 1) acc += s * lbx[(size_t)a];
            if (b == 0) acc -= s * lbx[(size_t)a];
        }
        // dby(i,j) = -(negEz(i,j) - negEz(i-1,j)) / dx for i >= 1 or periodic;
        // a non-periodic xlo column uses dby(1,j).
        if (a >= 1 || G.px) acc -= s * lby[k];
        {
            int in = a + 1;
            if (in < nx) acc += s * lby[k + 1];
            else if (G.px) acc += s * lby[(size_t)nx * b];
        }
        if (!G.px) {
            if (a == 1) acc -= s * lby[(size_t)nx * b];
            if (a == 0) acc += s * lby[(size_t)nx * b];
        }
        R.lnegEz[k] = acc;
    }
}

// Reverse of the thermal check, the rho/P face floor and MOOD for one face state of
// cell k (side s): lam of the face -> lam of the MUSCL candidate (cand) or of the
// cell value (lreset).
__device__ __forceinline__ void face_state_reverse(double lam[6], int s, bool to_cell,
                                                   unsigned long long bits, double lreset[6],
                                                   double *const cand[6], size_t k) {
    if (bits & (1ull << (24 + s))) {   // thermal reset: every field took the cell value
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            lreset[f] += lam[f];
            lam[f] = 0.0;
        }
    }
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        double v = lam[f];
        if (f == 0 || f == 3) {
            const int kk = f == 0 ? 0 : 1;
            if (bits & (1ull << (32 + 4 * kk + s))) v = 0.0;   // floored to 1e-8
        }
        const bool reset = (bits >> (4 * f + s)) & 1ull;
        if (reset || to_cell) {
            lreset[f] += v;
            cand[f][k] = 0.0;
        } else {
            cand[f][k] = v;
        }
    }
}

// Per face owner cell, one direction: lam of the face flux (from the update of the
// two cells beside it and from the corner EMFs), the Riemann VJP, then the reverse
// of the face-state construction.  XDIR: face (i+1/2, j); else (i, j+1/2), rotated.
template <int SOLVER, bool XDIR>
__device__ __forceinline__ void face_reverse(const AdjGrid &G, const StageBufs &S, const RevBufs &R,
                                             double dt, int i, int j, size_t k,
                                             unsigned long long bits, double lreset[6]) {
    const int nx = G.nx, ny = G.ny;
    const double dtdx = dt * G.dx;
    auto lEz = [&](int a, int b) { return -R.lnegEz[(size_t)a + (size_t)nx * b]; };
    const bool has_next = XDIR ? (i < nx - 1 || G.px) : (j < ny - 1 || G.py);
    const size_t kn = XDIR ? (i < nx - 1 ? k + 1 : (size_t)nx * j) : (j < ny - 1 ? k + nx : (size_t)i);
    double lF[5];  // storage order Mass, Momx, Momy, Energy, B flux
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = -dtdx * R.lUp[q][k];
        if (has_next) v += dtdx * R.lUp[q][kn];
        lF[q] = v;
    }
    if (XDIR) {  // F_x(By) enters Ez(i,j) and Ez(i,j-1) with -1/4
        double v = -0.25 * lEz(i, j);
        if (j >= 1) v += -0.25 * lEz(i, j - 1);
        else if (G.py) v += -0.25 * lEz(i, ny - 1);
        if (!G.py && j == ny - 1) v += -0.25 * lEz(i, j);
        lF[4] = v;
    } else {     // F_y(Bx) enters Ez(i,j) and Ez(i-1,j) with +1/4
        double v = 0.25 * lEz(i, j);
        if (i >= 1) v += 0.25 * lEz(i - 1, j);
        else if (G.px) v += 0.25 * lEz(nx - 1, j);
        if (!G.px && i == nx - 1) v += 0.25 * lEz(i, j);
        lF[4] = v;
    }
    constexpr int sL = XDIR ? 0 : 2, sR = XDIR ? 1 : 3;
    constexpr int fu = XDIR ? 1 : 2, fw = XDIR ? 2 : 1, fbn = XDIR ? 4 : 5, fbt = XDIR ? 5 : 4;
    FaceState a{S.face[4 * 0 + sL][k], S.face[4 * fu + sL][k], S.face[4 * fw + sL][k],
                S.face[4 * 3 + sL][k], S.face[4 * fbn + sL][k], S.face[4 * fbt + sL][k]};
    FaceState b{S.face[4 * 0 + sR][k], S.face[4 * fu + sR][k], S.face[4 * fw + sR][k],
                S.face[4 * 3 + sR][k], S.face[4 * fbn + sR][k], S.face[4 * fbt + sR][k]};
    // Riemann output order mass, mom_n, mom_t, energy, bt
    double lFr[5] = {lF[0], XDIR ? lF[1] : lF[2], XDIR ? lF[2] : lF[1], lF[3], lF[4]};
    double lLr[6], lRr[6];
    riemann_vjp<SOLVER, 1>(a, b, G.gamma, lFr, lLr, lRr);
    double lL[6], lR[6];   // field order rho, vx, vy, P, Bx, By
    lL[0] = lLr[0]; lL[fu] = lLr[1]; lL[fw] = lLr[2]; lL[3] = lLr[3]; lL[fbn] = lLr[4]; lL[fbt] = lLr[5];
    lR[0] = lRr[0]; lR[fu] = lRr[1]; lR[fw] = lRr[2]; lR[3] = lRr[3]; lR[fbn] = lRr[4]; lR[fbt] = lRr[5];
    const bool lastf = XDIR ? (!G.px && i == nx - 1) : (!G.py && j == ny - 1);
    double *const *candL = XDIR ? R.lXL : R.lYL;
    double *const *candR = XDIR ? R.lXR : R.lYR;
    face_state_reverse(lL, sL, false, bits, lreset, candL, k);
    face_state_reverse(lR, sR, lastf, bits, lreset, candR, k);
}

template <int SOLVER>
__global__ __launch_bounds__(256) void ar_faces(AdjGrid G, StageBufs S, RevBufs R, double dt) {
    const int nx = G.nx, ny = G.ny;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        unsigned long long bits = S.bits[k];
        double lreset[6] = {0, 0, 0, 0, 0, 0};
        face_reverse<SOLVER, true>(G, S, R, dt, i, j, k, bits, lreset);
        face_reverse<SOLVER, false>(G, S, R, dt, i, j, k, bits, lreset);
#pragma unroll
        for (int f = 0; f < 6; ++f) R.lPr[f][k] = lreset[f];
    }
}

// lam of the slopes and of the cell values reached directly.
__global__ void ar_slopes(AdjGrid G, RevBufs R) {
    const int nx = G.nx, ny = G.ny;
    const double hdx = 0.5 * G.dx;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        bool hl = i > 0 || G.px, hb = j > 0 || G.py;   // a face on the lo side exists
        size_t kl = i > 0 ? k - 1 : k + nx - 1;
        size_t kb = j > 0 ? k - nx : k + (size_t)nx * (ny - 1);
        for (int f = 0; f < 6; ++f) {
            double xr = hl ? R.lXR[f][kl] : 0.0;   // the right state of face i-1/2 came from (i,j)
            double yr = hb ? R.lYR[f][kb] : 0.0;
            R.lSx[f][k] = hdx * (R.lXL[f][k] - xr);
            R.lSy[f][k] = hdx * (R.lYL[f][k] - yr);
            R.lP[f][k] = R.lXL[f][k] + xr + R.lYL[f][k] + yr + R.lPr[f][k] + R.lPd[f][k];
        }
    }
}

// Reverse of the limiter (gather over the three slopes that read this cell), then of
// cons_to_prim; writes lam of the conserved stage input and the total lam of cell B.
template <int LIMITER>
__global__ void ar_prims(AdjGrid G, const double *const *U, StageBufs S, RevBufs R,
                         double *const *lin) {
    const int nx = G.nx, ny = G.ny;
    const double inv_dx = 1.0 / G.dx, inv_2dx = 1.0 / (2.0 * G.dx);
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        double lp6[6];
        for (int f = 0; f < 6; ++f) {
            double acc = R.lP[f][k];
            const double *a = S.P[f];
            for (int dir = 0; dir < 2; ++dir) {
                const int n = dir == 0 ? nx : ny, per = dir == 0 ? G.px : G.py;
                const int c = dir == 0 ? i : j;
                const double *ls = dir == 0 ? R.lSx[f] : R.lSy[f];
                for (int o = -1; o <= 1; ++o) {
                    int kc = c + o;
                    if (per) kc = (kc + n) % n;
                    else if (kc < 0 || kc >= n) continue;
                    if (o != 0 && kc == c) continue;   // tiny periodic axes
                    int km = wrap_or_clamp(kc - 1, n, per), kp = wrap_or_clamp(kc + 1, n, per);
                    auto at = [&](int t) {
                        return dir == 0 ? a[(size_t)t + (size_t)nx * j] : a[(size_t)i + (size_t)nx * t];
                    };
                    size_t ks = dir == 0 ? (size_t)kc + (size_t)nx * j : (size_t)i + (size_t)nx * kc;
                    double lam = ls[ks];
                    if (lam == 0.0) continue;
                    double dm, d0, dp;
                    if (G.use_lim) {
                        limited_slope_grad<LIMITER>(at(km), at(kc), at(kp), inv_dx, dm, d0, dp);
                    } else {
                        dm = -inv_2dx;
                        d0 = 0.0;
                        dp = inv_2dx;
                    }
                    double w = (km == c ? dm : 0.0) + (kc == c ? d0 : 0.0) + (kp == c ? dp : 0.0);
                    acc += lam * w;
                }
            }
            lp6[f] = acc;
        }
        double Bx = S.P[4][k], By = S.P[5][k];
        double lU[4] = {0, 0, 0, 0};
        double lBx = lp6[4] + R.lBc[0][k], lBy = lp6[5] + R.lBc[1][k];
        cons_to_prim_vjp(U[0][k], U[1][k], U[2][k], U[3][k], Bx, By, G.gamma, G.vol, lp6, lU, lBx, lBy);
        for (int q = 0; q < 4; ++q) lin[q][k] += lU[q];
        R.lBcT[0][k] = lBx;
        R.lBcT[1][k] = lBy;
    }
}

// Reverse of the face -> cell average (src/mhd_field_ops.f90:120-156).
__global__ void ar_bavg(AdjGrid G, RevBufs R, double *const *lin) {
    const int nx = G.nx, ny = G.ny;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        // b_x(i) enters Bx(i) and Bx(i+1)
        double ax = (i == 0 && !G.px) ? R.lBcT[0][k] : 0.5 * R.lBcT[0][k];
        if (i + 1 < nx) ax += 0.5 * R.lBcT[0][k + 1];
        else if (G.px) ax += 0.5 * R.lBcT[0][(size_t)nx * j];
        double ay = (j == 0 && !G.py) ? R.lBcT[1][k] : 0.5 * R.lBcT[1][k];
        if (j + 1 < ny) ay += 0.5 * R.lBcT[1][k + nx];
        else if (G.py) ay += 0.5 * R.lBcT[1][(size_t)i];
        lin[4][k] += ax;
        lin[5][k] += ay;
    }
}

__global__ void add_kernel(double *const *a, const double *const *b, size_t n) {
    CELL_LOOP(k, n) {
        for (int q = 0; q < 6; ++q) a[q][k] += b[q][k];
    }
}

__global__ void zero6_kernel(double *const *a, size_t n) {
    CELL_LOOP(k, n) {
        for (int q = 0; q < 6; ++q) a[q][k] = 0.0;
    }
}

inline int nblocks(size_t n) {
    size_t b = (n + BS - 1) / BS;
    return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace
}  // namespace chimera

using namespace chimera;

struct Ptr6 {
    double *p[6];
};

struct chimera_adj {
    AdjGrid G{};
    chimera_scheme_t sch{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t n = 0;
    double *arena = nullptr;
    StageBufs S[3];
    RevBufs R;
    Ptr6 stage_state[3];   // U^n (copied), U^(1), U^(2)
    Ptr6 lam_a, lam_b, lam_u0, scratch_out;
    // device-side pointer tables (kernels take double *const *)
    double **dptr = nullptr;
};

namespace {

// device pointer tables: [0] stage_state[0..2], [3] lam_a, [4] lam_b, [5] lam_u0,
// [6] scratch_out, [7] caller input, [8] caller output, [9] caller lam in, [10] caller lam out
enum { TB_S0 = 0, TB_S1 = 1, TB_S2 = 2, TB_LA = 3, TB_LB = 4, TB_LU0 = 5, TB_SCR = 6, TB_IN = 7,
       TB_OUT = 8, TB_LIN = 9, TB_LOUT = 10, TB_COUNT = 11 };

double *const *table(chimera_adj *a, int t) { return a->dptr + 6 * t; }

// The caller's pointer arrays may be temporaries: copy them synchronously (after
// the work already queued on the stream).
int set_table(chimera_adj *a, int t, const double *const *p) {
    CH_HIP(hipStreamSynchronize(a->stream));
    CH_HIP(hipMemcpy(a->dptr + 6 * t, p, 6 * sizeof(double *), hipMemcpyHostToDevice));
    return CHIMERA_OK;
}

int run_forward_stage(chimera_adj *a, int stage, int t_in, int t_u0, int t_out, double dt) {
    StageBufs &S = a->S[stage - 1];
    const int nb = nblocks(a->n);
    hipLaunchKernelGGL(af_prims, dim3(nb), dim3(BS), 0, a->stream, a->G, table(a, t_in), S);
    if (a->sch.slope_limiter == LIM_MC)
        hipLaunchKernelGGL(af_faces<LIM_MC>, dim3(nb), dim3(BS), 0, a->stream, a->G, S);
    else
        hipLaunchKernelGGL(af_faces<LIM_VAN_LEER>, dim3(nb), dim3(BS), 0, a->stream, a->G, S);
    switch (a->sch.riemann_solver) {
    case RUSANOV: hipLaunchKernelGGL(af_fluxes<RUSANOV>, dim3(nb), dim3(BS), 0, a->stream, a->G, S); break;
    case HLLE: hipLaunchKernelGGL(af_fluxes<HLLE>, dim3(nb), dim3(BS), 0, a->stream, a->G, S); break;
    default: hipLaunchKernelGGL(af_fluxes<HLLD>, dim3(nb), dim3(BS), 0, a->stream, a->G, S); break;
    }
    hipLaunchKernelGGL(af_update, dim3(nb), dim3(BS), 0, a->stream, a->G, table(a, t_in),
                       table(a, t_u0), table(a, t_out), S, stage, dt);
    return launch_check("adjoint forward stage");
}

int run_reverse_stage(chimera_adj *a, int stage, int t_in, int t_u0, int t_lout, int t_lin,
                      double dt) {
    StageBufs &S = a->S[stage - 1];
    RevBufs &R = a->R;
    const int nb = nblocks(a->n);
    hipLaunchKernelGGL(ar_cell_tail, dim3(nb), dim3(BS), 0, a->stream, a->G, table(a, t_u0), S, R,
                       table(a, t_lout), table(a, t_lin), table(a, TB_LU0), stage, dt);
    hipLaunchKernelGGL(ar_corner, dim3(nb), dim3(BS), 0, a->stream, a->G, R, dt);
    switch (a->sch.riemann_solver) {
    case RUSANOV: hipLaunchKernelGGL(ar_faces<RUSANOV>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R, dt); break;
    case HLLE: hipLaunchKernelGGL(ar_faces<HLLE>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R, dt); break;
    default: hipLaunchKernelGGL(ar_faces<HLLD>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R, dt); break;
    }
    hipLaunchKernelGGL(ar_slopes, dim3(nb), dim3(BS), 0, a->stream, a->G, R);
    if (a->sch.slope_limiter == LIM_MC)
        hipLaunchKernelGGL(ar_prims<LIM_MC>, dim3(nb), dim3(BS), 0, a->stream, a->G, table(a, t_in), S, R,
                           table(a, t_lin));
    else
        hipLaunchKernelGGL(ar_prims<LIM_VAN_LEER>, dim3(nb), dim3(BS), 0, a->stream, a->G,
                           table(a, t_in), S, R, table(a, t_lin));
    hipLaunchKernelGGL(ar_bavg, dim3(nb), dim3(BS), 0, a->stream, a->G, R, table(a, t_lin));
    return launch_check("adjoint reverse stage");
}

}  // namespace

extern "C" int chimera_adj_destroy(chimera_adj_t *a) {
    if (!a) return CHIMERA_OK;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    (void)hipFree(a->arena);
    (void)hipFree(a->dptr);
    if (a->own_stream && a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
    return CHIMERA_OK;
}

extern "C" int chimera_adj_create(chimera_adj_t **out, int nx, int ny, double dx, double gamma,
                                  const chimera_bc_t *bc, const chimera_scheme_t *scheme,
                                  void *stream) {
    CH_REQUIRE(out, "out is NULL");
    CH_REQUIRE(nx >= 4 && ny >= 4, "grid must be at least 4 x 4");
    CH_REQUIRE(dx > 0.0 && gamma > 1.0, "dx must be > 0 and gamma > 1");
    int px, py;
    CH_TRY(check_bc(bc, &px, &py));
    CH_TRY(check_scheme(scheme));
    chimera_adj *a = new chimera_adj();
    a->G = AdjGrid{nx, ny, px, py, dx, dx * dx, gamma, scheme->use_slope_limiting,
                   scheme->fallback_2_mood};
    a->sch = *scheme;
    a->n = (size_t)nx * ny;
    if (stream) {
        a->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess) {
            delete a;
            set_error("hipStreamCreate failed");
            return CHIMERA_EHIP;
        }
        a->own_stream = true;
    }
    // per stage: 6 P + 24 faces + 1 bits + 10 fluxes + 4 Up + 4 U3 = 49
    // reverse: 4 + 2 + 2 + 6 + 1 + 30 + 12 + 6 + 2 = 65; states 3 x 6; lambdas 4 x 6
    const size_t per_stage = 49, rev = 65, extra = 18 + 24;
    const size_t total = 3 * per_stage + rev + extra;
    if (hipMalloc(&a->arena, total * a->n * sizeof(double)) != hipSuccess ||
        hipMalloc(&a->dptr, 6 * TB_COUNT * sizeof(double *)) != hipSuccess) {
        set_error("adjoint workspace allocation (%zu arrays of %zu cells) failed", total, a->n);
        chimera_adj_destroy(a);
        return CHIMERA_ENOMEM;
    }
    double *p = a->arena;
    auto take = [&]() {
        double *r = p;
        p += a->n;
        return r;
    };
    for (int s = 0; s < 3; ++s) {
        StageBufs &S = a->S[s];
        for (auto &x : S.P) x = take();
        for (auto &x : S.face) x = take();
        S.bits = reinterpret_cast<unsigned long long *>(take());
        for (auto &x : S.FX) x = take();
        for (auto &x : S.FY) x = take();
        for (auto &x : S.Up) x = take();
        for (auto &x : S.U3) x = take();
    }
    RevBufs &R = a->R;
    for (auto &x : R.lUp) x = take();
    for (auto &x : R.lBp) x = take();
    for (auto &x : R.lBc) x = take();
    for (auto &x : R.lPd) x = take();
    R.lnegEz = take();
    for (auto &x : R.lXL) x = take();
    for (auto &x : R.lXR) x = take();
    for (auto &x : R.lYL) x = take();
    for (auto &x : R.lYR) x = take();
    for (auto &x : R.lPr) x = take();
    for (auto &x : R.lSx) x = take();
    for (auto &x : R.lSy) x = take();
    for (auto &x : R.lP) x = take();
    for (auto &x : R.lBcT) x = take();
    for (auto &st : a->stage_state)
        for (auto &x : st.p) x = take();
    for (auto &x : a->lam_a.p) x = take();
    for (auto &x : a->lam_b.p) x = take();
    for (auto &x : a->lam_u0.p) x = take();
    for (auto &x : a->scratch_out.p) x = take();
    int rc = CHIMERA_OK;
    if (rc == CHIMERA_OK) rc = set_table(a, TB_S0, a->stage_state[0].p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_S1, a->stage_state[1].p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_S2, a->stage_state[2].p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_LA, a->lam_a.p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_LB, a->lam_b.p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_LU0, a->lam_u0.p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_SCR, a->scratch_out.p);
    if (rc == CHIMERA_OK && hipStreamSynchronize(a->stream) != hipSuccess) rc = CHIMERA_EHIP;
    if (rc != CHIMERA_OK) {
        chimera_adj_destroy(a);
        return rc;
    }
    *out = a;
    return CHIMERA_OK;
}

// One SSP-RK3 step with a given dt through the adjoint's forward kernels
// (device pointers).
extern "C" int chimera_adj_forward_step(chimera_adj_t *a, const double *const *state_in,
                                        double dt, double *const *state_out) {
    CH_REQUIRE(a && state_in && state_out, "NULL argument");
    size_t bytes = a->n * sizeof(double);
    for (int q = 0; q < 6; ++q)
        CH_HIP(hipMemcpyAsync(a->stage_state[0].p[q], state_in[q], bytes, hipMemcpyDefault, a->stream));
    CH_TRY(set_table(a, TB_OUT, (const double *const *)state_out));
    CH_TRY(run_forward_stage(a, 1, TB_S0, TB_S0, TB_S1, dt));
    CH_TRY(run_forward_stage(a, 2, TB_S1, TB_S0, TB_S2, dt));
    CH_TRY(run_forward_stage(a, 3, TB_S2, TB_S0, TB_OUT, dt));
    CH_HIP(hipStreamSynchronize(a->stream));
    return CHIMERA_OK;
}

// lam_in = (d U^{n+1} / d U^n)^T lam_out for one step from U^n = state_n with dt frozen.
extern "C" int chimera_adj_step_vjp(chimera_adj_t *a, const double *const *state_n, double dt,
                                    const double *const *lam_out, double *const *lam_in) {
    CH_REQUIRE(a && state_n && lam_out && lam_in, "NULL argument");
    size_t bytes = a->n * sizeof(double);
    for (int q = 0; q < 6; ++q) {
        CH_HIP(hipMemcpyAsync(a->stage_state[0].p[q], state_n[q], bytes, hipMemcpyDefault, a->stream));
        CH_HIP(hipMemcpyAsync(a->lam_a.p[q], lam_out[q], bytes, hipMemcpyDefault, a->stream));
    }
    CH_TRY(set_table(a, TB_LIN, (const double *const *)lam_in));
    const int nb = nblocks(a->n);
    // forward: keep every stage's intermediates
    CH_TRY(run_forward_stage(a, 1, TB_S0, TB_S0, TB_S1, dt));
    CH_TRY(run_forward_stage(a, 2, TB_S1, TB_S0, TB_S2, dt));
    CH_TRY(run_forward_stage(a, 3, TB_S2, TB_S0, TB_SCR, dt));
    hipLaunchKernelGGL(zero6_kernel, dim3(nb), dim3(BS), 0, a->stream, table(a, TB_LU0), a->n);
    // reverse: stage 3 (lam_a -> lam_b), stage 2 (lam_b -> lam_a), stage 1 (lam_a -> lam_b)
    CH_TRY(run_reverse_stage(a, 3, TB_S2, TB_S0, TB_LA, TB_LB, dt));
    CH_TRY(run_reverse_stage(a, 2, TB_S1, TB_S0, TB_LB, TB_LA, dt));
    CH_TRY(run_reverse_stage(a, 1, TB_S0, TB_S0, TB_LA, TB_LB, dt));
    hipLaunchKernelGGL(add_kernel, dim3(nb), dim3(BS), 0, a->stream, table(a, TB_LB),
                       (const double *const *)table(a, TB_LU0), a->n);
    CH_TRY(launch_check("adjoint accumulate"));
    for (int q = 0; q < 6; ++q)
        CH_HIP(hipMemcpyAsync(lam_in[q], a->lam_b.p[q], bytes, hipMemcpyDefault, a->stream));
    CH_HIP(hipStreamSynchronize(a->stream));
    return CHIMERA_OK;
}

extern "C" void *chimera_adj_stream(chimera_adj_t *a) { return a ? (void *)a->stream : nullptr; }
