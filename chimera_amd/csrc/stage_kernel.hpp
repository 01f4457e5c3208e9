// Fused SSP-RK3 stage kernel (gfx950): one workgroup advances a TX x TY tile of
// cells through a whole RK stage in one pass over HBM.
//
// Per tile:
//   phase 1  load the conserved state of the tile plus a 2-cell halo (face B with
//            one more cell on the lo side), apply the boundary mapping at load time
//            (periodic wrap / zero-gradient clamp, src/mhd_bc.f90), and store the
//            stage primitives (src/mhd_field_ops.f90:120-156 face->cell B,
//            src/mhd_change_states.f90:61-113) in LDS;
//   phase 2  for every x face (TX+1 x TY+2) and y face (TX+2 x TY+1) the tile needs:
//            limited slopes (src/mhd_derivatives.f90:50-127), MUSCL face states and
//            the MOOD bounds check (130-229), the thermal-pressure check (232-293) and
//            the Riemann flux (src/mhd_flux.f90:139-779); fluxes go to LDS;
//   phase 3  per cell: flux-difference update (mhd_flux.f90:20-76), conserved floors
//            with the stage's cell B (change_states:116-166), constrained transport of
//            the face B (mhd_flux.f90:79-136), the outflow boundary correction and the
//            RK blend with floors (src/main.f90:107-267); one write of the new state.
//
// The blockIdx -> tile map is XCD-aware: the 8 XCDs receive consecutive blocks
// round-robin, so block b is sent to tile (b % 8) * (grid / 8) + b / 8, giving each
// XCD a contiguous band of tiles whose halos then hit its own L2.
#pragma once

#include <cstdint>

#include "mhd_phys.hpp"

namespace chimera {

// One grid as the stage kernels see it.
struct GridView {
    const double *uin[6];  // stage input: Mass, Momx, Momy, Energy, b_x, b_y
    const double *u0[6];   // u^n (read only by stages 2 and 3)
    double *uout[6];       // stage output
    const double *dt;      // device scalar
    uint32_t *mask;        // optional MOOD / thermal-check bits (stage 1 diagnostics)
    int nx, ny;
    int px, py;            // periodic x / y
    double dx, vol, gamma;
};

// Stage intermediates the discrete adjoint reads back (owned cells only): all of them
// with STORE_ALL, only Up with STORE_UP (the segment recompute of the fused reverse).
enum StoreMode { STORE_NONE = 0, STORE_ALL = 1, STORE_UP = 2 };
struct StageStore {
    double *P[6];        // stage primitives rho, vx, vy, P, Bx, By
    double *face[24];    // field f, side s (XL, XR, YL, YR) at 4*f + s, after the thermal check
    unsigned long long *bits_x, *bits_y;  // MOOD 4f+s | thermal 24+s | face floor 32+4k+s (k: rho, P)
    double *Up[4];       // after the flux update, before the floors
    double *U3[4];       // after floors and the outflow correction, before the blend
};

// LDS layout of a tile.  The transverse flux (By on x faces, Bx on y faces) is kept on
// every face the corner EMFs read, one row / column beyond the tile on each side; the
// four conserved fluxes only on the faces of the tile's own rows / columns, which is
// all the flux difference reads.
template <int TX, int TY>
struct TileShape {
    static constexpr int PW = TX + 4, PH = TY + 4;     // primitives with a 2-cell halo
    static constexpr int FXW = TX + 1, FXH = TY + 2;   // x faces i in [-1,TX-1], j in [-1,TY]
    static constexpr int FYW = TX + 2, FYH = TY + 1;   // y faces i in [-1,TX],   j in [-1,TY-1]
    static constexpr int P_SZ = PW * PH;
    static constexpr int FX_SZ = FXW * FXH;            // transverse flux, all x faces
    static constexpr int FY_SZ = FYW * FYH;            // transverse flux, all y faces
    static constexpr int FX4_SZ = FXW * TY;            // conserved fluxes, x faces j in [0,TY-1]
    static constexpr int FY4_SZ = TX * FYH;            // conserved fluxes, y faces i in [0,TX-1]
    static constexpr int LDS_DOUBLES = 6 * P_SZ + 4 * FX4_SZ + FX_SZ + 4 * FY4_SZ + FY_SZ;
};

// Global index of a halo position: periodic wrap or zero-gradient clamp.  Halo
// positions lie in [-2, n + T + 1], so one conditional add / subtract wraps them
// whenever n >= T + 2; smaller grids fall back to the modulo.
__device__ __forceinline__ int map_index(int g, int n, int periodic) {
    if (periodic) {
        if (g < 0) g += n;
        if (g >= n) {
            g -= n;
            if (g >= n) g %= n;
        }
        return g;
    }
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

// Field order in LDS / face registers: rho, vx, vy, P, Bx, By.
enum { F_RHO = 0, F_VX = 1, F_VY = 2, F_P = 3, F_BX = 4, F_BY = 5 };

// Limited (or central) slope of a cell from its two neighbours along one direction.
template <int LIMITER>
__device__ __forceinline__ double cell_slope(double m, double c, double p, bool use_lim,
                                             double inv_dx, double inv_2dx) {
    return use_lim ? limited_slope<LIMITER>(m, c, p, inv_dx) : central_slope(m, p, inv_2dx);
}

// Face states of one field at one face from the slopes of its two cells: the MUSCL
// left state (cell c extrapolated forward), the right state (next cell extrapolated
// backward, or the cell value itself at a non-periodic hi boundary), then MOOD against
// the 5-point stencil of cell c (neighbours m, p, nA, nB) and the rho/P floor.  Returns
// the MOOD bits (1: left reset, 2: right reset); floor_bits likewise for the floor.
__device__ __forceinline__ unsigned face_states(double m, double c, double p, double nA,
                                                double nB, double sl, double sr, bool last,
                                                bool mood, bool rho_or_p, double hdx, double &fl,
                                                double &fr, unsigned &floor_bits) {
    double l = c + hdx * sl;
    double r = last ? c : p - hdx * sr;
    unsigned bits = 0;
    if (mood) {
        double lo = fmin(fmin(fmin(fmin(c, p), m), nA), nB);
        double hi = fmax(fmax(fmax(fmax(c, p), m), nA), nB);
        if (rho_or_p) lo = fmax(1e-9, lo);
        if (l < lo || l > hi) { l = c; bits |= 1u; }
        if (r < lo || r > hi) { r = c; bits |= 2u; }
    }
    floor_bits = 0u;
    if (rho_or_p) {
        floor_bits = (1.0e-8 > l ? 1u : 0u) | (1.0e-8 > r ? 2u : 0u);
        l = fmax(l, 1.0e-8);
        r = fmax(r, 1.0e-8);
    }
    fl = l;
    fr = r;
    return bits;
}

// Work split of phase 2: every face line (x faces along a row, y faces along a column)
// is cut into runs of consecutive faces, as many runs as threads (runs as even as
// possible, the longest lines cut most), ordered longest first so that runs of one
// length fill whole waves.  Built at compile time per tile shape.
struct FaceRun {
    int dir, line, start, len;
};

template <int TX, int TY>
struct FaceRunTable {
    static constexpr int NT = TX * TY;
    static constexpr int NXL = TY + 2, XLEN = TX + 1;   // x-face rows
    static constexpr int NYL = TX + 2, YLEN = TY + 1;   // y-face columns
    static constexpr int NL = NXL + NYL;
    FaceRun r[NT];
    constexpr FaceRunTable() : r{} {
        int cnt[NL] = {};
        int lmax = 1;
        // the shortest maximum run length whose runs fit in NT threads
        for (;; ++lmax) {
            int total = 0;
            for (int l = 0; l < NL; ++l) total += (len_of(l) + lmax - 1) / lmax;
            if (total <= NT) break;
        }
        int total = 0;
        for (int l = 0; l < NL; ++l) {
            cnt[l] = (len_of(l) + lmax - 1) / lmax;
            total += cnt[l];
        }
        while (total < NT) {   // one more run on the line with the longest runs
            int best = -1, longest = 1;
            for (int l = 0; l < NL; ++l) {
                int m = len_of(l);
                if (cnt[l] >= m) continue;
                int lr = (m + cnt[l] - 1) / cnt[l];
                if (lr > longest) {
                    longest = lr;
                    best = l;
                }
            }
            if (best < 0) break;
            ++cnt[best];
            ++total;
        }
        int t = 0;
        for (int want = lmax; want >= 1; --want)
            for (int l = 0; l < NL; ++l) {
                int m = len_of(l), c = cnt[l], base = m / c, extra = m % c, pos = 0;
                for (int k = 0; k < c; ++k) {
                    int len = base + (k < extra ? 1 : 0);
                    if (len == want && t < NT) {
                        r[t] = FaceRun{l < NXL ? 0 : 1, l < NXL ? l : l - NXL, pos, len};
                        ++t;
                    }
                    pos += len;
                }
            }
        for (; t < NT; ++t) r[t] = FaceRun{0, 0, 0, 0};
    }
    static constexpr int len_of(int l) { return l < NXL ? XLEN : YLEN; }
};

template <int TX, int TY>
__device__ constexpr FaceRunTable<TX, TY> kFaceRuns{};

// Thermal-pressure check of one face (src/mhd_derivatives.f90:262-267): when the
// implied thermal pressure is below P_floor every field takes the cell value.
__device__ __forceinline__ bool thermal_reset(double f[6], const double c[6]) {
    double p_th = f[F_P] - 0.5 * (f[F_BX] * f[F_BX] + f[F_BY] * f[F_BY]);
    if (p_th < P_FLOOR) {
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] = c[q];
        return true;
    }
    return false;
}

template <int SOLVER, int LIMITER, int TX, int TY, int STORE = STORE_NONE>
__device__ __forceinline__ void stage_tile(const GridView &g, int stage, bool use_lim, bool mood,
                                           int tile_x, int tile_y, double *lds,
                                           const StageStore *store = nullptr) {
    using S = TileShape<TX, TY>;
    constexpr int NT = TX * TY;
    const int tid = threadIdx.x;
    const int nx = g.nx, ny = g.ny, px = g.px, py = g.py;
    const int x0 = tile_x * TX, y0 = tile_y * TY;
    double *sP = lds;                       // [6][PH][PW]
    double *sFX = lds + 6 * S::P_SZ;        // [4][TY][FXW]   conserved fluxes, x faces
    double *sFY = sFX + 4 * S::FX4_SZ;      // [4][FYH][TX]   conserved fluxes, y faces
    double *sBX = sFY + 4 * S::FY4_SZ;      // [FXH][FXW]     By flux, x faces
    double *sBY = sBX + S::FX_SZ;           // [FYH][FYW]     Bx flux, y faces
    const double dx = g.dx, vol = g.vol, gam = g.gamma;
    const double inv_dx = 1.0 / dx;
    const double inv_2dx = 1.0 / (2.0 * dx);
    const double hdx = 0.5 * dx;

    // ---- phase 1: halo load and stage primitives -------------------------------
    for (int idx = tid; idx < S::P_SZ; idx += NT) {
        int lx = idx % S::PW, ly = idx / S::PW;
        int gi = map_index(x0 - 2 + lx, nx, px);
        int gj = map_index(y0 - 2 + ly, ny, py);
        int gim = gi > 0 ? gi - 1 : (px ? nx - 1 : 0);
        int gjm = gj > 0 ? gj - 1 : (py ? ny - 1 : 0);
        size_t k = (size_t)gi + (size_t)nx * gj;
        double Bx = 0.5 * (g.uin[4][k] + g.uin[4][(size_t)gim + (size_t)nx * gj]);
        double By = 0.5 * (g.uin[5][k] + g.uin[5][(size_t)gi + (size_t)nx * gjm]);
        double rho, vx, vy, P;
        cons_to_prim(g.uin[0][k], g.uin[1][k], g.uin[2][k], g.uin[3][k], Bx, By, gam, vol, rho, vx,
                     vy, P);
        sP[0 * S::P_SZ + idx] = rho;
        sP[1 * S::P_SZ + idx] = vx;
        sP[2 * S::P_SZ + idx] = vy;
        sP[3 * S::P_SZ + idx] = P;
        sP[4 * S::P_SZ + idx] = Bx;
        sP[5 * S::P_SZ + idx] = By;
        if (STORE == STORE_ALL) {
            int ci = lx - 2, cj = ly - 2;
            if (ci >= 0 && ci < TX && cj >= 0 && cj < TY && x0 + ci < nx && y0 + cj < ny) {
                size_t ko = (size_t)(x0 + ci) + (size_t)nx * (y0 + cj);
                store->P[0][ko] = rho;
                store->P[1][ko] = vx;
                store->P[2][ko] = vy;
                store->P[3][ko] = P;
                store->P[4][ko] = Bx;
                store->P[5][ko] = By;
            }
        }
    }
    __syncthreads();

    const bool want_mask = g.mask != nullptr;

    // ---- phase 2: every x face and y face the tile needs, as runs of consecutive faces
    //      along their normal (one run per thread): the slope of a run's right cell is
    //      the next face's left-cell slope, so each slope is evaluated once per run.  x
    //      and y runs share one code path (strides and field roles by direction), so a
    //      wave never splits on the direction. ------------------------------------------
    {
        const FaceRun run = kFaceRuns<TX, TY>.r[tid];
        const int dir = run.dir;                        // 0: x faces, 1: y faces (rotated)
        const int sA = dir ? S::PW : 1;                 // LDS stride along the face normal
        const int sT = dir ? 1 : S::PW;                 // LDS stride across it
        // slope of the first face's left cell; afterwards each face's right-cell slope
        // is carried to the next face
        double sprev[6];
        if (run.len > 0) {
            const int fx0 = dir ? run.line : run.start, fy0 = dir ? run.start : run.line;
            const int lpos0 = (fy0 + 1) * S::PW + (fx0 + 1);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double *b = sP + q * S::P_SZ + lpos0;
                sprev[q] = cell_slope<LIMITER>(b[-sA], b[0], b[sA], use_lim, inv_dx, inv_2dx);
            }
        }
        for (int t = 0; t < run.len; ++t) {
            // x face: i = fx-1 in row j = fy-1, left cell at LDS (fx+1, fy+1);
            // y face: column i = fx-1, j = fy-1, lower cell at LDS (fx+1, fy+1)
            const int fx = dir ? run.line : run.start + t;
            const int fy = dir ? run.start + t : run.line;
            const int lpos = (fy + 1) * S::PW + (fx + 1);
            const bool last = dir ? (!py && map_index(y0 + fy - 1, ny, py) == ny - 1)
                                  : (!px && map_index(x0 + fx - 1, nx, px) == nx - 1);
            double L[6], R[6], C[6];
            unsigned bits = 0, fbits = 0;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double *b = sP + q * S::P_SZ + lpos;
                const double m = b[-sA], c = b[0], p = b[sA], pp = b[2 * sA];
                C[q] = c;
                const double sl = sprev[q];
                const double sr = cell_slope<LIMITER>(c, p, pp, use_lim, inv_dx, inv_2dx);
                sprev[q] = sr;
                unsigned fb;
                bits |= face_states(m, c, p, b[sT], b[-sT], sl, sr, last, mood,
                                    q == F_RHO || q == F_P, hdx, L[q], R[q], fb)
                        << (4 * q + 2 * dir);
                if (q == F_RHO) fbits |= fb << (2 * dir);
                if (q == F_P) fbits |= fb << (4 + 2 * dir);
            }
            if (thermal_reset(L, C)) bits |= 1u << (24 + 2 * dir);
            if (thermal_reset(R, C)) bits |= 1u << (25 + 2 * dir);
            const int i = fx - 1, j = fy - 1;
            const bool owned = i >= 0 && i < TX && j >= 0 && j < TY && x0 + i < nx && y0 + j < ny;
            if (STORE == STORE_ALL && owned) {
                size_t ko = (size_t)(x0 + i) + (size_t)nx * (y0 + j);
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    store->face[4 * q + 2 * dir][ko] = L[q];
                    store->face[4 * q + 2 * dir + 1][ko] = R[q];
                }
                unsigned long long w = (unsigned long long)bits | ((unsigned long long)fbits << 32);
                if (dir) store->bits_y[ko] = w;
                else store->bits_x[ko] = w;
            }
            FaceState a{L[F_RHO], dir ? L[F_VY] : L[F_VX], dir ? L[F_VX] : L[F_VY], L[F_P],
                        dir ? L[F_BY] : L[F_BX], dir ? L[F_BX] : L[F_BY]};
            FaceState bR{R[F_RHO], dir ? R[F_VY] : R[F_VX], dir ? R[F_VX] : R[F_VY], R[F_P],
                         dir ? R[F_BY] : R[F_BX], dir ? R[F_BX] : R[F_BY]};
            Flux5 f = riemann<SOLVER>(a, bR, gam);
            // transverse flux: By flux (x faces) / Bx flux (y faces), every face
            (dir ? sBY : sBX)[dir ? fy * S::FYW + fx : fy * S::FXW + fx] = f.bt;
            // conserved fluxes: faces of the tile's own rows (x) / columns (y)
            if (dir ? (fx >= 1 && fx <= TX) : (fy >= 1 && fy <= TY)) {
                double *sF = dir ? sFY : sFX;
                const int fsz = dir ? S::FY4_SZ : S::FX4_SZ;
                const int idx = dir ? fy * TX + (fx - 1) : (fy - 1) * S::FXW + fx;
                sF[0 * fsz + idx] = f.mass;
                sF[1 * fsz + idx] = dir ? f.mom_t : f.mom_n;   // Momx
                sF[2 * fsz + idx] = dir ? f.mom_n : f.mom_t;   // Momy
                sF[3 * fsz + idx] = f.energy;
            }
            if (want_mask && owned && bits)
                atomicOr(&g.mask[(size_t)(x0 + i) + (size_t)nx * (y0 + j)], bits);
        }
    }
    __syncthreads();

    // ---- phase 3: update, floors, CT, outflow correction, blend -----------------
    for (int c = tid; c < TX * TY; c += NT) {
        int ci = c % TX, cj = c / TX;
        int gi = x0 + ci, gj = y0 + cj;
        if (gi >= nx || gj >= ny) continue;
        size_t k = (size_t)gi + (size_t)nx * gj;
        const double dt = *g.dt;
        const double dtdx = dt * dx;
        // conserved fluxes: x faces FX(i, j) at [j][i+1], y faces FY(i, j) at [j+1][i]
        const int xr = cj * S::FXW + (ci + 1);
        const int yr = (cj + 1) * TX + ci;
        const bool xlo_open = (gi == 0) && !px;
        const bool ylo_open = (gj == 0) && !py;
        double U[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double v = g.uin[q][k];
            double fr = sFX[q * S::FX4_SZ + xr];
            if (xlo_open) v = v - dtdx * fr;
            else v = v - dtdx * fr + dtdx * sFX[q * S::FX4_SZ + xr - 1];
            double ft = sFY[q * S::FY4_SZ + yr];
            if (ylo_open) v = v - dtdx * ft;
            else v = v - dtdx * ft + dtdx * sFY[q * S::FY4_SZ + yr - TX];
            U[q] = v;
            if (STORE != STORE_NONE) store->Up[q][k] = v;
        }
        const int pc = (cj + 2) * S::PW + (ci + 2);
        const double Bxc = sP[4 * S::P_SZ + pc], Byc = sP[5 * S::P_SZ + pc];
        conserved_floors(U[0], U[1], U[2], U[3], Bxc, Byc, vol, gam);

        // corner EMF: Ez(a,b) = 1/4 (-F_x(By)(a,b) - F_x(By)(a,b+1) + F_y(Bx)(a,b) + F_y(Bx)(a+1,b))
        auto negEz = [&](int a, int b) -> double {
            const double *fbx = sBX;
            const double *fby = sBY;
            double f0 = fbx[(b + 1) * S::FXW + (a + 1)];
            double up = (!py && y0 + b == ny - 1) ? f0 : fbx[(b + 2) * S::FXW + (a + 1)];
            double e0 = fby[(b + 1) * S::FYW + (a + 1)];
            double rt = (!px && x0 + a == nx - 1) ? e0 : fby[(b + 1) * S::FYW + (a + 2)];
            double Ez = 0.25 * (-f0 - up + e0 + rt);
            return -Ez;
        };
        double dbx, dby;
        if (ylo_open) dbx = (negEz(ci, cj + 1) - negEz(ci, cj)) * inv_dx;
        else dbx = (negEz(ci, cj) - negEz(ci, cj - 1)) * inv_dx;
        if (xlo_open) dby = -((negEz(ci + 1, cj) - negEz(ci, cj)) * inv_dx);
        else dby = -((negEz(ci, cj) - negEz(ci - 1, cj)) * inv_dx);
        double bxn = g.uin[4][k] + dt * dbx;
        double byn = g.uin[5][k] + dt * dby;

        // outflow boundary flux correction (src/main.f90:116-148), stage primitives
        if (xlo_open || ylo_open) {
            double rho = sP[0 * S::P_SZ + pc], vx = sP[1 * S::P_SZ + pc];
            double vy = sP[2 * S::P_SZ + pc], P = sP[3 * S::P_SZ + pc];
            double h = 0.5 * (Bxc * Bxc + Byc * Byc);
            double en = (P - h) / (gam - 1.0) + 0.5 * rho * (vx * vx + vy * vy) + h;
            if (xlo_open) {
                U[0] = U[0] + dtdx * rho * vx;
                U[1] = U[1] + dtdx * (rho * (vx * vx) + P - Bxc * Bxc);
                U[2] = U[2] + dtdx * (rho * vx * vy - Bxc * Byc);
                U[3] = U[3] + dtdx * ((en + P) * vx - Bxc * (Bxc * vx + Byc * vy));
            }
            if (ylo_open) {
                U[0] = U[0] + dtdx * rho * vy;
                U[2] = U[2] + dtdx * (rho * (vy * vy) + P - Byc * Byc);
                U[1] = U[1] + dtdx * (rho * vy * vx - Byc * Bxc);
                U[3] = U[3] + dtdx * ((en + P) * vy - Byc * (Byc * vy + Bxc * vx));
            }
        }

        if (STORE == STORE_ALL) {
#pragma unroll
            for (int q = 0; q < 4; ++q) store->U3[q][k] = U[q];
        }
        if (stage > 1) {
            const double c0 = (stage == 2) ? 0.75 : 1.0 / 3.0;
            const double c1 = (stage == 2) ? 0.25 : 2.0 / 3.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) U[q] = c0 * g.u0[q][k] + c1 * U[q];
            bxn = c0 * g.u0[4][k] + c1 * bxn;
            byn = c0 * g.u0[5][k] + c1 * byn;
            conserved_floors(U[0], U[1], U[2], U[3], Bxc, Byc, vol, gam);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) g.uout[q][k] = U[q];
        g.uout[4][k] = bxn;
        g.uout[5][k] = byn;
    }
}

}  // namespace chimera
