// Fused reverse of one SSP-RK3 stage (discrete adjoint, frozen dt) on one tile: the
// device functions behind the single-grid adjoint (adjoint.hip) and the batched one
// (batch_adjoint.hip).
#pragma once

#include "adjoint_phys.hpp"
#include "riemann_adj.hpp"
#include "stage_kernel.hpp"

namespace chimera {

// One grid as the reverse kernels see it.
struct AdjGrid {
    int nx, ny, px, py;
    double dx, vol, gamma;
    int use_lim, mood;
};

// The arrays the two fused reverse kernels exchange for one grid, contiguous, n cells
// apart: stage primitives, lam of the four face-state candidates, of the cell resets,
// of the outflow primitives and of the stage's cell B.
struct FusedBufs {
    double *base;
    size_t n;
    static constexpr int ARRAYS = 44;
    __device__ __host__ double *P(int f) const { return base + (size_t)f * n; }
    __device__ __host__ double *XL(int f) const { return base + (size_t)(6 + f) * n; }
    __device__ __host__ double *XR(int f) const { return base + (size_t)(12 + f) * n; }
    __device__ __host__ double *YL(int f) const { return base + (size_t)(18 + f) * n; }
    __device__ __host__ double *YR(int f) const { return base + (size_t)(24 + f) * n; }
    __device__ __host__ double *Pr(int f) const { return base + (size_t)(30 + f) * n; }
    __device__ __host__ double *Pd(int f) const { return base + (size_t)(36 + f) * n; }
    __device__ __host__ double *Bc(int q) const { return base + (size_t)(42 + q) * n; }
};

// Reverse of the thermal check, the rho/P face floor and MOOD for one face state of
// cell k (side s): lam of the face -> lam of the MUSCL candidate (cand, six fields
// `stride` apart) or of the cell value (lreset).
__device__ __forceinline__ void face_state_reverse(double lam[6], int s, bool to_cell,
                                                   unsigned long long bits, double lreset[6],
                                                   double *cand, size_t stride, size_t k) {
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
            cand[(size_t)f * stride + k] = 0.0;
        } else {
            cand[(size_t)f * stride + k] = v;
        }
    }
}

// ---- fused reverse -----------------------------------------------------------------
//
// The reverse in two LDS-tiled passes that recompute the forward stage -- all of it
// but the flux update, whose result Up the segment recompute keeps -- instead of
// reading stored intermediates (the stored-intermediate kernels ar_* of adjoint.hip are
// the cross-check).
//
// rev_tile_a (one TX x TY tile of owned cells per workgroup) replaces the STORE_ALL
// recompute, ar_cell_tail, ar_corner and ar_faces: it loads the primitives of cells
// [-2, TX+2] x [-2, TY+2] together with lam of -Ez at the corners around the owned
// faces (one load latency for both), takes Up of cells [0, TX] x [0, TY] (the state
// after the flux update, before the floors) as the segment recompute stored it,
// recomputes their floors, outflow correction and blend with the reverse of that tail
// (lam of Up into LDS), and finally, per owned cell, the face-state recompute, the
// Riemann VJP and the reverse of the face construction of its x and y faces.  It
// writes what the slope / primitive reverse needs (stage primitives, lam of the
// face-state candidates, lam of the cell resets, of the outflow primitives and of the
// cell B) and the update's and CT's part of lam_in; lam_u0 is accumulated.
//
// rev_tile_b replaces ar_slopes, ar_prims and ar_bavg: a few fields per round it
// gathers the slope cotangents of cells [-1, TX+1] x [-1, TY+1] into LDS, sums each
// cell's own and its neighbours' parts for cells [0, TX] x [0, TY], then reverses
// cons_to_prim and the face -> cell B average and writes lam_in once per cell.  Every
// value is computed with the arithmetic, order and branch of the unfused kernels, so
// both paths give the same bits.

template <int TX, int TY>
struct RevShapeA {
    static constexpr int PW = TX + 5, PH = TY + 5, P_SZ = PW * PH;        // cells [-2, T+2]
    static constexpr int CW = TX + 1, CH = TY + 1, C_SZ = CW * CH;        // lam Up: cells [0, TX] x [0, TY]
    static constexpr int EW = TX + 1, EH = TY + 1, E_SZ = EW * EH;        // lam -Ez: corners [-1, TX) x [-1, TY)
    static constexpr int LDS_DOUBLES = 6 * P_SZ + 4 * C_SZ + E_SZ;
};

// Face states of one face from the primitives in LDS exactly as the forward stage
// builds them (slopes of both cells, MUSCL, MOOD, rho/P floor, thermal check).  b0
// points at the left / lower cell's rho; sA / sT are the strides along / across the
// face normal; returns the MOOD and thermal bits and the floor bits at the positions
// of the stored words (direction dir).
template <int LIMITER>
__device__ __forceinline__ unsigned long long face_pair(const double *b0, int psz, int sA, int sT,
                                                        bool last, bool mood, bool use_lim,
                                                        double inv_dx, double inv_2dx, double hdx,
                                                        int dir, double L[6], double R[6]) {
    double C[6];
    unsigned bits = 0, fbits = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double *b = b0 + q * psz;
        const double m = b[-sA], c = b[0], p = b[sA], pp = b[2 * sA];
        C[q] = c;
        const double sl = cell_slope<LIMITER>(m, c, p, use_lim, inv_dx, inv_2dx);
        const double sr = cell_slope<LIMITER>(c, p, pp, use_lim, inv_dx, inv_2dx);
        unsigned fb;
        bits |= face_states(m, c, p, b[sT], b[-sT], sl, sr, last, mood, q == F_RHO || q == F_P,
                            hdx, L[q], R[q], fb)
                << (4 * q + 2 * dir);
        if (q == F_RHO) fbits |= fb << (2 * dir);
        if (q == F_P) fbits |= fb << (4 + 2 * dir);
    }
    bits |= thermal_pair(L, R, C) << (24 + 2 * dir);
    return (unsigned long long)bits | ((unsigned long long)fbits << 32);
}

// OPAQUE_TID: the thread index made opaque per tile, for a caller that loops over tiles
// (the batched persistent kernel): the compiler would otherwise hoist thread-dependent
// values of the tile out of the loop and keep them live through it (HLLD: 18 registers).
// A one-tile caller leaves it off (there it only blocks folding: 5.6% slower).
template <int SOLVER, int LIMITER, int TX, int TY, bool OPAQUE_TID = false>
__device__ __forceinline__ void rev_tile_a(const GridView &g, const AdjGrid &G, const FusedBufs &F,
                                           const double *const *Up, const double *const *lout,
                                           double *const *lin, double *const *lU0, int stage,
                                           double dt, int tile_x, int tile_y, double *lds) {
    using A = RevShapeA<TX, TY>;
    constexpr int NT = TX * TY;
    int tid = threadIdx.x;
    if constexpr (OPAQUE_TID) asm volatile("" : "+v"(tid));
    const int nx = g.nx, ny = g.ny, px = g.px, py = g.py;
    const int x0 = tile_x * TX, y0 = tile_y * TY;
    double *sP = lds;                          // [6][PH][PW]
    double *sLU = sP + 6 * A::P_SZ;            // [4][CH][CW]
    double *sLE = sLU + 4 * A::C_SZ;           // [EH][EW]
    const double dx = g.dx, vol = g.vol, gam = g.gamma;
    const double inv_dx = 1.0 / dx, inv_2dx = 1.0 / (2.0 * dx), hdx = 0.5 * dx;
    const bool use_lim = G.use_lim != 0, mood = G.mood != 0;
    const bool vol_p2 = pow2_value(vol);   // the divisions by vol are then exact products
    const double ivol = 1.0 / vol;

    // loads of phase 3 (Up as the forward stage left it, u^n and lam_out of cells
    // [0, TX] x [0, TY]), issued first so that they arrive while phase 1 computes
    constexpr int CPT = (A::C_SZ + NT - 1) / NT;
    const double c0 = (stage == 2) ? 0.75 : 1.0 / 3.0;
    const double c1 = (stage == 2) ? 0.25 : 2.0 / 3.0;
    size_t ck[CPT];
    bool cvalid[CPT];
    double cup[CPT][4], cu0[CPT][4], clo[CPT][6], cl0[CPT][6];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int c = tid + u * NT;
        const int gx = x0 + c % A::CW, gy = y0 + c / A::CW;
        cvalid[u] = c < A::C_SZ && !((!px && gx >= nx) || (!py && gy >= ny));
        ck[u] = cvalid[u] ? (size_t)map_index(gx, nx, px) + (size_t)nx * map_index(gy, ny, py) : 0;
        if (cvalid[u]) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                cup[u][q] = Up[q][ck[u]];
                cu0[u][q] = stage > 1 ? g.u0[q][ck[u]] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) clo[u][q] = lout[q][ck[u]];
            // stage 2 adds to the lam_u0 that stage 3 wrote: read now with the other
            // loads (only this thread's owned cell k reads and writes lU0 at k)
            if (stage == 2) {
                const int ci = c % A::CW, cj = c / A::CW;
                if (ci < TX && cj < TY) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) cl0[u][q] = lU0[q][ck[u]];
                }
            }
        }
    }

    // ---- 1: primitives of cells [-2, TX+2] x [-2, TY+2] and lam of -Ez at the corners
    // [-1, TX) x [-1, TY) (as ar_corner); every load of the phase is issued first ---------
    constexpr int PPA = (A::P_SZ + NT - 1) / NT, EPA = (A::E_SZ + NT - 1) / NT;
    double pl[PPA][8];
    size_t pk[PPA];
#pragma unroll
    for (int u = 0; u < PPA; ++u) {
        const int idx = tid + u * NT;
        if (idx >= A::P_SZ) break;
        const int lx = idx % A::PW, ly = idx / A::PW;
        const int gi = map_index(x0 - 2 + lx, nx, px);
        const int gj = map_index(y0 - 2 + ly, ny, py);
        const int gim = gi > 0 ? gi - 1 : (px ? nx - 1 : 0);
        const int gjm = gj > 0 ? gj - 1 : (py ? ny - 1 : 0);
        const size_t k = (size_t)gi + (size_t)nx * gj;
        pk[u] = k;
#pragma unroll
        for (int q = 0; q < 6; ++q) pl[u][q] = g.uin[q][k];
        pl[u][6] = g.uin[4][(size_t)gim + (size_t)nx * gj];
        pl[u][7] = g.uin[5][(size_t)gi + (size_t)nx * gjm];
    }
    // per corner: lam_out of the x faces above and below it, of the y faces right and
    // left of it, and of the faces a clamped edge repeats
    double el[EPA][6];
    {
        const double *lox = lout[4], *loy = lout[5];
#pragma unroll
        for (int u = 0; u < EPA; ++u) {
            const int e = tid + u * NT;
            if (e >= A::E_SZ) break;
            const int gx = x0 + e % A::EW - 1, gy = y0 + e / A::EW - 1;
            if ((!px && (gx < 0 || gx >= nx)) || (!py && (gy < 0 || gy >= ny))) continue;
            const int a = map_index(gx, nx, px), bb = map_index(gy, ny, py);
            const size_t k = (size_t)a + (size_t)nx * bb;
            el[u][0] = lox[k];
            el[u][1] = lox[bb + 1 < ny ? k + nx : (size_t)a];
            el[u][2] = lox[(size_t)a];
            el[u][3] = loy[k];
            el[u][4] = loy[a + 1 < nx ? k + 1 : (size_t)nx * bb];
            el[u][5] = loy[(size_t)nx * bb];
        }
    }
#pragma unroll
    for (int u = 0; u < PPA; ++u) {
        const int idx = tid + u * NT;
        if (idx >= A::P_SZ) break;
        const int lx = idx % A::PW, ly = idx / A::PW;
        const double Bx = 0.5 * (pl[u][4] + pl[u][6]);
        const double By = 0.5 * (pl[u][5] + pl[u][7]);
        double rho, vx, vy, P;
        if (vol_p2)
            cons_to_prim_v<true>(pl[u][0], pl[u][1], pl[u][2], pl[u][3], Bx, By, gam, vol, ivol, rho,
                                 vx, vy, P);
        else
            cons_to_prim_v<false>(pl[u][0], pl[u][1], pl[u][2], pl[u][3], Bx, By, gam, vol, ivol,
                                  rho, vx, vy, P);
        sP[0 * A::P_SZ + idx] = rho;
        sP[1 * A::P_SZ + idx] = vx;
        sP[2 * A::P_SZ + idx] = vy;
        sP[3 * A::P_SZ + idx] = P;
        sP[4 * A::P_SZ + idx] = Bx;
        sP[5 * A::P_SZ + idx] = By;
        const int ci = lx - 2, cj = ly - 2;
        if (ci >= 0 && ci < TX && cj >= 0 && cj < TY && x0 + ci < nx && y0 + cj < ny) {
            const size_t ko = pk[u];
            F.P(0)[ko] = rho;
            F.P(1)[ko] = vx;
            F.P(2)[ko] = vy;
            F.P(3)[ko] = P;
            F.P(4)[ko] = Bx;
            F.P(5)[ko] = By;
        }
    }
    {
        const double s = dt / dx;
        const double cb = stage > 1 ? c1 : 1.0;   // lam of the CT-updated face B
#pragma unroll
        for (int u = 0; u < EPA; ++u) {
            const int e = tid + u * NT;
            if (e >= A::E_SZ) break;
            const int gx = x0 + e % A::EW - 1, gy = y0 + e / A::EW - 1;
            if ((!px && (gx < 0 || gx >= nx)) || (!py && (gy < 0 || gy >= ny))) {
                sLE[e] = 0.0;
                continue;
            }
            const int a = map_index(gx, nx, px), bb = map_index(gy, ny, py);
            double acc = 0.0;
            if (bb >= 1 || py) acc += s * (cb * el[u][0]);
            if (bb + 1 < ny || py) acc -= s * (cb * el[u][1]);
            if (!py) {
                if (bb == 1) acc += s * (cb * el[u][2]);
                if (bb == 0) acc -= s * (cb * el[u][2]);
            }
            if (a >= 1 || px) acc -= s * (cb * el[u][3]);
            if (a + 1 < nx || px) acc += s * (cb * el[u][4]);
            if (!px) {
                if (a == 1) acc -= s * (cb * el[u][5]);
                if (a == 0) acc += s * (cb * el[u][5]);
            }
            sLE[e] = acc;
        }
    }
    __syncthreads();

    // ---- 3: forward tail and its reverse for cells [0, TX] x [0, TY] ------------------
    const double dtdx = dt * dx;
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int c = tid + u * NT;
        if (c >= A::C_SZ) break;
        const int ci = c % A::CW, cj = c / A::CW;
        const int gx = x0 + ci, gy = y0 + cj;
        if (!cvalid[u]) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sLU[q * A::C_SZ + c] = 0.0;
            continue;
        }
        const int gi = map_index(gx, nx, px), gj = map_index(gy, ny, py);
        const size_t k = ck[u];
        const bool xlo_open = (gi == 0) && !px;
        const bool ylo_open = (gj == 0) && !py;
        const double *const Upc = cup[u];
        const int pc = (cj + 2) * A::PW + (ci + 2);
        double P6[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) P6[q] = sP[q * A::P_SZ + pc];
        const double Bx = P6[4], By = P6[5];
        double U3[4] = {Upc[0], Upc[1], Upc[2], Upc[3]};
        conserved_floors_fast(U3[0], U3[1], U3[2], U3[3], Bx, By, vol, gam);
        if (xlo_open || ylo_open) {
            const double rho = P6[0], vx = P6[1], vy = P6[2], P = P6[3];
            const double h = 0.5 * (Bx * Bx + By * By);
            const double en = (P - h) / (gam - 1.0) + 0.5 * rho * (vx * vx + vy * vy) + h;
            if (xlo_open) {
                U3[0] = U3[0] + dtdx * rho * vx;
                U3[1] = U3[1] + dtdx * (rho * (vx * vx) + P - Bx * Bx);
                U3[2] = U3[2] + dtdx * (rho * vx * vy - Bx * By);
                U3[3] = U3[3] + dtdx * ((en + P) * vx - Bx * (Bx * vx + By * vy));
            }
            if (ylo_open) {
                U3[0] = U3[0] + dtdx * rho * vy;
                U3[2] = U3[2] + dtdx * (rho * (vy * vy) + P - By * By);
                U3[1] = U3[1] + dtdx * (rho * vy * vx - By * Bx);
                U3[3] = U3[3] + dtdx * ((en + P) * vy - By * (By * vy + Bx * vx));
            }
        }
        // reverse (as ar_cell_tail)
        const bool owned = ci < TX && cj < TY && gx < nx && gy < ny;
        double lBx = 0.0, lBy = 0.0;
        double l3[4], lb[2];
        if (stage > 1) {
            double V[4], lo[4], lV[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                V[q] = c0 * cu0[u][q] + c1 * U3[q];
                lo[q] = clo[u][q];
            }
            floors_vjp(V, Bx, By, vol, gam, lo, lV, lBx, lBy);
            // stage 3 starts lam_u0 (written), stage 2 adds to it
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (owned) lU0[q][k] = stage == 3 ? c0 * lV[q] : cl0[u][q] + c0 * lV[q];
                l3[q] = c1 * lV[q];
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                if (owned)
                    lU0[4 + q][k] =
                        stage == 3 ? c0 * clo[u][4 + q] : cl0[u][4 + q] + c0 * clo[u][4 + q];
                lb[q] = c1 * clo[u][4 + q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) l3[q] = clo[u][q];
#pragma unroll
            for (int q = 0; q < 2; ++q) lb[q] = clo[u][4 + q];
        }
        double lPd[6] = {0, 0, 0, 0, 0, 0};
        if (xlo_open) outflow_vjp(0, P6, gam, dtdx, l3, lPd);
        if (ylo_open) outflow_vjp(1, P6, gam, dtdx, l3, lPd);
        double lUp[4];
        floors_vjp(Upc, Bx, By, vol, gam, l3, lUp, lBx, lBy);
#pragma unroll
        for (int q = 0; q < 4; ++q) sLU[q * A::C_SZ + c] = lUp[q];
        if (owned) {
#pragma unroll
            for (int q = 0; q < 4; ++q) lin[q][k] = lUp[q];
            lin[4][k] = lb[0];
            lin[5][k] = lb[1];
            F.Bc(0)[k] = lBx;
            F.Bc(1)[k] = lBy;
            if (xlo_open || ylo_open) {   // zero elsewhere: neither written nor read
#pragma unroll
                for (int f = 0; f < 6; ++f) F.Pd(f)[k] = lPd[f];
            }
        }
    }

    __syncthreads();

    // ---- 5: owned faces (as ar_faces): per owned cell its x face, then its y face ------
    {
        const int ci = tid % TX, cj = tid / TX;
        const int gi = x0 + ci, gj = y0 + cj;
        const bool active = gi < nx && gj < ny;
        const size_t k = active ? (size_t)gi + (size_t)nx * gj : 0;
        // lam of the two face states of the face dir of this cell, and its bits
        auto face_rev = [&](int dir, double lL[6], double lR[6]) -> unsigned long long {
            auto lEz = [&](int a, int bb) { return -sLE[(bb + 1) * A::EW + (a + 1)]; };
            const double *b0 = sP + (cj + 2) * A::PW + (ci + 2);
            const bool lastf = dir ? (!py && gj == ny - 1) : (!px && gi == nx - 1);
            double L[6], Rr[6];
            const unsigned long long bits =
                face_pair<LIMITER>(b0, A::P_SZ, dir ? A::PW : 1, dir ? 1 : A::PW, lastf, mood,
                                   use_lim, inv_dx, inv_2dx, hdx, dir, L, Rr);
            const bool has_next = dir ? (gj < ny - 1 || py) : (gi < nx - 1 || px);
            const int cn = dir ? (cj + 1) * A::CW + ci : cj * A::CW + ci + 1;
            const int cc = cj * A::CW + ci;
            double lF[5];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double v = -dtdx * sLU[q * A::C_SZ + cc];
                if (has_next) v += dtdx * sLU[q * A::C_SZ + cn];
                lF[q] = v;
            }
            if (!dir) {
                double v = -0.25 * lEz(ci, cj);
                if (gj >= 1 || py) v += -0.25 * lEz(ci, cj - 1);
                if (!py && gj == ny - 1) v += -0.25 * lEz(ci, cj);
                lF[4] = v;
            } else {
                double v = 0.25 * lEz(ci, cj);
                if (gi >= 1 || px) v += 0.25 * lEz(ci - 1, cj);
                if (!px && gi == nx - 1) v += 0.25 * lEz(ci, cj);
                lF[4] = v;
            }
            const int fu = dir ? 2 : 1, fw = dir ? 1 : 2, fbn = dir ? 5 : 4, fbt = dir ? 4 : 5;
            FaceState fa{L[0], L[fu], L[fw], L[3], L[fbn], L[fbt]};
            FaceState fb{Rr[0], Rr[fu], Rr[fw], Rr[3], Rr[fbn], Rr[fbt]};
            double lFr[5] = {lF[0], dir ? lF[2] : lF[1], dir ? lF[1] : lF[2], lF[3], lF[4]};
            double lLr[6], lRr[6];
            riemann_rev<SOLVER>(fa, fb, gam, lFr, lLr, lRr);
            lL[0] = lLr[0]; lL[fu] = lLr[1]; lL[fw] = lLr[2]; lL[3] = lLr[3]; lL[fbn] = lLr[4]; lL[fbt] = lLr[5];
            lR[0] = lRr[0]; lR[fu] = lRr[1]; lR[fw] = lRr[2]; lR[3] = lRr[3]; lR[fbn] = lRr[4]; lR[fbt] = lRr[5];
            return bits;
        };
        if (active) {
            double lreset[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int dir = 0; dir < 2; ++dir) {
                double lL[6], lR[6];
                const unsigned long long bits = face_rev(dir, lL, lR);
                const bool lastf = dir ? (!py && gj == ny - 1) : (!px && gi == nx - 1);
                face_state_reverse(lL, 2 * dir, false, bits, lreset, dir ? F.YL(0) : F.XL(0), F.n, k);
                face_state_reverse(lR, 2 * dir + 1, lastf, bits, lreset, dir ? F.YR(0) : F.XR(0), F.n, k);
            }
#pragma unroll
            for (int f = 0; f < 6; ++f) F.Pr(f)[k] = lreset[f];
        }
    }
}

template <int TX, int TY>
struct RevShapeB {
    static constexpr int NT = TX * TY;
    static constexpr int PW = TX + 5, PH = TY + 5, P_SZ = PW * PH;   // primitive: cells [-2, T+2]
    static constexpr int GW = TX + 3, GH = TY + 3, G_SZ = GW * GH;   // slope parts: cells [-1, T+1]
    static constexpr int CW = TX + 1, CH = TY + 1, C_SZ = CW * CH;   // cells [0, T]
    // one field's primitive and slope parts, lam of cell B, lam of the owned cells' U
    static constexpr int LDS_DOUBLES = P_SZ + 5 * G_SZ + 2 * C_SZ + 4 * NT;
    static constexpr int CPT = (C_SZ + NT - 1) / NT;                 // cells [0, T] per thread
};

// lU0_add: the step's lam_u0, added to lam_in at the end (stage 1), else nullptr.
// EARLY_TAIL issues the tail's loads with the last round and the fields' loads two
// rounds ahead (a lone small grid, whose few waves per SIMD cannot hide their
// latency); off, each field's loads are issued at the start of its own round and the
// tail's where they are used, which keeps ~110 VGPRs out of the field loop for the
// batched kernel, whose four waves per SIMD hide the latency instead.
template <int LIMITER, int TX, int TY, bool EARLY_TAIL = true>
__device__ __forceinline__ void rev_tile_b(const AdjGrid &G, const double *const *U,
                                           const FusedBufs &F, double *const *lin,
                                           const double *const *lU0_add, int tile_x, int tile_y,
                                           double *lds) {
    using B = RevShapeB<TX, TY>;
    constexpr int NT = B::NT;
    const int tid = threadIdx.x;
    const int nx = G.nx, ny = G.ny, px = G.px, py = G.py;
    const int x0 = tile_x * TX, y0 = tile_y * TY;
    double *sA = lds;                       // one primitive, cells [-2, T+2]
    double *sG = sA + B::P_SZ;              // gmx, gpx, gmy, gpy, lP of one field, cells [-1, T+1]
    double *sB = sG + 5 * B::G_SZ;          // lam of cell B (total), cells [0, T]
    double *sU = sB + 2 * B::C_SZ;          // lam of the owned cells' conserved U (cons_to_prim)
    const double hdx = 0.5 * G.dx, inv_dx = 1.0 / G.dx, inv_2dx = 1.0 / (2.0 * G.dx);
    // valid global cell of a local position, or -1 outside a non-periodic grid
    auto gcell = [&](int lx, int ly, int &gi, int &gj) -> bool {
        const int ux = x0 + lx, uy = y0 + ly;
        if ((!px && (ux < 0 || ux >= nx)) || (!py && (uy < 0 || uy >= ny))) return false;
        gi = map_index(ux, nx, px);
        gj = map_index(uy, ny, py);
        return true;
    };
    double lp6[B::CPT][6];

    // per-thread positions of the primitive items and of the slope items, fixed over the
    // six fields; each field's loads are issued one round ahead of its use
    constexpr int PPT = (B::P_SZ + NT - 1) / NT, GPT = (B::G_SZ + NT - 1) / NT;
    size_t kp[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const int idx = tid + u * NT;
        const int gi = map_index(x0 - 2 + idx % B::PW, nx, px);
        const int gj = map_index(y0 - 2 + idx / B::PW, ny, py);
        kp[u] = idx < B::P_SZ ? (size_t)gi + (size_t)nx * gj : 0;
    }
    struct GItem {
        size_t k, kl, kb;
        bool valid, hl, hb, cxm, cxp, cym, cyp;
    };
    GItem gt[GPT];
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        const int idx = tid + u * NT;
        GItem &t = gt[u];
        int i = 0, j = 0;
        t.valid = idx < B::G_SZ && gcell(idx % B::GW - 1, idx / B::GW - 1, i, j);
        t.k = (size_t)i + (size_t)nx * j;
        t.hl = i > 0 || px;
        t.hb = j > 0 || py;
        t.kl = i > 0 ? t.k - 1 : t.k + nx - 1;
        t.kb = j > 0 ? t.k - nx : t.k + (size_t)nx * (ny - 1);
        t.cxm = !px && i == 0;   // clamped neighbours
        t.cxp = !px && i == nx - 1;
        t.cym = !py && j == 0;
        t.cyp = !py && j == ny - 1;
    }
    constexpr int NBUF = EARLY_TAIL ? 3 : 1;   // rounds of loads in flight + 1
    double pv[NBUF][PPT], rv[NBUF][GPT][6];
    auto load = [&](int f, int buf) {
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            if (tid + u * NT < B::P_SZ) pv[buf][u] = F.P(f)[kp[u]];
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const GItem &t = gt[u];
            if (!t.valid) continue;
            rv[buf][u][0] = F.XL(f)[t.k];
            rv[buf][u][1] = t.hl ? F.XR(f)[t.kl] : 0.0;   // right state of face i-1/2
            rv[buf][u][2] = F.YL(f)[t.k];
            rv[buf][u][3] = t.hb ? F.YR(f)[t.kb] : 0.0;
            rv[buf][u][4] = F.Pr(f)[t.k];
            rv[buf][u][5] = (t.cxm || t.cym) ? F.Pd(f)[t.k] : 0.0;   // open low side
        }
    };
    // the tail's inputs: per cell of [0, T] its B, lam of its cell B and its conserved U;
    // for the owned cell of this thread lam_in (and lam_u0)
    size_t ck[B::CPT];
    bool cvalid[B::CPT];
    double tu[B::CPT][8];
    const int ox = tid % TX, oy = tid / TX;
    const bool oact = x0 + ox < nx && y0 + oy < ny;
    const size_t ko = oact ? (size_t)(x0 + ox) + (size_t)nx * (y0 + oy) : 0;
    double tl[12];
    auto load_owned = [&]() {
        if (oact) {
#pragma unroll
            for (int q = 0; q < 6; ++q) tl[q] = lin[q][ko];
            if (lU0_add) {
#pragma unroll
                for (int q = 0; q < 6; ++q) tl[6 + q] = lU0_add[q][ko];
            }
        }
    };
    auto load_cell = [&](int u) {
        const int c = tid + u * NT;
        int i = 0, j = 0;
        cvalid[u] = c < B::C_SZ && gcell(c % B::CW, c / B::CW, i, j);
        ck[u] = (size_t)i + (size_t)nx * j;
        if (!cvalid[u]) return;
        tu[u][0] = F.P(4)[ck[u]];
        tu[u][1] = F.P(5)[ck[u]];
        tu[u][2] = F.Bc(0)[ck[u]];
        tu[u][3] = F.Bc(1)[ck[u]];
#pragma unroll
        for (int q = 0; q < 4; ++q) tu[u][4 + q] = U[q][ck[u]];
    };
    auto load_cells = [&]() {
#pragma unroll
        for (int u = 0; u < B::CPT; ++u) load_cell(u);
    };
#pragma unroll
    for (int f = 0; f < NBUF - 1; ++f) load(f, f);

#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int cur = f % NBUF;
        if constexpr (NBUF == 1) load(f, 0);
        // the field's stage primitive over cells [-2, T+2] (clamped like the forward)
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            if (tid + u * NT < B::P_SZ) sA[tid + u * NT] = pv[cur][u];
        __syncthreads();
        if (NBUF > 1 && f + NBUF - 1 < 6) load(f + NBUF - 1, (f + NBUF - 1) % NBUF);
        if (f == 5 && EARLY_TAIL) {
            load_cells();
            load_owned();
        }
        // slope parts of cells [-1, T+1] (as ar_slopes)
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const int idx = tid + u * NT;
            if (idx >= B::G_SZ) break;
            const GItem &t = gt[u];
            if (!t.valid) {
#pragma unroll
                for (int c = 0; c < 5; ++c) sG[c * B::G_SZ + idx] = 0.0;
                continue;
            }
            const int lx = idx % B::GW - 1, ly = idx / B::GW - 1;
            const double *a = sA + (ly + 2) * B::PW + (lx + 2);
            const double am = t.cxm ? a[0] : a[-1], ap = t.cxp ? a[0] : a[1];
            const double bm = t.cym ? a[0] : a[-B::PW], bp = t.cyp ? a[0] : a[B::PW];
            const double *r6 = rv[cur][u];
            const double xr = r6[1], yr = r6[3];
            const double lSx = hdx * (r6[0] - xr);
            const double lSy = hdx * (r6[2] - yr);
            double lp = r6[0] + xr + r6[2] + yr + r6[4] + r6[5];
            double dm, d0, dp;
            if (G.use_lim) limited_slope_grad<LIMITER>(am, a[0], ap, inv_dx, dm, d0, dp);
            else { dm = -inv_2dx; d0 = 0.0; dp = inv_2dx; }
            double gm = lSx * dm, gp = lSx * dp;
            lp += lSx * d0;
            if (t.cxm) { lp += gm; gm = 0.0; }
            if (t.cxp) { lp += gp; gp = 0.0; }
            sG[0 * B::G_SZ + idx] = gm;
            sG[1 * B::G_SZ + idx] = gp;
            if (G.use_lim) limited_slope_grad<LIMITER>(bm, a[0], bp, inv_dx, dm, d0, dp);
            else { dm = -inv_2dx; d0 = 0.0; dp = inv_2dx; }
            gm = lSy * dm;
            gp = lSy * dp;
            lp += lSy * d0;
            if (t.cym) { lp += gm; gm = 0.0; }
            if (t.cyp) { lp += gp; gp = 0.0; }
            sG[2 * B::G_SZ + idx] = gm;
            sG[3 * B::G_SZ + idx] = gp;
            sG[4 * B::G_SZ + idx] = lp;
        }
        __syncthreads();
        // own part plus the neighbours' parts for cells [0, T] (as ar_prims)
#pragma unroll
        for (int u = 0; u < B::CPT; ++u) {
            const int c = tid + u * NT;
            if (c >= B::C_SZ) break;
            const int lx = c % B::CW, ly = c / B::CW;
            int i, j;
            if (!gcell(lx, ly, i, j)) {
                lp6[u][f] = 0.0;
                continue;
            }
            const bool hl = i > 0 || px, hr = i < nx - 1 || px;
            const bool hb = j > 0 || py, ht = j < ny - 1 || py;
            const int gidx = (ly + 1) * B::GW + (lx + 1);
            double acc = sG[4 * B::G_SZ + gidx];
            if (hl) acc += sG[1 * B::G_SZ + gidx - 1];
            if (hr) acc += sG[0 * B::G_SZ + gidx + 1];
            if (hb) acc += sG[3 * B::G_SZ + gidx - B::GW];
            if (ht) acc += sG[2 * B::G_SZ + gidx + B::GW];
            lp6[u][f] = acc;
        }
        // the next field overwrites sA and sG; the tail uses sB and sU only
        if (f < 5) __syncthreads();
    }
    // reverse of cons_to_prim for cells [0, T]; lam of the owned cells' U into sU
#pragma unroll
    for (int u = 0; u < B::CPT; ++u) {
        const int c = tid + u * NT;
        if (c >= B::C_SZ) break;
        if (!EARLY_TAIL) load_cell(u);
        const int lx = c % B::CW, ly = c / B::CW;
        if (!cvalid[u]) {
            sB[c] = 0.0;
            sB[B::C_SZ + c] = 0.0;
            continue;
        }
        double lU[4] = {0, 0, 0, 0};
        double lBx = lp6[u][4] + tu[u][2], lBy = lp6[u][5] + tu[u][3];
        cons_to_prim_vjp(tu[u][4], tu[u][5], tu[u][6], tu[u][7], tu[u][0], tu[u][1], G.gamma,
                         G.vol, lp6[u], lU, lBx, lBy);
        if (lx < TX && ly < TY && x0 + lx < nx && y0 + ly < ny) {
#pragma unroll
            for (int q = 0; q < 4; ++q) sU[q * NT + ly * TX + lx] = lU[q];
        }
        sB[c] = lBx;
        sB[B::C_SZ + c] = lBy;
    }
    __syncthreads();
    if (!EARLY_TAIL) load_owned();
    // face -> cell B average (as ar_bavg); lam_in of the owned cell written once:
    // ((lam_in + cell part) + lam_u0)
    if (oact) {
        const int i = x0 + ox, j = y0 + oy;
        const int c = oy * B::CW + ox;
        const double *lbx = sB, *lby = sB + B::C_SZ;
        double ax = (i == 0 && !px) ? lbx[c] : 0.5 * lbx[c];
        if (i + 1 < nx || px) ax += 0.5 * lbx[c + 1];
        double ay = (j == 0 && !py) ? lby[c] : 0.5 * lby[c];
        if (j + 1 < ny || py) ay += 0.5 * lby[c + B::CW];
        double v[6];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = tl[q] + sU[q * NT + tid];
        v[4] = tl[4] + ax;
        v[5] = tl[5] + ay;
        if (lU0_add) {
#pragma unroll
            for (int q = 0; q < 6; ++q) v[q] += tl[6 + q];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) lin[q][ko] = v[q];
    }
}

}  // namespace chimera
