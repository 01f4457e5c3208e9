// Discrete adjoint of the SSP-RK3 step (frozen dt): hand-written reverse kernels for
// the reconstruction (limiter, MOOD, face floors, thermal check), the Riemann fluxes,
// the conservative update, constrained transport, the outflow correction, the conserved
// floors and the RK blends.
//
// Per stage, the fused LDS-tiled stage kernel (stage_kernel.hpp, STORE instantiation)
// recomputes the stage exactly as the forward path does and keeps what the reverse
// needs -- primitives, face states after MOOD and the thermal check, branch bits, the
// pre-floor and pre-blend states.  Reverse kernels ar_* are written in gather form
// (each thread owns one output location and collects the contributions of every
// forward operation that read it), so there are no atomics and the gradient is
// deterministic.  Branches (limiter cases, MOOD resets, face and conserved floors,
// thermal resets, Riemann-solver regions and catches) are the ones the forward pass
// took.  Rusanov and HLLE are reversed by hand (riemann_adj.hpp); HLLD Jacobians come
// from forward-mode dual numbers (riemann_dual.hpp).
#include <vector>

#include "adjoint_phys.hpp"
#include "common.hpp"
#include "resident.hpp"
#include "riemann_adj.hpp"

namespace chimera {
namespace {

constexpr int BS = 256;

struct AdjGrid {
    int nx, ny, px, py;
    double dx, vol, gamma;
    int use_lim, mood;
};

using StageBufs = StageStore;

struct RevBufs {
    double *lUp[4], *lBp[2], *lBc[2], *lPd[6], *lnegEz;
    double *lXL[6], *lXR[6], *lYL[6], *lYR[6], *lPr[6];
    double *gmx[6], *gpx[6], *gmy[6], *gpy[6];  // limiter parts passed to the lo / hi neighbour
    double *lP[6];
    double *lBcT[2];
};

__device__ __forceinline__ int wrap_or_clamp(int g, int n, int periodic) {
    if (periodic) return g < 0 ? g + n : (g >= n ? g - n : g);
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

#define CELL_LOOP(k, n) \
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (n); \
         k += (size_t)gridDim.x * blockDim.x)

// ---- forward recompute: the fused stage kernel, keeping the intermediates -----------

constexpr int ATX = 32, ATY = 8;
// tile shapes of the fused reverse kernels (chimera_adj_set_tiles); smaller tiles give
// more workgroups on small grids at the price of more halo work
struct RevTile {
    int tx, ty;
};
constexpr int REV_NUM_SHAPES = 4;
constexpr RevTile REV_SHAPES[REV_NUM_SHAPES] = {{16, 16}, {16, 8}, {8, 8}, {32, 8}};

template <int SOLVER, int LIMITER, bool STORE>
__global__ __launch_bounds__(STAGE_THREADS) void adj_stage_kernel(GridView g, StageStore st,
                                                                 int stage, int use_lim, int mood,
                                                                 int tiles_x, int ntiles) {
    __shared__ double lds[TileShape<ATX, ATY>::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;  // nb is a multiple of 8
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    stage_tile<SOLVER, LIMITER, ATX, ATY, STORE>(g, stage, use_lim != 0, mood != 0, tile % tiles_x,
                                                 tile / tiles_x, lds, STORE ? &st : nullptr);
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
    riemann_rev<SOLVER>(a, b, G.gamma, lFr, lLr, lRr);
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
        unsigned long long bits = S.bits_x[k] | S.bits_y[k];
        double lreset[6] = {0, 0, 0, 0, 0, 0};
        face_reverse<SOLVER, true>(G, S, R, dt, i, j, k, bits, lreset);
        face_reverse<SOLVER, false>(G, S, R, dt, i, j, k, bits, lreset);
#pragma unroll
        for (int f = 0; f < 6; ++f) R.lPr[f][k] = lreset[f];
    }
}

// lam of each limited slope (from the face states built with it), its limiter gradient
// (once per slope), the parts passed to the lo / hi neighbour (gm / gp) and the cell's
// own part of lam of its primitives.
template <int LIMITER>
__global__ void ar_slopes(AdjGrid G, StageStore S, RevBufs R) {
    const int nx = G.nx, ny = G.ny;
    const double hdx = 0.5 * G.dx, inv_dx = 1.0 / G.dx, inv_2dx = 1.0 / (2.0 * G.dx);
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        bool hl = i > 0 || G.px, hb = j > 0 || G.py;   // a face on the lo side exists
        size_t kl = i > 0 ? k - 1 : k + nx - 1;
        size_t kb = j > 0 ? k - nx : k + (size_t)nx * (ny - 1);
        // neighbours the slopes of (i, j) read: periodic wrap or zero-gradient clamp
        size_t kxm = (size_t)wrap_or_clamp(i - 1, nx, G.px) + (size_t)nx * j;
        size_t kxp = (size_t)wrap_or_clamp(i + 1, nx, G.px) + (size_t)nx * j;
        size_t kym = (size_t)i + (size_t)nx * wrap_or_clamp(j - 1, ny, G.py);
        size_t kyp = (size_t)i + (size_t)nx * wrap_or_clamp(j + 1, ny, G.py);
        for (int f = 0; f < 6; ++f) {
            const double *a = S.P[f];
            double xr = hl ? R.lXR[f][kl] : 0.0;   // the right state of face i-1/2 came from (i,j)
            double yr = hb ? R.lYR[f][kb] : 0.0;
            double lSx = hdx * (R.lXL[f][k] - xr);
            double lSy = hdx * (R.lYL[f][k] - yr);
            double lp = R.lXL[f][k] + xr + R.lYL[f][k] + yr + R.lPr[f][k] + R.lPd[f][k];
            double dm, d0, dp;
            // x slope
            if (G.use_lim) limited_slope_grad<LIMITER>(a[kxm], a[k], a[kxp], inv_dx, dm, d0, dp);
            else { dm = -inv_2dx; d0 = 0.0; dp = inv_2dx; }
            double gm = lSx * dm, gp = lSx * dp;
            lp += lSx * d0;
            if (kxm == k) { lp += gm; gm = 0.0; }   // clamped at a non-periodic edge
            if (kxp == k) { lp += gp; gp = 0.0; }
            R.gmx[f][k] = gm;
            R.gpx[f][k] = gp;
            // y slope
            if (G.use_lim) limited_slope_grad<LIMITER>(a[kym], a[k], a[kyp], inv_dx, dm, d0, dp);
            else { dm = -inv_2dx; d0 = 0.0; dp = inv_2dx; }
            gm = lSy * dm;
            gp = lSy * dp;
            lp += lSy * d0;
            if (kym == k) { lp += gm; gm = 0.0; }
            if (kyp == k) { lp += gp; gp = 0.0; }
            R.gmy[f][k] = gm;
            R.gpy[f][k] = gp;
            R.lP[f][k] = lp;
        }
    }
}

// lam of the primitives (own part + the neighbours' slope parts), then the reverse of
// cons_to_prim; writes lam of the conserved stage input and the total lam of cell B.
__global__ void ar_prims(AdjGrid G, const double *const *U, StageStore S, RevBufs R,
                         double *const *lin) {
    const int nx = G.nx, ny = G.ny;
    CELL_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        bool hl = i > 0 || G.px, hr = i < nx - 1 || G.px;
        bool hb = j > 0 || G.py, ht = j < ny - 1 || G.py;
        size_t kl = i > 0 ? k - 1 : k + nx - 1;
        size_t kr = i < nx - 1 ? k + 1 : k - (nx - 1);
        size_t kb = j > 0 ? k - nx : k + (size_t)nx * (ny - 1);
        size_t kt = j < ny - 1 ? k + nx : k - (size_t)nx * (ny - 1);
        double lp6[6];
        for (int f = 0; f < 6; ++f) {
            double acc = R.lP[f][k];
            if (hl) acc += R.gpx[f][kl];
            if (hr) acc += R.gmx[f][kr];
            if (hb) acc += R.gpy[f][kb];
            if (ht) acc += R.gmy[f][kt];
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

// ---- fused reverse -----------------------------------------------------------------
//
// The same reverse in two LDS-tiled kernels that recompute the forward stage instead of
// reading stored intermediates.
//
// ar_tile_a (one TX x TY tile of owned cells per workgroup) replaces the STORE
// recompute, ar_cell_tail, ar_corner and ar_faces: it loads the primitives of cells
// [-2, TX+2] x [-2, TY+2], computes the fluxes of the faces of cells [0, TX] x [0, TY]
// (forward phase 2), the flux update, floors, outflow correction and blend of those
// cells with the reverse of the tail (lam of Up into LDS), lam of -Ez at the corners
// around the owned faces, and finally, per owned cell, the face-state recompute, the
// Riemann VJP and the reverse of the face construction of its x and y faces.  It
// writes what the slope / primitive reverse needs (stage primitives, lam of the
// face-state candidates, lam of the cell resets, of the outflow primitives and of the
// cell B) and the update's and CT's part of lam_in; lam_u0 is accumulated.
//
// ar_tile_b replaces ar_slopes, ar_prims and ar_bavg: field by field it gathers the
// slope cotangents of cells [-1, TX+1] x [-1, TY+1] into LDS, sums each cell's own and
// its neighbours' parts for cells [0, TX] x [0, TY], then reverses cons_to_prim and the
// face -> cell B average.  Every value is computed with the arithmetic, order and
// branch of the unfused kernels, so both paths give the same bits.

template <int TX, int TY>
struct RevShapeA {
    static constexpr int PW = TX + 5, PH = TY + 5, P_SZ = PW * PH;        // cells [-2, T+2]
    static constexpr int FXW = TX + 2, FXH = TY + 1, FX_SZ = FXW * FXH;   // x faces: left cell [-1, TX] x [0, TY]
    static constexpr int FYW = TX + 1, FYH = TY + 2, FY_SZ = FYW * FYH;   // y faces: lower cell [0, TX] x [-1, TY]
    static constexpr int CW = TX + 1, CH = TY + 1, C_SZ = CW * CH;        // lam Up: cells [0, TX] x [0, TY]
    static constexpr int EW = TX + 1, EH = TY + 1, E_SZ = EW * EH;        // lam -Ez: corners [-1, TX) x [-1, TY)
    static constexpr int LDS_DOUBLES = 6 * P_SZ + 4 * FX_SZ + 4 * FY_SZ + 4 * C_SZ + E_SZ;
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
    if (thermal_reset(L, C)) bits |= 1u << (24 + 2 * dir);
    if (thermal_reset(R, C)) bits |= 1u << (25 + 2 * dir);
    return (unsigned long long)bits | ((unsigned long long)fbits << 32);
}

template <int SOLVER, int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void ar_tile_a(GridView g, AdjGrid G, StageBufs S, RevBufs R,
                                                   const double *const *lout, double *const *lin,
                                                   double *const *lU0, int stage, double dt,
                                                   int tiles_x, int ntiles) {
    using A = RevShapeA<TX, TY>;
    constexpr int NT = TX * TY;
    __shared__ double lds[A::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;   // nb is a multiple of 8
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    const int tid = threadIdx.x;
    const int nx = g.nx, ny = g.ny, px = g.px, py = g.py;
    const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
    double *sP = lds;                          // [6][PH][PW]
    double *sFX = sP + 6 * A::P_SZ;            // [4][FXH][FXW] mass, Momx, Momy, energy
    double *sFY = sFX + 4 * A::FX_SZ;          // [4][FYH][FYW]
    double *sLU = sFY + 4 * A::FY_SZ;          // [4][CH][CW]
    double *sLE = sLU + 4 * A::C_SZ;           // [EH][EW]
    const double dx = g.dx, vol = g.vol, gam = g.gamma;
    const double inv_dx = 1.0 / dx, inv_2dx = 1.0 / (2.0 * dx), hdx = 0.5 * dx;
    const bool use_lim = G.use_lim != 0, mood = G.mood != 0;

    // loads of phase 3 (the stage input, u^n and lam_out of cells [0, TX] x [0, TY]),
    // issued first so that they arrive while phases 1 and 2 compute
    constexpr int CPT = (A::C_SZ + NT - 1) / NT;
    const double c0 = (stage == 2) ? 0.75 : 1.0 / 3.0;
    const double c1 = (stage == 2) ? 0.25 : 2.0 / 3.0;
    size_t ck[CPT];
    bool cvalid[CPT];
    double cin[CPT][4], cu0[CPT][4], clo[CPT][6];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
        const int c = tid + u * NT;
        const int gx = x0 + c % A::CW, gy = y0 + c / A::CW;
        cvalid[u] = c < A::C_SZ && !((!px && gx >= nx) || (!py && gy >= ny));
        ck[u] = cvalid[u] ? (size_t)map_index(gx, nx, px) + (size_t)nx * map_index(gy, ny, py) : 0;
        if (cvalid[u]) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                cin[u][q] = g.uin[q][ck[u]];
                cu0[u][q] = stage > 1 ? g.u0[q][ck[u]] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) clo[u][q] = lout[q][ck[u]];
        }
    }

    // ---- 1: primitives of cells [-2, TX+2] x [-2, TY+2] ------------------------------
    for (int idx = tid; idx < A::P_SZ; idx += NT) {
        const int lx = idx % A::PW, ly = idx / A::PW;
        const int gi = map_index(x0 - 2 + lx, nx, px);
        const int gj = map_index(y0 - 2 + ly, ny, py);
        const int gim = gi > 0 ? gi - 1 : (px ? nx - 1 : 0);
        const int gjm = gj > 0 ? gj - 1 : (py ? ny - 1 : 0);
        const size_t k = (size_t)gi + (size_t)nx * gj;
        const double Bx = 0.5 * (g.uin[4][k] + g.uin[4][(size_t)gim + (size_t)nx * gj]);
        const double By = 0.5 * (g.uin[5][k] + g.uin[5][(size_t)gi + (size_t)nx * gjm]);
        double rho, vx, vy, P;
        cons_to_prim(g.uin[0][k], g.uin[1][k], g.uin[2][k], g.uin[3][k], Bx, By, gam, vol, rho, vx,
                     vy, P);
        sP[0 * A::P_SZ + idx] = rho;
        sP[1 * A::P_SZ + idx] = vx;
        sP[2 * A::P_SZ + idx] = vy;
        sP[3 * A::P_SZ + idx] = P;
        sP[4 * A::P_SZ + idx] = Bx;
        sP[5 * A::P_SZ + idx] = By;
        const int ci = lx - 2, cj = ly - 2;
        if (ci >= 0 && ci < TX && cj >= 0 && cj < TY && x0 + ci < nx && y0 + cj < ny) {
            const size_t ko = (size_t)(x0 + ci) + (size_t)nx * (y0 + cj);
            S.P[0][ko] = rho;
            S.P[1][ko] = vx;
            S.P[2][ko] = vy;
            S.P[3][ko] = P;
            S.P[4][ko] = Bx;
            S.P[5][ko] = By;
        }
    }
    __syncthreads();

    // ---- 2: fluxes of the faces of cells [0, TX] x [0, TY] ---------------------------
    for (int f = tid; f < A::FX_SZ + A::FY_SZ; f += NT) {
        const int dir = f >= A::FX_SZ;
        int i, j;
        if (!dir) {
            i = f % A::FXW - 1;
            j = f / A::FXW;
        } else {
            i = (f - A::FX_SZ) % A::FYW;
            j = (f - A::FX_SZ) / A::FYW - 1;
        }
        const int lpos = (j + 2) * A::PW + (i + 2);
        const bool last = dir ? (!py && map_index(y0 + j, ny, py) == ny - 1)
                              : (!px && map_index(x0 + i, nx, px) == nx - 1);
        double L[6], Rs[6];
        (void)face_pair<LIMITER>(sP + lpos, A::P_SZ, dir ? A::PW : 1, dir ? 1 : A::PW, last, mood,
                                 use_lim, inv_dx, inv_2dx, hdx, dir, L, Rs);
        FaceState fa{L[F_RHO], dir ? L[F_VY] : L[F_VX], dir ? L[F_VX] : L[F_VY], L[F_P],
                     dir ? L[F_BY] : L[F_BX], dir ? L[F_BX] : L[F_BY]};
        FaceState fb{Rs[F_RHO], dir ? Rs[F_VY] : Rs[F_VX], dir ? Rs[F_VX] : Rs[F_VY], Rs[F_P],
                     dir ? Rs[F_BY] : Rs[F_BX], dir ? Rs[F_BX] : Rs[F_BY]};
        const Flux5 fl = riemann<SOLVER>(fa, fb, gam);
        double *sF = dir ? sFY : sFX;
        const int fsz = dir ? A::FY_SZ : A::FX_SZ;
        const int idx = dir ? f - A::FX_SZ : f;
        sF[0 * fsz + idx] = fl.mass;
        sF[1 * fsz + idx] = dir ? fl.mom_t : fl.mom_n;
        sF[2 * fsz + idx] = dir ? fl.mom_n : fl.mom_t;
        sF[3 * fsz + idx] = fl.energy;
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
        const int xr = cj * A::FXW + (ci + 1);         // x face right of the cell
        const int yr = (cj + 1) * A::FYW + ci;         // y face above the cell
        const bool xlo_open = (gi == 0) && !px;
        const bool ylo_open = (gj == 0) && !py;
        double Up[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double v = cin[u][q];
            const double fr = sFX[q * A::FX_SZ + xr];
            if (xlo_open) v = v - dtdx * fr;
            else v = v - dtdx * fr + dtdx * sFX[q * A::FX_SZ + xr - 1];
            const double ft = sFY[q * A::FY_SZ + yr];
            if (ylo_open) v = v - dtdx * ft;
            else v = v - dtdx * ft + dtdx * sFY[q * A::FY_SZ + yr - A::FYW];
            Up[q] = v;
        }
        const int pc = (cj + 2) * A::PW + (ci + 2);
        double P6[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) P6[q] = sP[q * A::P_SZ + pc];
        const double Bx = P6[4], By = P6[5];
        double U3[4] = {Up[0], Up[1], Up[2], Up[3]};
        conserved_floors(U3[0], U3[1], U3[2], U3[3], Bx, By, vol, gam);
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
                if (owned) {
                    if (stage == 3) lU0[q][k] = c0 * lV[q];
                    else lU0[q][k] += c0 * lV[q];
                }
                l3[q] = c1 * lV[q];
            }
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                if (owned) {
                    if (stage == 3) lU0[4 + q][k] = c0 * clo[u][4 + q];
                    else lU0[4 + q][k] += c0 * clo[u][4 + q];
                }
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
        floors_vjp(Up, Bx, By, vol, gam, l3, lUp, lBx, lBy);
#pragma unroll
        for (int q = 0; q < 4; ++q) sLU[q * A::C_SZ + c] = lUp[q];
        if (owned) {
#pragma unroll
            for (int q = 0; q < 4; ++q) lin[q][k] = lUp[q];
            lin[4][k] = lb[0];
            lin[5][k] = lb[1];
            R.lBc[0][k] = lBx;
            R.lBc[1][k] = lBy;
#pragma unroll
            for (int f = 0; f < 6; ++f) R.lPd[f][k] = lPd[f];
        }
    }

    // ---- 4: lam of -Ez at corners [-1, TX) x [-1, TY) (as ar_corner) -----------------
    {
        const double s = dt / dx;
        const double cb = stage > 1 ? c1 : 1.0;   // lam of the CT-updated face B
        const double *lox = lout[4], *loy = lout[5];
        for (int e = tid; e < A::E_SZ; e += NT) {
            const int la = e % A::EW - 1, lb_ = e / A::EW - 1;
            const int gx = x0 + la, gy = y0 + lb_;
            if ((!px && (gx < 0 || gx >= nx)) || (!py && (gy < 0 || gy >= ny))) {
                sLE[e] = 0.0;
                continue;
            }
            const int a = map_index(gx, nx, px), bb = map_index(gy, ny, py);
            const size_t k = (size_t)a + (size_t)nx * bb;
            double acc = 0.0;
            if (bb >= 1 || py) acc += s * (cb * lox[k]);
            {
                const int jn = bb + 1;
                if (jn < ny) acc -= s * (cb * lox[k + nx]);
                else if (py) acc -= s * (cb * lox[(size_t)a]);
            }
            if (!py) {
                if (bb == 1) acc += s * (cb * lox[(size_t)a]);
                if (bb == 0) acc -= s * (cb * lox[(size_t)a]);
            }
            if (a >= 1 || px) acc -= s * (cb * loy[k]);
            {
                const int in = a + 1;
                if (in < nx) acc += s * (cb * loy[k + 1]);
                else if (px) acc += s * (cb * loy[(size_t)nx * bb]);
            }
            if (!px) {
                if (a == 1) acc -= s * (cb * loy[(size_t)nx * bb]);
                if (a == 0) acc += s * (cb * loy[(size_t)nx * bb]);
            }
            sLE[e] = acc;
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
                face_state_reverse(lL, 2 * dir, false, bits, lreset, dir ? R.lYL : R.lXL, k);
                face_state_reverse(lR, 2 * dir + 1, lastf, bits, lreset, dir ? R.lYR : R.lXR, k);
            }
#pragma unroll
            for (int f = 0; f < 6; ++f) R.lPr[f][k] = lreset[f];
        }
    }
}

template <int TX, int TY>
struct RevShapeB {
    static constexpr int PW = TX + 5, PH = TY + 5, P_SZ = PW * PH;   // primitive: cells [-2, T+2]
    static constexpr int GW = TX + 3, GH = TY + 3, G_SZ = GW * GH;   // slope parts: cells [-1, T+1]
    static constexpr int CW = TX + 1, CH = TY + 1, C_SZ = CW * CH;   // cells [0, T]
    static constexpr int LDS_DOUBLES = P_SZ + 5 * G_SZ + 2 * C_SZ;
    static constexpr int CPT = (C_SZ + TX * TY - 1) / (TX * TY);     // cells [0, T] per thread
};

// lU0_add: the step's lam_u0, added to lam_in at the end (stage 1), else nullptr
template <int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void ar_tile_b(AdjGrid G, const double *const *U, StageStore S,
                                                   RevBufs R, double *const *lin,
                                                   const double *const *lU0_add, int tiles_x,
                                                   int ntiles) {
    using B = RevShapeB<TX, TY>;
    constexpr int NT = TX * TY;
    __shared__ double lds[B::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    const int tid = threadIdx.x;
    const int nx = G.nx, ny = G.ny, px = G.px, py = G.py;
    const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
    double *sA = lds;                       // one primitive, cells [-2, T+2]
    double *sG = sA + B::P_SZ;              // gmx, gpx, gmy, gpy, lP of one field, cells [-1, T+1]
    double *sB = sG + 5 * B::G_SZ;          // lam of cell B (total), cells [0, T]
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
    double pv[2][PPT], rv[2][GPT][6];
    auto load = [&](int f, int buf) {
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            if (tid + u * NT < B::P_SZ) pv[buf][u] = S.P[f][kp[u]];
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const GItem &t = gt[u];
            if (!t.valid) continue;
            rv[buf][u][0] = R.lXL[f][t.k];
            rv[buf][u][1] = t.hl ? R.lXR[f][t.kl] : 0.0;   // right state of face i-1/2
            rv[buf][u][2] = R.lYL[f][t.k];
            rv[buf][u][3] = t.hb ? R.lYR[f][t.kb] : 0.0;
            rv[buf][u][4] = R.lPr[f][t.k];
            rv[buf][u][5] = R.lPd[f][t.k];
        }
    };
    load(0, 0);

#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int cur = f & 1;
        // the field's stage primitive over cells [-2, T+2] (clamped like the forward)
#pragma unroll
        for (int u = 0; u < PPT; ++u)
            if (tid + u * NT < B::P_SZ) sA[tid + u * NT] = pv[cur][u];
        __syncthreads();
        if (f < 5) load(f + 1, cur ^ 1);
        // slope parts of cells [-1, T+1] (as ar_slopes)
#pragma unroll
        for (int u = 0; u < GPT; ++u) {
            const int idx = tid + u * NT;
            if (idx >= B::G_SZ) break;
            const GItem &t = gt[u];
            if (!t.valid) {
#pragma unroll
                for (int r = 0; r < 5; ++r) sG[r * B::G_SZ + idx] = 0.0;
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
        __syncthreads();
    }
    // reverse of cons_to_prim for cells [0, T]; lam of the conserved input for owned cells
#pragma unroll
    for (int u = 0; u < B::CPT; ++u) {
        const int c = tid + u * NT;
        if (c >= B::C_SZ) break;
        const int lx = c % B::CW, ly = c / B::CW;
        int i, j;
        if (!gcell(lx, ly, i, j)) {
            sB[c] = 0.0;
            sB[B::C_SZ + c] = 0.0;
            continue;
        }
        const size_t k = (size_t)i + (size_t)nx * j;
        const double Bx = S.P[4][k], By = S.P[5][k];
        double lU[4] = {0, 0, 0, 0};
        double lBx = lp6[u][4] + R.lBc[0][k], lBy = lp6[u][5] + R.lBc[1][k];
        cons_to_prim_vjp(U[0][k], U[1][k], U[2][k], U[3][k], Bx, By, G.gamma, G.vol, lp6[u], lU,
                         lBx, lBy);
        if (lx < TX && ly < TY && x0 + lx < nx && y0 + ly < ny) {
#pragma unroll
            for (int q = 0; q < 4; ++q) lin[q][k] += lU[q];
        }
        sB[c] = lBx;
        sB[B::C_SZ + c] = lBy;
    }
    __syncthreads();
    // face -> cell B average (as ar_bavg), owned cells
    {
        const int lx = tid % TX, ly = tid / TX;
        const int i = x0 + lx, j = y0 + ly;
        if (i < nx && j < ny) {
            const size_t k = (size_t)i + (size_t)nx * j;
            const int c = ly * B::CW + lx;
            const double *lbx = sB, *lby = sB + B::C_SZ;
            double ax = (i == 0 && !px) ? lbx[c] : 0.5 * lbx[c];
            if (i + 1 < nx || px) ax += 0.5 * lbx[c + 1];
            double ay = (j == 0 && !py) ? lby[c] : 0.5 * lby[c];
            if (j + 1 < ny || py) ay += 0.5 * lby[c + B::CW];
            lin[4][k] += ax;
            lin[5][k] += ay;
            if (lU0_add) {
#pragma unroll
                for (int q = 0; q < 6; ++q) lin[q][k] += lU0_add[q][k];
            }
        }
    }
}

// Face VJPs of one solver for n faces (x direction), by hand-written reverse (method 0)
// or forward-mode dual numbers (method 1): the cross-check of the hand-written
// solvers.  in / out hold (rho, u, w, P, Bn, Bt) as L, R pairs: [2*q + side][n].
template <int SOLVER>
__global__ void face_vjp_kernel(int method, int n, double g, const double *in, const double *lamF,
                                double *out) {
    CELL_LOOP(k, (size_t)n) {
        const FaceState L{in[0 * (size_t)n + k], in[2 * (size_t)n + k], in[4 * (size_t)n + k],
                          in[6 * (size_t)n + k], in[8 * (size_t)n + k], in[10 * (size_t)n + k]};
        const FaceState R{in[1 * (size_t)n + k], in[3 * (size_t)n + k], in[5 * (size_t)n + k],
                          in[7 * (size_t)n + k], in[9 * (size_t)n + k], in[11 * (size_t)n + k]};
        double lam[5], lL[6], lR[6];
        for (int c = 0; c < 5; ++c) lam[c] = lamF[(size_t)c * n + k];
        if (method == 0) riemann_rev<SOLVER>(L, R, g, lam, lL, lR);
        else riemann_vjp<SOLVER, 1>(L, R, g, lam, lL, lR);
        for (int q = 0; q < 6; ++q) {
            out[(size_t)(2 * q) * n + k] = lL[q];
            out[(size_t)(2 * q + 1) * n + k] = lR[q];
        }
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
    bool fused = true;                        // fused reverse (ar_tile_a / ar_tile_b)
    int rshape_a = 3, rshape_b = 3;           // their tile shapes (REV_SHAPES)
    chimera_scheme_t sch{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t n = 0;
    int tiles_x = 0, ntiles = 0, grid = 0;   // STORE stage kernel launch shape
    double *arena = nullptr;
    StageBufs S[3];
    RevBufs R;
    Ptr6 stage_state[3];   // U^n (copied by the one-step calls), U^(1), U^(2)
    Ptr6 lam_a, lam_b, lam_c, lam_u0, scratch_out;
    // device-side pointer tables (reverse kernels take double *const *)
    double **dptr = nullptr;
    // dt of the steps being processed (device; the stage kernel reads *dt)
    double *dt_dev = nullptr;
    int dt_cap = 0;
    // checkpoint recomputation: the recomputed stage inputs
    double *traj = nullptr;       // traj_cap x 6 fields
    double **traj_tab = nullptr;  // device tables, 6 pointers per slot
    int traj_cap = 0;
};

namespace {

// device pointer tables: [0..2] stage_state, [3] lam_a, [4] lam_b, [5] lam_u0,
// [6] scratch_out, [7] caller input, [8] caller output, [9] caller lam in,
// [10] caller lam out, [11] lam_c
enum { TB_S0 = 0, TB_S1 = 1, TB_S2 = 2, TB_LA = 3, TB_LB = 4, TB_LU0 = 5, TB_SCR = 6, TB_IN = 7,
       TB_OUT = 8, TB_LIN = 9, TB_LOUT = 10, TB_LC = 11, TB_COUNT = 12 };

using Tab = double *const *;

Tab table(chimera_adj *a, int t) { return a->dptr + 6 * t; }

// The caller's pointer arrays may be temporaries: copy them synchronously (after
// the work already queued on the stream).
int set_table(chimera_adj *a, int t, const double *const *p) {
    CH_HIP(hipStreamSynchronize(a->stream));
    CH_HIP(hipMemcpy(a->dptr + 6 * t, p, 6 * sizeof(double *), hipMemcpyHostToDevice));
    return CHIMERA_OK;
}

Ptr6 ptr6(const double *const *p) {
    Ptr6 r;
    for (int q = 0; q < 6; ++q) r.p[q] = const_cast<double *>(p[q]);
    return r;
}

// dts of the steps to come, on the device (synchronous upload)
int set_dts(chimera_adj *a, const double *dts, int k) {
    CH_HIP(hipStreamSynchronize(a->stream));
    if (k > a->dt_cap) {
        (void)hipFree(a->dt_dev);
        a->dt_dev = nullptr;
        a->dt_cap = 0;
        CH_HIP(hipMalloc(&a->dt_dev, (size_t)k * sizeof(double)));
        a->dt_cap = k;
    }
    CH_HIP(hipMemcpy(a->dt_dev, dts, (size_t)k * sizeof(double), hipMemcpyHostToDevice));
    return CHIMERA_OK;
}

using AdjStageFn = void (*)(GridView, StageStore, int, int, int, int, int);

AdjStageFn adj_stage_fn(int solver, int limiter, bool store) {
    static const AdjStageFn t[2][3][2] = {
        {{adj_stage_kernel<RUSANOV, LIM_VAN_LEER, false>, adj_stage_kernel<RUSANOV, LIM_MC, false>},
         {adj_stage_kernel<HLLE, LIM_VAN_LEER, false>, adj_stage_kernel<HLLE, LIM_MC, false>},
         {adj_stage_kernel<HLLD, LIM_VAN_LEER, false>, adj_stage_kernel<HLLD, LIM_MC, false>}},
        {{adj_stage_kernel<RUSANOV, LIM_VAN_LEER, true>, adj_stage_kernel<RUSANOV, LIM_MC, true>},
         {adj_stage_kernel<HLLE, LIM_VAN_LEER, true>, adj_stage_kernel<HLLE, LIM_MC, true>},
         {adj_stage_kernel<HLLD, LIM_VAN_LEER, true>, adj_stage_kernel<HLLD, LIM_MC, true>}}};
    return t[store ? 1 : 0][solver - 1][limiter - 1];
}

GridView grid_view(const chimera_adj *a, const Ptr6 &in, const Ptr6 &u0, const Ptr6 &out,
                   const double *dt_dev) {
    GridView g{};
    for (int q = 0; q < 6; ++q) {
        g.uin[q] = in.p[q];
        g.u0[q] = u0.p[q];
        g.uout[q] = out.p[q];
    }
    g.dt = dt_dev;
    g.mask = nullptr;
    g.nx = a->G.nx;
    g.ny = a->G.ny;
    g.px = a->G.px;
    g.py = a->G.py;
    g.dx = a->G.dx;
    g.vol = a->G.vol;
    g.gamma = a->G.gamma;
    return g;
}

// One stage through the fused kernel; with store, keeping the intermediates in
// S[stage - 1].
int run_forward_stage(chimera_adj *a, int stage, const Ptr6 &in, const Ptr6 &u0, const Ptr6 &out,
                      const double *dt_dev, bool store = true) {
    GridView g = grid_view(a, in, u0, out, dt_dev);
    hipLaunchKernelGGL(adj_stage_fn(a->sch.riemann_solver, a->sch.slope_limiter, store),
                       dim3(a->grid), dim3(STAGE_THREADS), 0, a->stream, g, a->S[stage - 1], stage,
                       a->G.use_lim, a->G.mood, a->tiles_x, a->ntiles);
    return launch_check("adjoint forward stage");
}

using TileAFn = void (*)(GridView, AdjGrid, StageBufs, RevBufs, const double *const *,
                         double *const *, double *const *, int, double, int, int);
using TileBFn = void (*)(AdjGrid, const double *const *, StageStore, RevBufs, double *const *,
                         const double *const *, int, int);

template <int TX, int TY>
constexpr TileAFn tile_a_of(int solver, int limiter) {
    return solver == RUSANOV ? (limiter == LIM_MC ? ar_tile_a<RUSANOV, LIM_MC, TX, TY>
                                                  : ar_tile_a<RUSANOV, LIM_VAN_LEER, TX, TY>)
         : solver == HLLE    ? (limiter == LIM_MC ? ar_tile_a<HLLE, LIM_MC, TX, TY>
                                                  : ar_tile_a<HLLE, LIM_VAN_LEER, TX, TY>)
                             : (limiter == LIM_MC ? ar_tile_a<HLLD, LIM_MC, TX, TY>
                                                  : ar_tile_a<HLLD, LIM_VAN_LEER, TX, TY>);
}

TileAFn tile_a_fn(int solver, int limiter, int shape) {
    switch (shape) {
    case 0: return tile_a_of<16, 16>(solver, limiter);
    case 1: return tile_a_of<16, 8>(solver, limiter);
    case 2: return tile_a_of<8, 8>(solver, limiter);
    default: return tile_a_of<32, 8>(solver, limiter);
    }
}

TileBFn tile_b_fn(int limiter, int shape) {
    const bool mc = limiter == LIM_MC;
    switch (shape) {
    case 0: return mc ? ar_tile_b<LIM_MC, 16, 16> : ar_tile_b<LIM_VAN_LEER, 16, 16>;
    case 1: return mc ? ar_tile_b<LIM_MC, 16, 8> : ar_tile_b<LIM_VAN_LEER, 16, 8>;
    case 2: return mc ? ar_tile_b<LIM_MC, 8, 8> : ar_tile_b<LIM_VAN_LEER, 8, 8>;
    default: return mc ? ar_tile_b<LIM_MC, 32, 8> : ar_tile_b<LIM_VAN_LEER, 32, 8>;
    }
}

// tiles_x, ntiles and the (multiple of 8) grid of one reverse tile shape
void rev_launch_shape(const chimera_adj *a, int shape, int &tiles_x, int &ntiles, int &grid) {
    const RevTile t = REV_SHAPES[shape];
    tiles_x = (a->G.nx + t.tx - 1) / t.tx;
    ntiles = tiles_x * ((a->G.ny + t.ty - 1) / t.ty);
    grid = (ntiles + 7) / 8 * 8;
}

// The reverse of one stage through the fused kernels: in is the stage input (host
// pointer set and device table tin), u0 the step input.  Stage 3 starts lam_u0,
// stage 2 adds to it and stage 1 adds it to lam_in.
int run_reverse_stage_fused(chimera_adj *a, int stage, const Ptr6 &in, Tab tin, const Ptr6 &u0,
                            Tab tlout, Tab tlin, double dt) {
    GridView g = grid_view(a, in, u0, in, nullptr);
    int tx, nt, grid;
    rev_launch_shape(a, a->rshape_a, tx, nt, grid);
    const RevTile ta = REV_SHAPES[a->rshape_a];
    hipLaunchKernelGGL(tile_a_fn(a->sch.riemann_solver, a->sch.slope_limiter, a->rshape_a),
                       dim3(grid), dim3(ta.tx * ta.ty), 0, a->stream, g, a->G, a->S[stage - 1],
                       a->R, tlout, tlin, table(a, TB_LU0), stage, dt, tx, nt);
    rev_launch_shape(a, a->rshape_b, tx, nt, grid);
    const RevTile tb = REV_SHAPES[a->rshape_b];
    hipLaunchKernelGGL(tile_b_fn(a->sch.slope_limiter, a->rshape_b), dim3(grid),
                       dim3(tb.tx * tb.ty), 0, a->stream, a->G, tin, a->S[stage - 1], a->R, tlin,
                       stage == 1 ? (const double *const *)table(a, TB_LU0) : nullptr, tx, nt);
    return launch_check("adjoint fused reverse stage");
}

// The reverse of one stage through the stored-intermediate kernels.
int run_reverse_stage(chimera_adj *a, int stage, Tab tin, Tab tu0, Tab tlout, Tab tlin, double dt) {
    StageBufs &S = a->S[stage - 1];
    RevBufs &R = a->R;
    const int nb = nblocks(a->n);
    hipLaunchKernelGGL(ar_cell_tail, dim3(nb), dim3(BS), 0, a->stream, a->G, tu0, S, R, tlout, tlin,
                       table(a, TB_LU0), stage, dt);
    hipLaunchKernelGGL(ar_corner, dim3(nb), dim3(BS), 0, a->stream, a->G, R, dt);
    switch (a->sch.riemann_solver) {
    case RUSANOV: hipLaunchKernelGGL(ar_faces<RUSANOV>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R, dt); break;
    case HLLE: hipLaunchKernelGGL(ar_faces<HLLE>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R, dt); break;
    default: hipLaunchKernelGGL(ar_faces<HLLD>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R, dt); break;
    }
    if (a->sch.slope_limiter == LIM_MC)
        hipLaunchKernelGGL(ar_slopes<LIM_MC>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R);
    else
        hipLaunchKernelGGL(ar_slopes<LIM_VAN_LEER>, dim3(nb), dim3(BS), 0, a->stream, a->G, S, R);
    hipLaunchKernelGGL(ar_prims, dim3(nb), dim3(BS), 0, a->stream, a->G, tin, S, R, tlin);
    hipLaunchKernelGGL(ar_bavg, dim3(nb), dim3(BS), 0, a->stream, a->G, R, tlin);
    return launch_check("adjoint reverse stage");
}

// lam_in = (dU^{n+1}/dU^n)^T lam_out for one step at U^n (host pointers hs, device
// table ts), dt = *dt_dev; asynchronous.  tlout and tlin must differ; neither may be
// LA, LB or LU0.
// The stage inputs U^(1), U^(2) of the step may be supplied (host pointer sets s1, s2
// and device tables t1, t2, e.g. recorded by the segment recompute); otherwise the
// fused path computes them with two forward stages.
int vjp_core(chimera_adj *a, const Ptr6 &hs, Tab ts, const double *dt_dev, double dt, Tab tlout,
             Tab tlin, const Ptr6 *s1 = nullptr, Tab t1 = nullptr, const Ptr6 *s2 = nullptr,
             Tab t2 = nullptr) {
    const int nb = nblocks(a->n);
    if (a->fused) {
        // the reverse kernels recompute everything else from the stage inputs
        if (!s1 || !s2) {
            CH_TRY(run_forward_stage(a, 1, hs, hs, a->stage_state[1], dt_dev, false));
            CH_TRY(run_forward_stage(a, 2, a->stage_state[1], hs, a->stage_state[2], dt_dev, false));
            s1 = &a->stage_state[1];
            s2 = &a->stage_state[2];
            t1 = table(a, TB_S1);
            t2 = table(a, TB_S2);
        }
        CH_TRY(run_reverse_stage_fused(a, 3, *s2, t2, hs, tlout, table(a, TB_LB), dt));
        CH_TRY(run_reverse_stage_fused(a, 2, *s1, t1, hs, table(a, TB_LB), table(a, TB_LA), dt));
        return run_reverse_stage_fused(a, 1, hs, ts, hs, table(a, TB_LA), tlin, dt);
    }
    // forward: keep every stage's intermediates
    CH_TRY(run_forward_stage(a, 1, hs, hs, a->stage_state[1], dt_dev));
    CH_TRY(run_forward_stage(a, 2, a->stage_state[1], hs, a->stage_state[2], dt_dev));
    CH_TRY(run_forward_stage(a, 3, a->stage_state[2], hs, a->scratch_out, dt_dev));
    hipLaunchKernelGGL(zero6_kernel, dim3(nb), dim3(BS), 0, a->stream, table(a, TB_LU0), a->n);
    // reverse: stage 3 (lam_out -> LB), stage 2 (LB -> LA), stage 1 (LA -> lam_in)
    CH_TRY(run_reverse_stage(a, 3, table(a, TB_S2), ts, tlout, table(a, TB_LB), dt));
    CH_TRY(run_reverse_stage(a, 2, table(a, TB_S1), ts, table(a, TB_LB), table(a, TB_LA), dt));
    CH_TRY(run_reverse_stage(a, 1, ts, ts, table(a, TB_LA), tlin, dt));
    hipLaunchKernelGGL(add_kernel, dim3(nb), dim3(BS), 0, a->stream, tlin,
                       (const double *const *)table(a, TB_LU0), a->n);
    return launch_check("adjoint accumulate");
}

int ensure_traj(chimera_adj *a, int slots) {
    if (slots <= a->traj_cap) return CHIMERA_OK;
    CH_HIP(hipStreamSynchronize(a->stream));
    (void)hipFree(a->traj);
    (void)hipFree(a->traj_tab);
    a->traj = nullptr;
    a->traj_tab = nullptr;
    a->traj_cap = 0;
    if (hipMalloc(&a->traj, (size_t)slots * 6 * a->n * sizeof(double)) != hipSuccess ||
        hipMalloc(&a->traj_tab, (size_t)slots * 6 * sizeof(double *)) != hipSuccess) {
        set_error("adjoint trajectory storage (%d states of %zu cells) failed", slots, a->n);
        return CHIMERA_ENOMEM;
    }
    std::vector<double *> h((size_t)slots * 6);
    for (size_t k = 0; k < h.size(); ++k) h[k] = a->traj + k * a->n;
    CH_HIP(hipMemcpy(a->traj_tab, h.data(), h.size() * sizeof(double *), hipMemcpyHostToDevice));
    a->traj_cap = slots;
    return CHIMERA_OK;
}

Ptr6 traj_slot(chimera_adj *a, int s) {
    Ptr6 r;
    for (int q = 0; q < 6; ++q) r.p[q] = a->traj + (6 * (size_t)s + q) * a->n;
    return r;
}

}  // namespace

extern "C" int chimera_adj_destroy(chimera_adj_t *a) {
    if (!a) return CHIMERA_OK;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    (void)hipFree(a->arena);
    (void)hipFree(a->dptr);
    (void)hipFree(a->dt_dev);
    (void)hipFree(a->traj);
    (void)hipFree(a->traj_tab);
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
    a->tiles_x = (nx + ATX - 1) / ATX;
    a->ntiles = a->tiles_x * ((ny + ATY - 1) / ATY);
    a->grid = (a->ntiles + 7) / 8 * 8;
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
    // per stage: 6 P + 24 faces + 2 bits + 4 Up + 4 U3 = 40
    // reverse: 4 + 2 + 2 + 6 + 1 + 24 + 6 + 24 + 6 + 2 = 77; states 3 x 6; lambdas 5 x 6
    const size_t per_stage = 40, rev = 77, extra = 18 + 30;
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
        S.bits_x = reinterpret_cast<unsigned long long *>(take());
        S.bits_y = reinterpret_cast<unsigned long long *>(take());
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
    for (auto &x : R.gmx) x = take();
    for (auto &x : R.gpx) x = take();
    for (auto &x : R.gmy) x = take();
    for (auto &x : R.gpy) x = take();
    for (auto &x : R.lP) x = take();
    for (auto &x : R.lBcT) x = take();
    for (auto &st : a->stage_state)
        for (auto &x : st.p) x = take();
    for (auto &x : a->lam_a.p) x = take();
    for (auto &x : a->lam_b.p) x = take();
    for (auto &x : a->lam_c.p) x = take();
    for (auto &x : a->lam_u0.p) x = take();
    for (auto &x : a->scratch_out.p) x = take();
    int rc = set_table(a, TB_S0, a->stage_state[0].p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_S1, a->stage_state[1].p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_S2, a->stage_state[2].p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_LA, a->lam_a.p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_LB, a->lam_b.p);
    if (rc == CHIMERA_OK) rc = set_table(a, TB_LC, a->lam_c.p);
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

// One SSP-RK3 step with a given dt through the adjoint's forward (the STORE
// instantiation of the fused stage kernel); device pointers.
extern "C" int chimera_adj_forward_step(chimera_adj_t *a, const double *const *state_in,
                                        double dt, double *const *state_out) {
    CH_REQUIRE(a && state_in && state_out, "NULL argument");
    CH_REQUIRE(dt > 0.0, "dt must be > 0");
    CH_TRY(set_dts(a, &dt, 1));
    size_t bytes = a->n * sizeof(double);
    for (int q = 0; q < 6; ++q)
        CH_HIP(hipMemcpyAsync(a->stage_state[0].p[q], state_in[q], bytes, hipMemcpyDefault, a->stream));
    Ptr6 out = ptr6((const double *const *)state_out);
    const Ptr6 &s0 = a->stage_state[0];
    CH_TRY(run_forward_stage(a, 1, s0, s0, a->stage_state[1], a->dt_dev));
    CH_TRY(run_forward_stage(a, 2, a->stage_state[1], s0, a->stage_state[2], a->dt_dev));
    CH_TRY(run_forward_stage(a, 3, a->stage_state[2], s0, out, a->dt_dev));
    CH_HIP(hipStreamSynchronize(a->stream));
    return CHIMERA_OK;
}

// lam_in = (d U^{n+1} / d U^n)^T lam_out for one step from U^n = state_n with dt frozen.
extern "C" int chimera_adj_step_vjp(chimera_adj_t *a, const double *const *state_n, double dt,
                                    const double *const *lam_out, double *const *lam_in) {
    CH_REQUIRE(a && state_n && lam_out && lam_in, "NULL argument");
    CH_REQUIRE(dt > 0.0, "dt must be > 0");
    CH_TRY(set_dts(a, &dt, 1));
    size_t bytes = a->n * sizeof(double);
    for (int q = 0; q < 6; ++q) {
        CH_HIP(hipMemcpyAsync(a->stage_state[0].p[q], state_n[q], bytes, hipMemcpyDefault, a->stream));
        CH_HIP(hipMemcpyAsync(a->lam_c.p[q], lam_out[q], bytes, hipMemcpyDefault, a->stream));
    }
    CH_TRY(set_table(a, TB_LIN, (const double *const *)lam_in));
    CH_TRY(vjp_core(a, a->stage_state[0], table(a, TB_S0), a->dt_dev, dt, table(a, TB_LC),
                    table(a, TB_LIN)));
    CH_HIP(hipStreamSynchronize(a->stream));
    return CHIMERA_OK;
}

extern "C" int chimera_adj_segment_vjp(chimera_adj_t *a, const double *const *state0,
                                       const double *dt_in, int nsteps, double *const *lam) {
    CH_REQUIRE(a && state0 && dt_in && lam, "NULL argument");
    CH_REQUIRE(nsteps >= 1, "nsteps must be >= 1");
    for (int s = 0; s < nsteps; ++s)
        CH_REQUIRE(dt_in[s] > 0.0, "dt_in[%d] = %g is not positive", s, dt_in[s]);
    CH_TRY(ensure_traj(a, 3 * nsteps));
    CH_TRY(set_dts(a, dt_in, nsteps));
    CH_TRY(set_table(a, TB_IN, state0));
    CH_TRY(set_table(a, TB_LIN, (const double *const *)lam));
    // recompute the segment, keeping every stage input: slot 3s holds U^s (s >= 1;
    // U^0 is the caller's state0), slots 3s+1 and 3s+2 the stage outputs U^(1), U^(2)
    const Ptr6 h0 = ptr6(state0);
    for (int s = 0; s < nsteps; ++s) {
        const Ptr6 in = s == 0 ? h0 : traj_slot(a, 3 * s);
        const double *dts = a->dt_dev + s;
        CH_TRY(run_forward_stage(a, 1, in, in, traj_slot(a, 3 * s + 1), dts, false));
        CH_TRY(run_forward_stage(a, 2, traj_slot(a, 3 * s + 1), in, traj_slot(a, 3 * s + 2), dts,
                                 false));
        if (s + 1 < nsteps)
            CH_TRY(run_forward_stage(a, 3, traj_slot(a, 3 * s + 2), in, traj_slot(a, 3 * s + 3), dts,
                                     false));
    }
    auto slot_tab = [&](int k) { return (Tab)(a->traj_tab + 6 * (size_t)k); };
    // lam ping-pongs between the caller's arrays and lam_c
    Tab cur = table(a, TB_LIN), other = table(a, TB_LC);
    for (int s = nsteps - 1; s >= 0; --s) {
        const Ptr6 hs = s == 0 ? h0 : traj_slot(a, 3 * s);
        Tab ts = s == 0 ? table(a, TB_IN) : slot_tab(3 * s);
        const Ptr6 s1 = traj_slot(a, 3 * s + 1), s2 = traj_slot(a, 3 * s + 2);
        CH_TRY(vjp_core(a, hs, ts, a->dt_dev + s, dt_in[s], cur, other, &s1, slot_tab(3 * s + 1),
                        &s2, slot_tab(3 * s + 2)));
        Tab t = cur;
        cur = other;
        other = t;
    }
    if (cur != table(a, TB_LIN)) {
        size_t bytes = a->n * sizeof(double);
        for (int q = 0; q < 6; ++q)
            CH_HIP(hipMemcpyAsync(lam[q], a->lam_c.p[q], bytes, hipMemcpyDeviceToDevice, a->stream));
    }
    CH_HIP(hipStreamSynchronize(a->stream));
    return CHIMERA_OK;
}

extern "C" void *chimera_adj_stream(chimera_adj_t *a) { return a ? (void *)a->stream : nullptr; }

// Tile shapes of the two fused reverse kernels (indices into REV_SHAPES; -1 keeps).
extern "C" int chimera_adj_set_tiles(chimera_adj_t *a, int shape_a, int shape_b) {
    CH_REQUIRE(a, "NULL adjoint");
    CH_REQUIRE(shape_a >= -1 && shape_a < REV_NUM_SHAPES && shape_b >= -1 &&
                   shape_b < REV_NUM_SHAPES,
               "tile shapes must be -1 or 0..%d, got %d, %d", REV_NUM_SHAPES - 1, shape_a, shape_b);
    CH_HIP(hipStreamSynchronize(a->stream));
    if (shape_a >= 0) a->rshape_a = shape_a;
    if (shape_b >= 0) a->rshape_b = shape_b;
    return CHIMERA_OK;
}

// 1 (default): the fused reverse kernels; 0: the stored-intermediate kernels.  Both
// give the same bits; the second is kept as the cross-check of the first.
extern "C" int chimera_adj_set_fused(chimera_adj_t *a, int fused) {
    CH_REQUIRE(a, "NULL adjoint");
    CH_REQUIRE(fused == 0 || fused == 1, "fused must be 0 or 1, got %d", fused);
    CH_HIP(hipStreamSynchronize(a->stream));
    a->fused = fused != 0;
    return CHIMERA_OK;
}

extern "C" int chimera_face_vjp(int solver, int method, int n, double gamma, const double *in,
                                const double *lam_flux, double *out) {
    CH_REQUIRE(in && lam_flux && out, "NULL argument");
    CH_REQUIRE(solver >= 1 && solver <= 3, "solver must be 1-3, got %d", solver);
    CH_REQUIRE(method == 0 || method == 1, "method must be 0 or 1, got %d", method);
    CH_REQUIRE(n >= 1 && gamma > 1.0, "need n >= 1 and gamma > 1");
    const size_t nn = (size_t)n;
    double *d = nullptr;
    CH_HIP(hipMalloc(&d, 29 * nn * sizeof(double)));
    double *din = d, *dlam = d + 12 * nn, *dout = d + 17 * nn;
    int rc = CHIMERA_OK;
    if (hipMemcpy(din, in, 12 * nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dlam, lam_flux, 5 * nn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("face_vjp: upload failed");
        rc = CHIMERA_EHIP;
    }
    if (rc == CHIMERA_OK) {
        const int nb = nblocks(nn);
        switch (solver) {
        case RUSANOV: hipLaunchKernelGGL(face_vjp_kernel<RUSANOV>, dim3(nb), dim3(BS), 0, 0, method, n, gamma, din, dlam, dout); break;
        case HLLE: hipLaunchKernelGGL(face_vjp_kernel<HLLE>, dim3(nb), dim3(BS), 0, 0, method, n, gamma, din, dlam, dout); break;
        default: hipLaunchKernelGGL(face_vjp_kernel<HLLD>, dim3(nb), dim3(BS), 0, 0, method, n, gamma, din, dlam, dout); break;
        }
        rc = launch_check("face_vjp_kernel");
    }
    if (rc == CHIMERA_OK &&
        hipMemcpy(out, dout, 12 * nn * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("face_vjp: download failed");
        rc = CHIMERA_EHIP;
    }
    (void)hipFree(d);
    return rc;
}
