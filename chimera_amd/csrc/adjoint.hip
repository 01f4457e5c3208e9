// Discrete adjoint of the SSP-RK3 step (frozen dt): hand-written reverse kernels for
// the reconstruction (limiter, MOOD, face floors, thermal check), the Riemann fluxes,
// the conservative update, constrained transport, the outflow correction, the conserved
// floors and the RK blends.
//
// Two reverse paths give the same bits.  The default, fused one: the segment recompute
// (stage_tile, STORE_UP) keeps every stage input and each stage's Up, and per stage two
// LDS-tiled kernels (ar_tile_a / ar_tile_b over adjoint_tile.hpp) recompute the rest of
// the stage tile by tile and reverse it.  The cross-check: the STORE_ALL instantiation
// keeps every intermediate the reverse needs -- primitives, face states after MOOD and
// the thermal check, branch bits, the pre-floor and pre-blend states -- and the
// kernels ar_* read them back.  Both are in gather form (each thread owns one output
// location and collects the contributions of every forward operation that read it), so
// there are no atomics and the gradient is deterministic.  Branches (limiter cases,
// MOOD resets, face and conserved floors, thermal resets, Riemann-solver regions and
// catches) are the ones the forward pass took.  The three Riemann solvers are reversed
// by hand (riemann_adj.hpp); their forward-mode dual-number VJPs (riemann_dual.hpp) are
// the cross-check.
//
// A whole reverse sweep (chimera_adj_sweep) overlaps the recompute of checkpoint
// segment k-1, on a second stream, with the reverse of segment k: the reverse kernels
// leave SIMD slots idle at sizes where one thread per cell gives a few waves per SIMD.
#include <algorithm>
#include <vector>

#include "adjoint_phys.hpp"
#include "adjoint_tile.hpp"
#include "common.hpp"
#include "resident.hpp"
#include "riemann_adj.hpp"

namespace chimera {
namespace {

constexpr int BS = 256;

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

// tile of the recompute kernel: 32 x 8 (a 32 x 16 tile, faster for the plain stage
// kernel, made the STORE_UP recompute slower on the 512 x 256 field loop: 174 vs 168 ms
// per 1000-step backward)
constexpr int ATX = 32, ATY = 8;
// tile shapes of the fused reverse kernels (chimera_adj_set_tiles); smaller tiles give
// more workgroups on small grids at the price of more halo work
struct RevTile {
    int tx, ty;
};
constexpr int REV_NUM_SHAPES = 4;
constexpr RevTile REV_SHAPES[REV_NUM_SHAPES] = {{16, 16}, {16, 8}, {8, 8}, {32, 8}};

// HLLD at three waves per SIMD (at most 168 registers, as the tile's LDS allows) but for
// the stored-intermediate cross-check (STORE_ALL, 193 registers)
template <int SOLVER, int LIMITER, int STORE>
__global__ __launch_bounds__(ATX * ATY)
__attribute__((amdgpu_waves_per_eu((SOLVER == HLLD && STORE != STORE_ALL) ? 3 : 1)))
void adj_stage_kernel(GridView g, StageStore st, int stage, int use_lim, int mood, int tiles_x,
                      int ntiles) {
    __shared__ double lds[TileShape<ATX, ATY>::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;  // nb is a multiple of 8
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    stage_tile<SOLVER, LIMITER, ATX, ATY, STORE>(g, stage, use_lim != 0, mood != 0, tile % tiles_x,
                                                 tile / tiles_x, lds, STORE != STORE_NONE ? &st : nullptr);
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
    const size_t n = (size_t)nx * ny;   // the reverse arrays are contiguous, n apart
    face_state_reverse(lL, sL, false, bits, lreset, candL[0], n, k);
    face_state_reverse(lR, sR, lastf, bits, lreset, candR[0], n, k);
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

// ---- fused reverse kernels (one tile per workgroup, XCD-aware tile map) --------------

// HLLD at two waves per SIMD: its face reverse needs ~330 registers a lane and held one
// wave; the excess spills to scratch (OT 512^2 segment VJP 3% faster, same bits).
template <int SOLVER, int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX *TY)
__attribute__((amdgpu_waves_per_eu(SOLVER == HLLD ? 2 : 1)))
void ar_tile_a(GridView g, AdjGrid G, FusedBufs F, const double *const *Up,
               const double *const *lout, double *const *lin, double *const *lU0, int stage,
               double dt, int tiles_x, int ntiles) {
    __shared__ double lds[RevShapeA<TX, TY>::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;   // nb is a multiple of 8
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    rev_tile_a<SOLVER, LIMITER, TX, TY>(g, G, F, Up, lout, lin, lU0, stage, dt, tile % tiles_x,
                                        tile / tiles_x, lds);
}

template <int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void ar_tile_b(AdjGrid G, const double *const *U, FusedBufs F,
                                                   double *const *lin,
                                                   const double *const *lU0_add, int tiles_x,
                                                   int ntiles) {
    __shared__ double lds[RevShapeB<TX, TY>::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    rev_tile_b<LIMITER, TX, TY>(G, U, F, lin, lU0_add, tile % tiles_x, tile / tiles_x, lds);
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
    FusedBufs F{nullptr, 0};                  // their exchange arrays
    int rshape_a = 3, rshape_b = 3;           // their tile shapes (REV_SHAPES)
    chimera_scheme_t sch{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t n = 0;
    int tiles_x = 0, ntiles = 0, grid = 0;   // adjoint stage kernel launch shape
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
    // checkpoint recomputation: the recomputed stage inputs and each stage's Up
    double *traj = nullptr;       // traj_cap x 6 fields
    double **traj_tab = nullptr;  // device tables, 6 pointers per slot
    int traj_cap = 0;
    double *uptraj = nullptr;     // up_cap x 4 fields
    double **up_tab = nullptr;    // device tables, 4 pointers per slot
    int up_cap = 0;
    // the whole-run sweep: segment k-1 is recomputed on rstream while segment k is
    // reversed on stream (two trajectory halves, alternating by segment parity)
    hipStream_t rstream = nullptr;
    hipEvent_t ev_start = nullptr, ev_traj[2] = {nullptr, nullptr}, ev_rev[2] = {nullptr, nullptr};
    double **ck_tab = nullptr;    // device tables of the checkpoints, 6 pointers each
    int ck_cap = 0;
    // the last segment of a run recorded by chimera_adj_record for the next sweep
    struct Recorded {
        bool valid = false;
        const double *state0[6] = {};
        int nsteps = 0, K = 0, nseg = 0;
        std::vector<double> dts;
    } rec;
};

namespace {

// device pointer tables: [0..2] stage_state, [3] lam_a, [4] lam_b, [5] lam_u0,
// [6] scratch_out, [7] caller input, [8] caller output, [9] caller lam in,
// [10] caller lam out, [11] lam_c, [12..14] Up of stages 1..3 (S[].Up, 4 pointers)
enum { TB_S0 = 0, TB_S1 = 1, TB_S2 = 2, TB_LA = 3, TB_LB = 4, TB_LU0 = 5, TB_SCR = 6, TB_IN = 7,
       TB_OUT = 8, TB_LIN = 9, TB_LOUT = 10, TB_LC = 11, TB_UP1 = 12, TB_UP2 = 13, TB_UP3 = 14,
       TB_COUNT = 15 };

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

template <int MODE>
constexpr AdjStageFn adj_stage_of(int solver, int limiter) {
    return solver == RUSANOV ? (limiter == LIM_MC ? adj_stage_kernel<RUSANOV, LIM_MC, MODE>
                                                  : adj_stage_kernel<RUSANOV, LIM_VAN_LEER, MODE>)
         : solver == HLLE    ? (limiter == LIM_MC ? adj_stage_kernel<HLLE, LIM_MC, MODE>
                                                  : adj_stage_kernel<HLLE, LIM_VAN_LEER, MODE>)
                             : (limiter == LIM_MC ? adj_stage_kernel<HLLD, LIM_MC, MODE>
                                                  : adj_stage_kernel<HLLD, LIM_VAN_LEER, MODE>);
}

AdjStageFn adj_stage_fn(int solver, int limiter, int mode) {
    return mode == STORE_ALL  ? adj_stage_of<STORE_ALL>(solver, limiter)
         : mode == STORE_UP   ? adj_stage_of<STORE_UP>(solver, limiter)
                              : adj_stage_of<STORE_NONE>(solver, limiter);
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

// One stage through the fused kernel on stream st (default: the adjoint's stream).
// STORE_ALL keeps the intermediates in S[stage - 1]; STORE_UP keeps only Up, in the
// four arrays up (host pointer set).
int run_forward_stage(chimera_adj *a, int stage, const Ptr6 &in, const Ptr6 &u0, const Ptr6 &out,
                      const double *dt_dev, int mode = STORE_ALL, hipStream_t st = nullptr,
                      double *const *up = nullptr) {
    GridView g = grid_view(a, in, u0, out, dt_dev);
    StageStore store = a->S[stage - 1];
    if (mode == STORE_UP)
        for (int q = 0; q < 4; ++q) store.Up[q] = up[q];
    hipLaunchKernelGGL(adj_stage_fn(a->sch.riemann_solver, a->sch.slope_limiter, mode),
                       dim3(a->grid), dim3(ATX * ATY), 0, st ? st : a->stream, g, store, stage,
                       a->G.use_lim, a->G.mood, a->tiles_x, a->ntiles);
    return launch_check("adjoint forward stage");
}

using TileAFn = void (*)(GridView, AdjGrid, FusedBufs, const double *const *,
                         const double *const *, double *const *, double *const *, int, double, int,
                         int);
using TileBFn = void (*)(AdjGrid, const double *const *, FusedBufs, double *const *,
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
// pointer set and device table tin), u0 the step input, tup the stage's Up (device
// table of 4).  Stage 3 starts lam_u0, stage 2 adds to it and stage 1 adds it to lam_in.
int run_reverse_stage_fused(chimera_adj *a, int stage, const Ptr6 &in, Tab tin, const Ptr6 &u0,
                            Tab tup, Tab tlout, Tab tlin, double dt) {
    GridView g = grid_view(a, in, u0, in, nullptr);
    int tx, nt, grid;
    rev_launch_shape(a, a->rshape_a, tx, nt, grid);
    const RevTile ta = REV_SHAPES[a->rshape_a];
    hipLaunchKernelGGL(tile_a_fn(a->sch.riemann_solver, a->sch.slope_limiter, a->rshape_a),
                       dim3(grid), dim3(ta.tx * ta.ty), 0, a->stream, g, a->G, a->F, tup, tlout,
                       tlin, table(a, TB_LU0), stage, dt, tx, nt);
    rev_launch_shape(a, a->rshape_b, tx, nt, grid);
    const RevTile tb = REV_SHAPES[a->rshape_b];
    hipLaunchKernelGGL(tile_b_fn(a->sch.slope_limiter, a->rshape_b), dim3(grid),
                       dim3(tb.tx * tb.ty), 0, a->stream, a->G, tin, a->F, tlin,
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
// The stage inputs U^(1), U^(2) of the step and the Up of its three stages may be
// supplied (host pointer sets s1, s2, device tables t1, t2 and up[3], as the segment
// recompute records them); otherwise the fused path computes them with three forward
// stages.
int vjp_core(chimera_adj *a, const Ptr6 &hs, Tab ts, const double *dt_dev, double dt, Tab tlout,
             Tab tlin, const Ptr6 *s1 = nullptr, Tab t1 = nullptr, const Ptr6 *s2 = nullptr,
             Tab t2 = nullptr, const Tab *up = nullptr) {
    const int nb = nblocks(a->n);
    if (a->fused) {
        // the reverse kernels recompute everything else from the stage inputs and Up
        Tab ups[3];
        if (!s1 || !s2 || !up) {
            CH_TRY(run_forward_stage(a, 1, hs, hs, a->stage_state[1], dt_dev, STORE_UP, nullptr,
                                     a->S[0].Up));
            CH_TRY(run_forward_stage(a, 2, a->stage_state[1], hs, a->stage_state[2], dt_dev,
                                     STORE_UP, nullptr, a->S[1].Up));
            CH_TRY(run_forward_stage(a, 3, a->stage_state[2], hs, a->scratch_out, dt_dev, STORE_UP,
                                     nullptr, a->S[2].Up));
            s1 = &a->stage_state[1];
            s2 = &a->stage_state[2];
            t1 = table(a, TB_S1);
            t2 = table(a, TB_S2);
            for (int k = 0; k < 3; ++k) ups[k] = table(a, TB_UP1 + k);
        } else {
            for (int k = 0; k < 3; ++k) ups[k] = up[k];
        }
        CH_TRY(run_reverse_stage_fused(a, 3, *s2, t2, hs, ups[2], tlout, table(a, TB_LB), dt));
        CH_TRY(run_reverse_stage_fused(a, 2, *s1, t1, hs, ups[1], table(a, TB_LB), table(a, TB_LA),
                                       dt));
        return run_reverse_stage_fused(a, 1, hs, ts, hs, ups[0], table(a, TB_LA), tlin, dt);
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

// slots x fields arrays of n cells in one allocation, with a device table of fields
// pointers per slot (grown on demand, kept)
int ensure_slots(chimera_adj *a, double *&buf, double **&tab, int &cap, int slots, int fields) {
    if (slots <= cap) return CHIMERA_OK;
    a->rec.valid = false;   // a recorded segment does not survive a reallocation
    CH_HIP(hipStreamSynchronize(a->stream));
    if (a->rstream) CH_HIP(hipStreamSynchronize(a->rstream));
    (void)hipFree(buf);
    (void)hipFree(tab);
    buf = nullptr;
    tab = nullptr;
    cap = 0;
    if (hipMalloc(&buf, (size_t)slots * fields * a->n * sizeof(double)) != hipSuccess ||
        hipMalloc(&tab, (size_t)slots * fields * sizeof(double *)) != hipSuccess) {
        set_error("adjoint trajectory storage (%d x %d arrays of %zu cells) failed", slots, fields,
                  a->n);
        return CHIMERA_ENOMEM;
    }
    std::vector<double *> h((size_t)slots * fields);
    for (size_t k = 0; k < h.size(); ++k) h[k] = buf + k * a->n;
    CH_HIP(hipMemcpy(tab, h.data(), h.size() * sizeof(double *), hipMemcpyHostToDevice));
    cap = slots;
    return CHIMERA_OK;
}

// trajectories of `halves` segments of K steps: per half 3K + 1 states (the stage
// inputs and the last stage's output) and 3K Up sets
int ensure_traj(chimera_adj *a, int K, int halves) {
    CH_TRY(ensure_slots(a, a->traj, a->traj_tab, a->traj_cap, halves * (3 * K + 1), 6));
    return ensure_slots(a, a->uptraj, a->up_tab, a->up_cap, halves * 3 * K, 4);
}

Ptr6 traj_slot(chimera_adj *a, int s) {
    Ptr6 r;
    for (int q = 0; q < 6; ++q) r.p[q] = a->traj + (6 * (size_t)s + q) * a->n;
    return r;
}

Tab slot_tab(chimera_adj *a, int s) { return (Tab)(a->traj_tab + 6 * (size_t)s); }

// Up slot s: four host pointers / its device table
struct Ptr4 {
    double *p[4];
};
Ptr4 up_slot(chimera_adj *a, int s) {
    Ptr4 r;
    for (int q = 0; q < 4; ++q) r.p[q] = a->uptraj + (4 * (size_t)s + q) * a->n;
    return r;
}
Tab up_tab(chimera_adj *a, int s) { return (Tab)(a->up_tab + 4 * (size_t)s); }

// Recompute steps [s_begin, s_end) of a checkpoint segment that starts at h0, on
// stream st, keeping every stage input in the trajectory slots from base on -- slot
// base+3s holds U^s (s >= 1; U^0 is h0), slots base+3s+1 and base+3s+2 the stage
// outputs U^(1), U^(2), slot base+3K the last step's output -- and the Up of step s,
// stage st in Up slot upbase+3s+st-1.  dt_dev: the segment's dts.  Step s reads slot
// base+3s, so the steps must be enqueued in order.
int recompute_steps(chimera_adj *a, hipStream_t st, const Ptr6 &h0, const double *dt_dev,
                    int s_begin, int s_end, int base, int upbase) {
    for (int s = s_begin; s < s_end; ++s) {
        const Ptr6 in = s == 0 ? h0 : traj_slot(a, base + 3 * s);
        const double *dts = dt_dev + s;
        for (int stage = 1; stage <= 3; ++stage) {
            const Ptr6 sin = stage == 1 ? in : traj_slot(a, base + 3 * s + stage - 1);
            const Ptr4 up = up_slot(a, upbase + 3 * s + stage - 1);
            CH_TRY(run_forward_stage(a, stage, sin, in, traj_slot(a, base + 3 * s + stage), dts,
                                     STORE_UP, st, up.p));
        }
    }
    return CHIMERA_OK;
}
int recompute_segment(chimera_adj *a, hipStream_t st, const Ptr6 &h0, const double *dt_dev,
                      int nsteps, int base, int upbase) {
    return recompute_steps(a, st, h0, dt_dev, 0, nsteps, base, upbase);
}

// The reverse of steps s_end-1 down to s_begin of a segment whose trajectory is in
// the slots from base / upbase; cur holds lam of step s_end on entry and lam of step
// s_begin on return (cur and other swap once per step).
int reverse_steps(chimera_adj *a, const Ptr6 &h0, Tab t0, const double *dt_dev,
                  const double *dt_host, int s_begin, int s_end, int base, int upbase, Tab &cur,
                  Tab &other) {
    for (int s = s_end - 1; s >= s_begin; --s) {
        const Ptr6 hs = s == 0 ? h0 : traj_slot(a, base + 3 * s);
        Tab ts = s == 0 ? t0 : slot_tab(a, base + 3 * s);
        const Ptr6 s1 = traj_slot(a, base + 3 * s + 1), s2 = traj_slot(a, base + 3 * s + 2);
        const Tab up[3] = {up_tab(a, upbase + 3 * s), up_tab(a, upbase + 3 * s + 1),
                           up_tab(a, upbase + 3 * s + 2)};
        CH_TRY(vjp_core(a, hs, ts, dt_dev + s, dt_host[s], cur, other, &s1,
                        slot_tab(a, base + 3 * s + 1), &s2, slot_tab(a, base + 3 * s + 2), up));
        Tab t = cur;
        cur = other;
        other = t;
    }
    return CHIMERA_OK;
}
int reverse_segment(chimera_adj *a, const Ptr6 &h0, Tab t0, const double *dt_dev,
                    const double *dt_host, int nsteps, int base, int upbase, Tab &cur,
                    Tab &other) {
    return reverse_steps(a, h0, t0, dt_dev, dt_host, 0, nsteps, base, upbase, cur, other);
}

// lam back into the caller's arrays when the ping-pong ended in lam_c
int finish_lam(chimera_adj *a, Tab cur, double *const *lam) {
    if (cur != table(a, TB_LIN)) {
        const size_t bytes = a->n * sizeof(double);
        for (int q = 0; q < 6; ++q)
            CH_HIP(hipMemcpyAsync(lam[q], a->lam_c.p[q], bytes, hipMemcpyDeviceToDevice, a->stream));
    }
    return CHIMERA_OK;
}

// what a sweep needs of a recorded last segment
struct Recorded_check_t {
    int nseg, K, nsteps;
    const double *const *state0;
    const double *dts;
};

bool recorded_matches(const chimera_adj *a, const Recorded_check_t &c) {
    const auto &r = a->rec;
    if (!r.valid || r.nseg != c.nseg || r.K != c.K || r.nsteps != c.nsteps) return false;
    for (int q = 0; q < 6; ++q)
        if (r.state0[q] != c.state0[q]) return false;
    return std::equal(r.dts.begin(), r.dts.end(), c.dts);
}

int ensure_sweep(chimera_adj *a, int nseg) {
    if (!a->rstream) {
        CH_HIP(hipStreamCreateWithFlags(&a->rstream, hipStreamNonBlocking));
        CH_HIP(hipEventCreateWithFlags(&a->ev_start, hipEventDisableTiming));
        for (int b = 0; b < 2; ++b) {
            CH_HIP(hipEventCreateWithFlags(&a->ev_traj[b], hipEventDisableTiming));
            CH_HIP(hipEventCreateWithFlags(&a->ev_rev[b], hipEventDisableTiming));
        }
    }
    if (nseg > a->ck_cap) {
        (void)hipFree(a->ck_tab);
        a->ck_tab = nullptr;
        a->ck_cap = 0;
        CH_HIP(hipMalloc(&a->ck_tab, (size_t)nseg * 6 * sizeof(double *)));
        a->ck_cap = nseg;
    }
    return CHIMERA_OK;
}

}  // namespace

extern "C" int chimera_adj_destroy(chimera_adj_t *a) {
    if (!a) return CHIMERA_OK;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    (void)hipFree(a->arena);
    (void)hipFree(a->F.base);
    (void)hipFree(a->dptr);
    (void)hipFree(a->dt_dev);
    (void)hipFree(a->traj);
    (void)hipFree(a->traj_tab);
    (void)hipFree(a->uptraj);
    (void)hipFree(a->up_tab);
    (void)hipFree(a->ck_tab);
    if (a->rstream) {
        (void)hipStreamSynchronize(a->rstream);
        (void)hipStreamDestroy(a->rstream);
    }
    for (hipEvent_t e : {a->ev_start, a->ev_traj[0], a->ev_traj[1], a->ev_rev[0], a->ev_rev[1]})
        if (e) (void)hipEventDestroy(e);
    if (a->own_stream && a->stream) (void)hipStreamDestroy(a->stream);
    delete a;
    return CHIMERA_OK;
}

extern "C" int chimera_adj_create(chimera_adj_t **out, int nx, int ny, double dx, double gamma,
                                  const chimera_bc_t *bc, const chimera_scheme_t *scheme,
                                  void *stream) {
    CH_REQUIRE(out, "out is NULL");
    CH_REQUIRE(nx >= 4 && ny >= 4, "grid must be at least 4 x 4");
    CH_REQUIRE((long long)nx * ny <= MAX_GRID_CELLS, "grid of %d x %d cells exceeds %lld cells", nx, ny,
               MAX_GRID_CELLS);
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
    a->F.n = a->n;
    if (hipMalloc(&a->arena, total * a->n * sizeof(double)) != hipSuccess ||
        hipMalloc(&a->F.base, FusedBufs::ARRAYS * a->n * sizeof(double)) != hipSuccess ||
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
    for (int st = 0; st < 3 && rc == CHIMERA_OK; ++st) {
        const double *up[6] = {a->S[st].Up[0], a->S[st].Up[1], a->S[st].Up[2], a->S[st].Up[3],
                               nullptr, nullptr};
        rc = set_table(a, TB_UP1 + st, up);
    }
    if (rc == CHIMERA_OK && hipStreamSynchronize(a->stream) != hipSuccess) rc = CHIMERA_EHIP;
    if (rc != CHIMERA_OK) {
        chimera_adj_destroy(a);
        return rc;
    }
    *out = a;
    return CHIMERA_OK;
}

// One SSP-RK3 step with a given dt through the adjoint's forward (the STORE_ALL
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
    a->rec.valid = false;   // this call overwrites the trajectory
    CH_TRY(ensure_traj(a, nsteps, 1));
    CH_TRY(set_dts(a, dt_in, nsteps));
    CH_TRY(set_table(a, TB_IN, state0));
    CH_TRY(set_table(a, TB_LIN, (const double *const *)lam));
    const Ptr6 h0 = ptr6(state0);
    CH_TRY(recompute_segment(a, a->stream, h0, a->dt_dev, nsteps, 0, 0));
    // lam ping-pongs between the caller's arrays and lam_c
    Tab cur = table(a, TB_LIN), other = table(a, TB_LC);
    CH_TRY(reverse_segment(a, h0, table(a, TB_IN), a->dt_dev, dt_in, nsteps, 0, 0, cur, other));
    CH_TRY(finish_lam(a, cur, lam));
    CH_HIP(hipStreamSynchronize(a->stream));
    return CHIMERA_OK;
}

extern "C" int chimera_adj_record(chimera_adj_t *a, const double *const *state0,
                                  const double *dt_in, int nsteps, int seg_steps, int nseg,
                                  double *const *state_out) {
    CH_REQUIRE(a && state0 && dt_in && state_out, "NULL argument");
    CH_REQUIRE(nsteps >= 1 && nseg >= 1 && seg_steps >= nsteps,
               "need 1 <= nsteps <= seg_steps and nseg >= 1 (got %d, %d, %d)", nsteps, seg_steps,
               nseg);
    for (int s = 0; s < nsteps; ++s)
        CH_REQUIRE(dt_in[s] > 0.0, "dt_in[%d] = %g is not positive", s, dt_in[s]);
    a->rec.valid = false;
    // the layout chimera_adj_sweep uses for segment nseg - 1 of this run
    const int K = nseg > 1 ? seg_steps : nsteps, h = (nseg - 1) & 1;
    CH_TRY(ensure_traj(a, K, nseg > 1 ? 2 : 1));
    CH_TRY(set_dts(a, dt_in, nsteps));
    CH_TRY(recompute_segment(a, a->stream, ptr6(state0), a->dt_dev, nsteps, h * (3 * K + 1),
                             h * 3 * K));
    const Ptr6 last = traj_slot(a, h * (3 * K + 1) + 3 * nsteps);
    for (int q = 0; q < 6; ++q)
        CH_HIP(hipMemcpyAsync(state_out[q], last.p[q], a->n * sizeof(double),
                              hipMemcpyDeviceToDevice, a->stream));
    CH_HIP(hipStreamSynchronize(a->stream));
    a->rec.valid = true;
    for (int q = 0; q < 6; ++q) a->rec.state0[q] = state0[q];
    a->rec.nsteps = nsteps;
    a->rec.K = K;
    a->rec.nseg = nseg;
    a->rec.dts.assign(dt_in, dt_in + nsteps);
    return CHIMERA_OK;
}

extern "C" int chimera_adj_sweep(chimera_adj_t *a, const double *const *states, const double *dt_in,
                                 int nsteps, int seg_steps, double *const *lam) {
    CH_REQUIRE(a && states && dt_in && lam, "NULL argument");
    CH_REQUIRE(nsteps >= 1 && seg_steps >= 1, "nsteps and seg_steps must be >= 1");
    for (int s = 0; s < nsteps; ++s)
        CH_REQUIRE(dt_in[s] > 0.0, "dt_in[%d] = %g is not positive", s, dt_in[s]);
    const int nseg = (nsteps + seg_steps - 1) / seg_steps;
    for (int k = 0; k < 6 * nseg; ++k) CH_REQUIRE(states[k], "states[%d] is NULL", k);
    const int K = std::min(seg_steps, nsteps);
    auto seg_len = [&](int k) { return std::min(seg_steps, nsteps - k * seg_steps); };
    // one segment: the trajectory halves alternate, so two are enough
    CH_TRY(ensure_traj(a, K, nseg > 1 ? 2 : 1));
    CH_TRY(ensure_sweep(a, nseg));
    // the last segment recorded by chimera_adj_record for exactly this run (checked after
    // the allocations, which drop a recorded segment when they reallocate)?
    const Recorded_check_t rc_last = {nseg, K, seg_len(nseg - 1), states + 6 * (size_t)(nseg - 1),
                                      dt_in + (size_t)(nseg - 1) * seg_steps};
    const bool have_last = recorded_matches(a, rc_last);
    a->rec.valid = false;
    CH_TRY(set_dts(a, dt_in, nsteps));
    CH_TRY(set_table(a, TB_LIN, (const double *const *)lam));
    CH_HIP(hipMemcpy(a->ck_tab, states, (size_t)nseg * 6 * sizeof(double *), hipMemcpyHostToDevice));
    // the recompute stream starts after everything queued on the main stream
    CH_HIP(hipEventRecord(a->ev_start, a->stream));
    CH_HIP(hipStreamWaitEvent(a->rstream, a->ev_start, 0));
    // one step s of segment k: its recompute (on the second stream, into half k & 1) and
    // its reverse (on the main stream)
    auto rec_step = [&](int k, int s) -> int {
        const int b = k & 1;
        return recompute_steps(a, a->rstream, ptr6(states + 6 * (size_t)k),
                               a->dt_dev + (size_t)k * seg_steps, s, s + 1, b * (3 * K + 1),
                               b * 3 * K);
    };
    Tab cur = table(a, TB_LIN), other = table(a, TB_LC);
    auto rev_step = [&](int k, int s) -> int {
        const int b = k & 1;
        return reverse_steps(a, ptr6(states + 6 * (size_t)k), (Tab)(a->ck_tab + 6 * (size_t)k),
                             a->dt_dev + (size_t)k * seg_steps, dt_in + (size_t)k * seg_steps, s,
                             s + 1, b * (3 * K + 1), b * 3 * K, cur, other);
    };
    if (!have_last)   // else the last segment's trajectory is already there
        for (int s = 0; s < seg_len(nseg - 1); ++s) CH_TRY(rec_step(nseg - 1, s));
    CH_HIP(hipEventRecord(a->ev_traj[(nseg - 1) & 1], a->rstream));
    for (int k = nseg - 1; k >= 0; --k) {
        CH_HIP(hipStreamWaitEvent(a->stream, a->ev_traj[k & 1], 0));
        // segment k-1 goes into the half segment k+1 used: after that reverse
        if (k >= 1 && k + 1 <= nseg - 1)
            CH_HIP(hipStreamWaitEvent(a->rstream, a->ev_rev[(k + 1) & 1], 0));
        // the reverse of segment k (steps in descending order) and the recompute of
        // segment k-1 are enqueued step by step in turn, so that both streams always
        // have work queued whatever the host's launch rate
        const int nrev = seg_len(k), nrec = k >= 1 ? seg_len(k - 1) : 0;
        for (int i = 0; i < std::max(nrev, nrec); ++i) {
            if (i < nrev) CH_TRY(rev_step(k, nrev - 1 - i));
            if (i < nrec) CH_TRY(rec_step(k - 1, i));
        }
        if (k >= 1) CH_HIP(hipEventRecord(a->ev_traj[(k - 1) & 1], a->rstream));
        CH_HIP(hipEventRecord(a->ev_rev[k & 1], a->stream));
    }
    CH_TRY(finish_lam(a, cur, lam));
    CH_HIP(hipStreamSynchronize(a->stream));
    CH_HIP(hipStreamSynchronize(a->rstream));
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
