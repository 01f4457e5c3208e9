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
#include <type_traits>

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
    // primitives of cells [-2, TX+2) x [-2, TY+2) (a 2-cell halo); rows PW apart, PW odd
    // so that the rows of neighbouring face runs fall into different LDS banks
    static constexpr int LW = TX + 4, PH = TY + 4;
    static constexpr int PW = LW | 1;
    static constexpr int FXW = TX + 1, FXH = TY + 2;   // x faces i in [-1,TX-1], j in [-1,TY]
    static constexpr int FYW = TX + 2, FYH = TY + 1;   // y faces i in [-1,TX],   j in [-1,TY-1]
    static constexpr int P_SZ = PW * PH;
    static constexpr int FX_SZ = FXW * FXH;            // transverse flux, all x faces
    static constexpr int FY_SZ = FYW * FYH;            // transverse flux, all y faces
    static constexpr int FX4_SZ = FXW * TY;            // conserved fluxes, x faces j in [0,TY-1]
    static constexpr int FY4_SZ = TX * FYH;            // conserved fluxes, y faces i in [0,TX-1]
    // one byte per wave (up to 16): some cell of the wave's phase-1 share is out of the
    // range of the short division and square root
    static constexpr int FLAG_OFF = 6 * P_SZ + 4 * FX4_SZ + FX_SZ + 4 * FY4_SZ + FY_SZ;
    static constexpr int LDS_DOUBLES = FLAG_OFF + 2;
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
// The field that holds role q (rho, u, w, P, Bn, Bt) of a y face: vx <-> vy, Bx <-> By.
__device__ constexpr int kSwapUW[6] = {0, 2, 1, 3, 5, 4};

// Limited (or central) slope of a cell from its two neighbours along one direction.
template <int LIMITER>
__device__ __forceinline__ double cell_slope(double m, double c, double p, bool use_lim,
                                             double inv_dx, double inv_2dx) {
    return use_lim ? limited_slope<LIMITER>(m, c, p, inv_dx) : central_slope(m, p, inv_2dx);
}

// Raw FP64 min / max (v_min_f64 / v_max_f64).  fmin / fmax of values the compiler
// cannot prove canonical (LDS loads, selects of them) get a canonicalising v_max of
// each operand first, to quiet signalling NaNs; the stage never holds a signalling
// NaN, and for every other input these are fmin / fmax exactly (IEEE mode: a quiet
// NaN operand yields the other one).  The MOOD bounds and face floors of a face use
// 16-20 of them.
__device__ __forceinline__ double vmin(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double vmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// the same with a wave-uniform second operand (a constant) held in SGPRs
__device__ __forceinline__ double vmax_s(double a, double s) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "s"(s));
    return r;
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
        double lo = vmin(vmin(vmin(vmin(c, p), m), nA), nB);
        const double hi = vmax(vmax(vmax(vmax(c, p), m), nA), nB);
        if (rho_or_p) lo = vmax_s(lo, 1e-9);
        if (l < lo || l > hi) { l = c; bits |= 1u; }
        if (r < lo || r > hi) { r = c; bits |= 2u; }
    }
    floor_bits = 0u;
    if (rho_or_p) {
        floor_bits = (1.0e-8 > l ? 1u : 0u) | (1.0e-8 > r ? 2u : 0u);
        l = vmax_s(l, 1.0e-8);
        r = vmax_s(r, 1.0e-8);
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
        // order: run length (longest first), direction, index k of the run within its
        // line, line -- so neighbouring lanes take the same run of neighbouring lines
        // (LDS addresses one pitch or one cell apart: few bank conflicts)
        int kmax = 0;
        for (int l = 0; l < NL; ++l) kmax = cnt[l] > kmax ? cnt[l] : kmax;
        int t = 0;
        for (int want = lmax; want >= 1; --want)
            for (int dir = 0; dir < 2; ++dir)
                for (int k = 0; k < kmax; ++k)
                    for (int l = dir ? NXL : 0; l < (dir ? NL : NXL); ++l) {
                        const int m = len_of(l), c = cnt[l], base = m / c, extra = m % c;
                        if (k >= c) continue;
                        const int len = base + (k < extra ? 1 : 0);
                        const int pos = k * base + (k < extra ? k : extra);
                        if (len == want && t < NT) {
                            r[t] = FaceRun{dir, dir ? l - NXL : l, pos, len};
                            ++t;
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

// The thermal check of both states of a face (the same values as two thermal_reset
// calls); the resets are rare, so their selects run only in waves where a face resets.
// Returns bit 0 for the left state, bit 1 for the right.
__device__ __forceinline__ unsigned thermal_pair(double L[6], double R[6], const double C[6]) {
    const double pthL = L[F_P] - 0.5 * (L[F_BX] * L[F_BX] + L[F_BY] * L[F_BY]);
    const double pthR = R[F_P] - 0.5 * (R[F_BX] * R[F_BX] + R[F_BY] * R[F_BY]);
    const bool tl = pthL < P_FLOOR, tr = pthR < P_FLOOR;
    if (__builtin_amdgcn_ballot_w64(tl || tr)) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (tl) L[q] = C[q];
            if (tr) R[q] = C[q];
        }
    }
    return (tl ? 1u : 0u) | (tr ? 2u : 0u);
}

// Face states of one field at one face without the diagnostic bits: the MUSCL left
// state c + dx/2 sl and right state p - dx/2 sr (src/mhd_derivatives.f90:150-229).
// At a non-periodic hi boundary the reference takes the cell value itself for the
// right state; here the clamped halo makes p = pp = c there, so the right cell's slope
// is exactly 0 (limited: dfL * dfR = 0; central: (c - c) / 2dx) and p - hdx * 0 = c,
// the same bits without a test.
//
// MOOD resets a state only outside [lo, hi], the extremes of the 5-point stencil of
// the left cell (Q1: the right state too), which holds c and p; a state inside
// [min(c, p), max(c, p)] (and for rho and P at or above the 1e-9 floor of lo) is
// therefore never reset.  The limited states of smooth data are inside: the return
// value says whether they are, and only in waves where some state is not does
// mood_full read the other three stencil members and make the full test (the same
// result for every lane).
__device__ __forceinline__ bool face_states_inside(double c, double p, double sl, double sr,
                                                   bool rho_or_p, double hdx, double &fl,
                                                   double &fr) {
    const double l = c + hdx * sl;
    const double r = p - hdx * sr;
    const double mn = vmin(c, p), mx = vmax(c, p);
    bool inside = l >= mn && l <= mx && r >= mn && r <= mx;
    if (rho_or_p) inside = inside && l >= 1e-9 && r >= 1e-9;
    fl = l;
    fr = r;
    return inside;
}
// The full MOOD test of one field's states; b points at the left cell c in LDS, p =
// b[sA] is the next cell along the face normal, the other stencil members are b[-sA],
// b[sT] and b[-sT].
__device__ __forceinline__ void mood_full(const double *b, int sA, int sT, double c,
                                          bool rho_or_p, double &fl, double &fr) {
    const double p = b[sA], m = b[-sA], nA = b[sT], nB = b[-sT];
    double lo = vmin(vmin(vmin(vmin(c, p), m), nA), nB);
    const double hi = vmax(vmax(vmax(vmax(c, p), m), nA), nB);
    if (rho_or_p) lo = vmax_s(lo, 1e-9);
    if (fl < lo || fl > hi) fl = c;
    if (fr < lo || fr > hi) fr = c;
}

// Global loads / stores at a 32-bit byte offset from a wave-uniform base (the
// scalar-base form of the global instructions: no 64-bit address arithmetic per
// access).  Grids of the fused path stay below 2^29 cells per field.  The base is
// taken in the global address space: every state array is device memory, and a
// pointer the compiler cannot place (one read from a descriptor table, as in the
// batched kernels) would otherwise become a flat access -- a 64-bit address per lane,
// and a load that also counts against the LDS wait counter.
typedef const char __attribute__((address_space(1))) *gchar_cptr;
typedef char __attribute__((address_space(1))) *gchar_ptr;
__device__ __forceinline__ double ldg(const double *base, uint32_t off) {
    const gchar_cptr p = (gchar_cptr)(reinterpret_cast<const char *>(base)) + off;
    return *(const double __attribute__((address_space(1))) *)p;
}
__device__ __forceinline__ void stg(double *base, uint32_t off, double v) {
    const gchar_ptr p = (gchar_ptr)(reinterpret_cast<char *>(base)) + off;
    *(double __attribute__((address_space(1))) *)p = v;
}

// True when v is a positive power of two (division by it is then exactly a
// multiplication by 1 / v, for every dividend).
__device__ __forceinline__ bool pow2_value(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long e = (b >> 52) & 0x7ffull;
    return (b >> 63) == 0 && e != 0 && e != 0x7ff && (b & 0xfffffffffffffull) == 0;
}

// get_primitive (src/mhd_change_states.f90:91-110) with the two divisions by the cell
// volume done as multiplications when vol is a power of two (the same bits).
template <bool VOL_POW2>
__device__ __forceinline__ void cons_to_prim_v(double M, double Mx, double My, double E, double Bx,
                                               double By, double g, double vol, double ivol,
                                               double &rho, double &vx, double &vy, double &P) {
    double r = fmax(VOL_POW2 ? M * ivol : M / vol, RHO_FLOOR);
    double u = Mx / (r * vol);
    double w = My / (r * vol);
    double halfB2 = 0.5 * (Bx * Bx + By * By);
    double ke = 0.5 * r * (u * u + w * w);
    double p = ((VOL_POW2 ? E * ivol : E / vol) - ke - halfB2) * (g - 1.0) + halfB2;
    p = fmax(p, P_FLOOR);
    p = fmax(p, halfB2 + E_FLOOR_FRAC * (ke + halfB2));
    rho = r;
    vx = u;
    vy = w;
    P = p;
}

// apply_conserved_floors (src/mhd_change_states.f90:143-164) as conserved_floors, with
// the velocity cap's square root and division skipped when |Mom|^2 < 49^2 M^2: then
// |Mom| / M < 49 (1 + 2^-53)^2 after every rounding, so the cap test is false and the
// result is unchanged.  A NaN or infinite |Mom|^2 fails the pre-test and takes the
// reference's own test.
__device__ __forceinline__ void conserved_floors_fast(double &M, double &Mx, double &My, double &E,
                                                      double Bx, double By, double vol, double g) {
    const double V_MAX = 50.0;
    if (M < RHO_FLOOR * vol) {
        double halfB2 = 0.5 * (Bx * Bx + By * By);
        Mx = 0.0;
        My = 0.0;
        M = RHO_FLOOR * vol;
        E = fmax(E, (P_FLOOR / (g - 1.0) + halfB2) * vol);
    }
    const double s2 = Mx * Mx + My * My;
    if (!(s2 < 2401.0 * (M * M))) {
        double mom = sqrt(s2);
        if (mom / M > V_MAX) {
            double scale = V_MAX * M / mom;
            Mx = Mx * scale;
            My = My * scale;
        }
    }
}

// MASK: also OR the MOOD / thermal bits of owned faces into g.mask (the stage-1
// diagnostics); the bits are assembled only with MASK or STORE_ALL.
// LM: slope limiting and MOOD known to be on (the usual scheme): no per-field branch
// on either flag.
template <int SOLVER, int LIMITER, int TX, int TY, int STORE, bool MASK, bool LM>
__device__ __forceinline__ void stage_tile_impl(const GridView &g, int stage, bool use_lim_arg,
                                                bool mood_arg, int tile_x, int tile_y,
                                                double *lds, const StageStore *store) {
    const bool use_lim = LM || use_lim_arg, mood = LM || mood_arg;
    using S = TileShape<TX, TY>;
    constexpr int NT = TX * TY;
    constexpr bool BITS = MASK || STORE == STORE_ALL;
    // the thread index made opaque per call: in a persistent loop over tiles the
    // compiler would otherwise keep every thread-dependent value of the tile (run
    // table entry, LDS pointers, strides) live in registers across the whole loop
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
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
    // Rusanov and HLLE take the short division and square root (the same bits) in a
    // tile whose stencil cells all have rho + P + |vx| + |vy| + |Bx| + |By| <= 2^64
    // (rho, P >= 1e-12 by the floors), with dx and g - 1 in [2^-100, 2^100].  The face
    // states are then within 2^66 in magnitude (limited or central slopes move a state
    // by at most one neighbour difference), rho and P at least 1e-8, and the one-sided
    // differences / dx within 2^165: the ranges the short forms need.
    // (Not in the diagnostic variants, whose mask bits could not be taken back.)
    static_assert(NT % 64 == 0 && NT / 64 <= 16, "wave flags hold 16 waves");
    constexpr bool SHORT_FORMS = SOLVER != HLLD && !BITS;
    const bool launch_short = SHORT_FORMS && dx >= 0x1p-100 && dx <= 0x1p100 &&
                              gam - 1.0 >= 0x1p-100 && gam <= 0x1p100;
    bool out_of_range = false;

    // ---- phase 1: halo load and stage primitives -------------------------------
    auto phase1 = [&](auto vol_pow2) {
        constexpr bool P2 = decltype(vol_pow2)::value;
        constexpr int NC = S::LW * S::PH, IT = (NC + NT - 1) / NT;
        static_assert(NC >= NT, "every thread has a first halo cell");
        const double ivol = 1.0 / vol;   // exact when vol is a power of two
        // every pass's loads first (M, Mx, My, E, the two x faces and the two y faces
        // of the cell), so that the passes share one round of memory latency
        double raw[IT][8];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int n = tid + it * NT;
            if (it == 0 || n < NC) {
                const int lx = n % S::LW, ly = n / S::LW;
                const int gi = map_index(x0 - 2 + lx, nx, px);
                const int gj = map_index(y0 - 2 + ly, ny, py);
                const int gim = gi > 0 ? gi - 1 : (px ? nx - 1 : 0);
                const int gjm = gj > 0 ? gj - 1 : (py ? ny - 1 : 0);
                const uint32_t row = (uint32_t)nx * (uint32_t)gj;
                const uint32_t k = (row + (uint32_t)gi) * 8u;
                const uint32_t kim = (row + (uint32_t)gim) * 8u;
                const uint32_t kjm = ((uint32_t)nx * (uint32_t)gjm + (uint32_t)gi) * 8u;
#pragma unroll
                for (int q = 0; q < 4; ++q) raw[it][q] = ldg(g.uin[q], k);
                raw[it][4] = ldg(g.uin[4], k);
                raw[it][5] = ldg(g.uin[4], kim);
                raw[it][6] = ldg(g.uin[5], k);
                raw[it][7] = ldg(g.uin[5], kjm);
            }
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int n = tid + it * NT;
            if (it > 0 && n >= NC) continue;
            const int lx = n % S::LW, ly = n / S::LW;
            const int idx = ly * S::PW + lx;
            const double Bx = 0.5 * (raw[it][4] + raw[it][5]);
            const double By = 0.5 * (raw[it][6] + raw[it][7]);
            double rho, vx, vy, P;
            cons_to_prim_v<P2>(raw[it][0], raw[it][1], raw[it][2], raw[it][3], Bx, By, gam, vol,
                               ivol, rho, vx, vy, P);
            // a NaN or infinity fails the test too
            out_of_range |=
                !(rho + P + fabs(vx) + fabs(vy) + fabs(Bx) + fabs(By) <= 0x1p64);
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
    };
    if (pow2_value(vol)) phase1(std::true_type{});
    else phase1(std::false_type{});
    unsigned char *wave_flag = reinterpret_cast<unsigned char *>(lds + S::FLAG_OFF);
    {
        const bool wave_out = __builtin_amdgcn_ballot_w64(out_of_range) != 0;
        if ((tid & 63) == 0) wave_flag[tid >> 6] = wave_out ? 1 : 0;
    }
    __syncthreads();
    bool tile_short;
    {
        constexpr int NW = NT / 64;
        const unsigned long long *wf = reinterpret_cast<const unsigned long long *>(wave_flag);
        unsigned long long f = wf[0];
        if (NW > 8) f |= wf[1];
        if (NW < 8) f &= (1ull << (8 * NW)) - 1;
        tile_short = __builtin_amdgcn_readfirstlane(launch_short && f == 0 ? 1 : 0) != 0;
    }

    // ---- phase 2: every x face and y face the tile needs, as runs of consecutive faces
    //      along their normal (one run per thread): the slope of a run's right cell is
    //      the next face's left-cell slope, so each slope is evaluated once per run.  x
    //      and y runs share one code path (strides and field roles by direction), so a
    //      wave never splits on the direction. ------------------------------------------
    auto phase2 = [&](auto short_forms) {
        constexpr bool SH = decltype(short_forms)::value;
        const Recip rg = SH ? recip_uniform(gam - 1.0) : Recip{};
        const FaceRun run = kFaceRuns<TX, TY>.r[tid];
        const int dir = run.dir;                        // 0: x faces, 1: y faces (rotated)
        const int sA = dir ? S::PW : 1;                 // LDS stride along the face normal
        const int sT = dir ? 1 : S::PW;                 // LDS stride across it
        // x face: i = fx-1 in row j = fy-1, left cell at LDS (fx+1, fy+1);
        // y face: column i = fx-1, j = fy-1, lower cell at LDS (fx+1, fy+1)
        const int fx0 = dir ? run.line : run.start, fy0 = dir ? run.start : run.line;
        const double *b0 = sP + (fy0 + 1) * S::PW + (fx0 + 1);   // left cell, advances by sA
        // transverse-flux slot of the face (every face) and its step along the run
        double *bt = dir ? sBY + fy0 * S::FYW + fx0 : sBX + fy0 * S::FXW + fx0;
        const int bt_step = dir ? S::FYW : 1;
        // conserved-flux slots: faces of the tile's own rows (x) / columns (y) only; the
        // normal / tangential momentum fluxes go to Momx / Momy by direction
        const bool cons = run.line >= 1 && run.line <= (dir ? TX : TY);
        const int fsz = dir ? S::FY4_SZ : S::FX4_SZ;
        double *cf = dir ? sFY + fy0 * TX + (fx0 - 1) : sFX + (fy0 - 1) * S::FXW + fx0;
        const int cf_step = dir ? TX : 1;
        const int off_n = dir ? 2 * fsz : fsz, off_t = dir ? fsz : 2 * fsz;
        // The face states are built in the Riemann solvers' roles (rho, u, w, P, Bn, Bt):
        // u / Bn the normal, w / Bt the tangential components, so y faces read vy for u,
        // vx for w, By for Bn and Bx for Bt -- an offset of one field in the LDS address
        // instead of a select of every value.
        const int so = dir ? S::P_SZ : 0;
        auto role = [&](const double *bf, int q) -> const double * {
            return (q == 1 || q == 4) ? bf + q * S::P_SZ + so
                                      : ((q == 2 || q == 5) ? bf + q * S::P_SZ - so : bf + q * S::P_SZ);
        };
        // slope of the first face's left cell; afterwards each face's right-cell slope
        // is carried to the next face, and with limiting the right cell's one-sided
        // difference dfR, which is the next right cell's dfL
        double sprev[6], dprev[6];
        if (run.len > 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double *b = role(b0, q);
                if (use_lim) {
                    const double dR = (b[sA] - b[0]) * inv_dx;
                    sprev[q] = limiter_of<LIMITER>((b[0] - b[-sA]) * inv_dx, dR);
                    dprev[q] = dR;
                } else {
                    sprev[q] = central_slope(b[-sA], b[sA], inv_2dx);
                }
            }
        }
        for (int t = 0; t < run.len; ++t) {
            const double *bf = b0 + t * sA;
            double L[6], R[6], C[6];
            unsigned bits = 0, fbits = 0;
            // Van Leer (limiting on): the six fields' quotients 2 dfL dfR / (dfL + dfR)
            // side by side in one branch, taken in waves where some field's slope is
            // not zero (limiter_of's value, field by field)
            double vl[6];
            if constexpr (LM && LIMITER == LIM_VAN_LEER && !BITS) {
                double num[6], den[6];
                bool nz[6], any = false;
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const double *b = role(bf, q);
                    const double dR = (b[2 * sA] - b[sA]) * inv_dx;
                    const double dL = dprev[q];
                    const double prod = dL * dR;
                    den[q] = dL + dR;
                    num[q] = 2.0 * prod;
                    nz[q] = !(prod <= 0.0 || fabs(den[q]) < 1e-12);
                    any = any || nz[q];
                    dprev[q] = dR;
                }
                if (__builtin_amdgcn_ballot_w64(any)) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) vl[q] = nz[q] ? num[q] / den[q] : 0.0;
                } else {
#pragma unroll
                    for (int q = 0; q < 6; ++q) vl[q] = 0.0;
                }
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const double *b = role(bf, q);
                const double c = b[0], p = b[sA], pp = b[2 * sA];
                C[q] = c;
                const double sl = sprev[q];
                double sr;
                if (LM && LIMITER == LIM_VAN_LEER && !BITS) {
                    sr = vl[q];
                } else if (use_lim) {
                    const double dR = (pp - p) * inv_dx;
                    sr = limiter_of<LIMITER>(dprev[q], dR);
                    dprev[q] = dR;
                } else {
                    sr = central_slope(c, pp, inv_2dx);
                }
                sprev[q] = sr;
                if (BITS) {
                    // bits by field: role q of a y face is field kSwapUW[q]
                    const int fq = dir ? kSwapUW[q] : q;
                    unsigned fb;
                    bits |= face_states(b[-sA], c, p, b[sT], b[-sT], sl, sr, false, mood,
                                        q == F_RHO || q == F_P, hdx, L[q], R[q], fb)
                            << (4 * fq + 2 * dir);
                    if (q == F_RHO) fbits |= fb << (2 * dir);
                    if (q == F_P) fbits |= fb << (4 + 2 * dir);
                } else {
                    // MOOD's full test only in waves where some state is not inside
                    // [c, p]; then the rho / P floor of the states
                    const bool rp = q == F_RHO || q == F_P;
                    const bool in = face_states_inside(c, p, sl, sr, rp, hdx, L[q], R[q]);
                    if (mood && __builtin_amdgcn_ballot_w64(!in))
                        mood_full(b, sA, sT, c, rp, L[q], R[q]);
                    if (rp) {
                        L[q] = vmax_s(L[q], 1.0e-8);
                        R[q] = vmax_s(R[q], 1.0e-8);
                    }
                }
            }
            // thermal-pressure check (src/mhd_derivatives.f90:262-267): both faces reset
            // to the cell values below P_floor; rare, so the selects run only in waves
            // where some face resets.  In roles the magnetic term is Bn^2 + Bt^2, for y
            // faces By^2 + Bx^2: the same sum (IEEE addition commutes).
            {
                const unsigned tb = thermal_pair(L, R, C);
                if (BITS) bits |= tb << (24 + 2 * dir);
            }
            const int fx = dir ? run.line : run.start + t;
            const int fy = dir ? run.start + t : run.line;
            const int i = fx - 1, j = fy - 1;
            if (STORE == STORE_ALL || MASK) {
                const bool owned = i >= 0 && i < TX && j >= 0 && j < TY && x0 + i < nx && y0 + j < ny;
                if (STORE == STORE_ALL && owned) {
                    size_t ko = (size_t)(x0 + i) + (size_t)nx * (y0 + j);
#pragma unroll
                    for (int q = 0; q < 6; ++q) {
                        const int fq = dir ? kSwapUW[q] : q;
                        store->face[4 * fq + 2 * dir][ko] = L[q];
                        store->face[4 * fq + 2 * dir + 1][ko] = R[q];
                    }
                    unsigned long long w = (unsigned long long)bits | ((unsigned long long)fbits << 32);
                    if (dir) store->bits_y[ko] = w;
                    else store->bits_x[ko] = w;
                }
                if (MASK && owned && bits)
                    atomicOr(&g.mask[(size_t)(x0 + i) + (size_t)nx * (y0 + j)], bits);
            }
            FaceState a{L[0], L[1], L[2], L[3], L[4], L[5]};
            FaceState bR{R[0], R[1], R[2], R[3], R[4], R[5]};
            Flux5 f = riemann<SOLVER, SH>(a, bR, gam, rg);
            // transverse flux: By flux (x faces) / Bx flux (y faces), every face
            bt[t * bt_step] = f.bt;
            // conserved fluxes: faces of the tile's own rows (x) / columns (y)
            if (cons) {
                double *c0 = cf + t * cf_step;
                c0[0] = f.mass;
                c0[off_n] = f.mom_n;
                c0[off_t] = f.mom_t;
                c0[3 * fsz] = f.energy;
            }
        }
    };
    if constexpr (SHORT_FORMS) {
        if (tile_short) phase2(std::true_type{});
        else phase2(std::false_type{});
    } else {
        phase2(std::false_type{});
    }
    // phase 3's global loads (the cell's stage input and u^n, which this launch does not
    // write) are issued before the barrier, so that their latency overlaps the wait for
    // the other waves' faces; the face loop's registers are free by now
    const int ci = tid % TX, cj = tid / TX;
    const int gi = x0 + ci, gj = y0 + cj;
    const bool cell_active = gi < nx && gj < ny;
    const uint32_t k = ((uint32_t)nx * (uint32_t)gj + (uint32_t)gi) * 8u;
    double Uc[6], U0[6];
    if (cell_active) {
#pragma unroll
        for (int q = 0; q < 6; ++q) Uc[q] = ldg(g.uin[q], k);
        if (stage > 1) {
#pragma unroll
            for (int q = 0; q < 6; ++q) U0[q] = ldg(g.u0[q], k);
        }
    }
    const double dt = ldg(g.dt, 0u);
    __syncthreads();

    // ---- phase 3: update, floors, CT, outflow correction, blend (one cell per thread)
    {
        if (!cell_active) return;
        const double dtdx = dt * dx;
        // conserved fluxes: x faces FX(i, j) at [j][i+1], y faces FY(i, j) at [j+1][i];
        // the lower faces' slots exist for every cell, so they are read unconditionally
        // and dropped by a select at an open lower boundary
        const int xr = cj * S::FXW + (ci + 1);
        const int yr = (cj + 1) * TX + ci;
        const bool xlo_open = (gi == 0) && !px;
        const bool ylo_open = (gj == 0) && !py;
        double U[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double fr = sFX[q * S::FX4_SZ + xr], fl = sFX[q * S::FX4_SZ + xr - 1];
            const double ft = sFY[q * S::FY4_SZ + yr], fb = sFY[q * S::FY4_SZ + yr - TX];
            double v = Uc[q] - dtdx * fr;
            v = xlo_open ? v : v + dtdx * fl;
            v = v - dtdx * ft;
            v = ylo_open ? v : v + dtdx * fb;
            U[q] = v;
            if (STORE != STORE_NONE) stg(store->Up[q], k, v);
        }
        const int pc = (cj + 2) * S::PW + (ci + 2);
        const double Bxc = sP[4 * S::P_SZ + pc], Byc = sP[5 * S::P_SZ + pc];
        conserved_floors_fast(U[0], U[1], U[2], U[3], Bxc, Byc, vol, gam);

        // corner EMF: Ez(a,b) = 1/4 (-F_x(By)(a,b) - F_x(By)(a,b+1) + F_y(Bx)(a,b) + F_y(Bx)(a+1,b));
        // at a closed hi boundary the outer flux is the inner one (both slots are read:
        // for the corners used below they lie inside the flux arrays)
        auto negEz = [&](int a, int b) -> double {
            const double f0 = sBX[(b + 1) * S::FXW + (a + 1)];
            const double f1 = sBX[(b + 2) * S::FXW + (a + 1)];
            const double e0 = sBY[(b + 1) * S::FYW + (a + 1)];
            const double e1 = sBY[(b + 1) * S::FYW + (a + 2)];
            const double up = (!py && y0 + b == ny - 1) ? f0 : f1;
            const double rt = (!px && x0 + a == nx - 1) ? e0 : e1;
            double Ez = 0.25 * (-f0 - up + e0 + rt);
            return -Ez;
        };
        double dbx, dby;
        if (ylo_open) dbx = (negEz(ci, cj + 1) - negEz(ci, cj)) * inv_dx;
        else dbx = (negEz(ci, cj) - negEz(ci, cj - 1)) * inv_dx;
        if (xlo_open) dby = -((negEz(ci + 1, cj) - negEz(ci, cj)) * inv_dx);
        else dby = -((negEz(ci, cj) - negEz(ci - 1, cj)) * inv_dx);
        double bxn = Uc[4] + dt * dbx;
        double byn = Uc[5] + dt * dby;

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
            for (int q = 0; q < 4; ++q) stg(store->U3[q], k, U[q]);
        }
        if (stage > 1) {
            const double c0 = (stage == 2) ? 0.75 : 1.0 / 3.0;
            const double c1 = (stage == 2) ? 0.25 : 2.0 / 3.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) U[q] = c0 * U0[q] + c1 * U[q];
            bxn = c0 * U0[4] + c1 * bxn;
            byn = c0 * U0[5] + c1 * byn;
            conserved_floors_fast(U[0], U[1], U[2], U[3], Bxc, Byc, vol, gam);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) stg(g.uout[q], k, U[q]);
        stg(g.uout[4], k, bxn);
        stg(g.uout[5], k, byn);
    }
}

template <int SOLVER, int LIMITER, int TX, int TY, int STORE = STORE_NONE, bool MASK = false>
__device__ __forceinline__ void stage_tile(const GridView &g, int stage, bool use_lim, bool mood,
                                           int tile_x, int tile_y, double *lds,
                                           const StageStore *store = nullptr) {
    if (use_lim && mood)
        stage_tile_impl<SOLVER, LIMITER, TX, TY, STORE, MASK, true>(g, stage, true, true, tile_x,
                                                                    tile_y, lds, store);
    else
        stage_tile_impl<SOLVER, LIMITER, TX, TY, STORE, MASK, false>(g, stage, use_lim, mood,
                                                                     tile_x, tile_y, lds, store);
}

}  // namespace chimera
