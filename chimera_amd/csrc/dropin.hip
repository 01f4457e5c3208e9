// Drop-in replacements for the reference's hot-path routines: one GPU kernel per
// Fortran routine (mhd_bc, mhd_derivatives, mhd_flux, mhd_field_ops,
// mhd_change_states), with the reference's argument meaning.  The chimera_* entry
// points take host pointers and copy per call; chimera_dev_* take device pointers.
#include <vector>

#include "common.hpp"
#include "mhd_phys.hpp"

namespace chimera {

namespace {

constexpr int BS = 256;

inline int blocks_for(size_t n) {
    size_t b = (n + BS - 1) / BS;
    if (b < 1) b = 1;
    if (b > 65535) b = 65535;
    return (int)b;
}

#define GRID_LOOP(idx, n) \
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (n); \
         idx += (size_t)gridDim.x * blockDim.x)

__device__ __forceinline__ bool is_fixed_like(int f) { return f == 2 || f == 4; }

// fill_ghost_cells / fill_corners / corner_val, src/mhd_bc.f90:36-170
__global__ void k_fill_ghost(const double *f, double *pad, const double *amb, int nx, int ny,
                             int xlo, int xhi, int ylo, int yhi) {
    const int W = nx + 2;
    GRID_LOOP(idx, (size_t)(nx + 2) * (ny + 2)) {
        int pi = (int)(idx % W), pj = (int)(idx / W);
        int i = pi - 1, j = pj - 1;
        bool xin = i >= 0 && i < nx, yin = j >= 0 && j < ny;
        double v;
        if (xin && yin) {
            v = f[i + (size_t)nx * j];
        } else if (yin) {  // left / right ghost column
            if (i < 0) {
                if (xlo == BC_PERIODIC) v = f[(nx - 1) + (size_t)nx * j];
                else if (is_fixed_like(xlo) && amb) v = amb[0 + (size_t)nx * j];
                else v = f[0 + (size_t)nx * j];
            } else {
                if (xhi == BC_PERIODIC) v = f[0 + (size_t)nx * j];
                else if (is_fixed_like(xhi) && amb) v = amb[(nx - 1) + (size_t)nx * j];
                else v = f[(nx - 1) + (size_t)nx * j];
            }
        } else if (xin) {  // bottom / top ghost row
            if (j < 0) {
                if (ylo == BC_PERIODIC) v = f[i + (size_t)nx * (ny - 1)];
                else if (is_fixed_like(ylo) && amb) v = amb[i];
                else v = f[i];
            } else {
                if (yhi == BC_PERIODIC) v = f[i];
                else if (is_fixed_like(yhi) && amb) v = amb[i + (size_t)nx * (ny - 1)];
                else v = f[i + (size_t)nx * (ny - 1)];
            }
        } else {  // corner: nearest interior corner, wrapped along periodic sides
            int ix = i < 0 ? 0 : nx - 1, iy = j < 0 ? 0 : ny - 1;
            int bcx = i < 0 ? xlo : xhi, bcy = j < 0 ? ylo : yhi;
            int sx = (bcx == BC_PERIODIC) ? nx - 1 - ix : ix;
            int sy = (bcy == BC_PERIODIC) ? ny - 1 - iy : iy;
            v = f[sx + (size_t)nx * sy];
        }
        pad[idx] = v;
    }
}

// compute_gradients, src/mhd_derivatives.f90:15-47
__global__ void k_gradients(const double *pad, double dx, int nx, int ny, double *gx,
                            double *gy) {
    const int W = nx + 2;
    const double inv_2dx = 1.0 / (2.0 * dx);
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        size_t c = (size_t)(i + 1) + (size_t)W * (j + 1);
        gx[k] = (pad[c + 1] - pad[c - 1]) * inv_2dx;
        gy[k] = (pad[c + W] - pad[c - W]) * inv_2dx;
    }
}

// apply_slope_limiter, src/mhd_derivatives.f90:50-127
template <int LIMITER>
__global__ void k_limiter(const double *pad, double dx, int nx, int ny, double *gx,
                          double *gy) {
    const int W = nx + 2;
    const double inv_dx = 1.0 / dx;
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        size_t c = (size_t)(i + 1) + (size_t)W * (j + 1);
        gx[k] = limited_slope<LIMITER>(pad[c - 1], pad[c], pad[c + 1], inv_dx);
        gy[k] = limited_slope<LIMITER>(pad[c - W], pad[c], pad[c + W], inv_dx);
    }
}

// reconstruction, src/mhd_derivatives.f90:130-229
__global__ void k_reconstruction(const double *f, const double *gx, const double *gy, double dx,
                                 int nx, int ny, double *XL, double *XR, double *YL, double *YR,
                                 int rho_or_p, int xlo, int xhi, int ylo, int yhi, int mood,
                                 uint32_t *mask, int field) {
    const double hdx = 0.5 * dx;
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        double c = f[k];
        double xl = c + hdx * gx[k];
        double yl = c + hdx * gy[k];
        double xr, yr;
        if (i < nx - 1) xr = f[k + 1] - hdx * gx[k + 1];
        else if (xhi == BC_PERIODIC) xr = f[(size_t)nx * j] - hdx * gx[(size_t)nx * j];
        else xr = c;
        if (j < ny - 1) yr = f[k + nx] - hdx * gy[k + nx];
        else if (yhi == BC_PERIODIC) yr = f[i] - hdx * gy[i];
        else yr = c;
        unsigned bits = 0;
        if (mood) {
            int ip1 = (xhi == BC_PERIODIC) ? (i + 1) % nx : min(i + 1, nx - 1);
            int im1 = (xlo == BC_PERIODIC) ? (i - 1 + nx) % nx : max(i - 1, 0);
            int jp1 = (yhi == BC_PERIODIC) ? (j + 1) % ny : min(j + 1, ny - 1);
            int jm1 = (ylo == BC_PERIODIC) ? (j - 1 + ny) % ny : max(j - 1, 0);
            double a1 = f[ip1 + (size_t)nx * j], a2 = f[im1 + (size_t)nx * j];
            double a3 = f[i + (size_t)nx * jp1], a4 = f[i + (size_t)nx * jm1];
            double lo = fmin(fmin(fmin(fmin(c, a1), a2), a3), a4);
            double hi = fmax(fmax(fmax(fmax(c, a1), a2), a3), a4);
            if (rho_or_p) lo = fmax(1e-9, lo);
            if (xl < lo || xl > hi) { xl = c; bits |= 1u; }
            if (xr < lo || xr > hi) { xr = c; bits |= 2u; }
            if (yl < lo || yl > hi) { yl = c; bits |= 4u; }
            if (yr < lo || yr > hi) { yr = c; bits |= 8u; }
        }
        if (rho_or_p) {
            xl = fmax(xl, 1.0e-8);
            xr = fmax(xr, 1.0e-8);
            yl = fmax(yl, 1.0e-8);
            yr = fmax(yr, 1.0e-8);
        }
        XL[k] = xl;
        XR[k] = xr;
        YL[k] = yl;
        YR[k] = yr;
        if (mask && bits) mask[k] |= bits << (4 * field);
    }
}

struct ThermalArgs {
    double *face[24];        // P, Bx, By, rho, vx, vy groups of XL, XR, YL, YR
    const double *cell[6];   // rho, vx, vy, P, Bx, By
};

// thermal_pressure_check, src/mhd_derivatives.f90:232-293
__global__ void k_thermal(ThermalArgs a, size_t n, uint32_t *mask) {
    const int grp_cell[6] = {3, 4, 5, 0, 1, 2};
    GRID_LOOP(k, n) {
        unsigned bits = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            double P = a.face[s][k], Bx = a.face[4 + s][k], By = a.face[8 + s][k];
            double p_th = P - 0.5 * (Bx * Bx + By * By);
            if (p_th < P_FLOOR) {
#pragma unroll
                for (int g = 0; g < 6; ++g) a.face[4 * g + s][k] = a.cell[grp_cell[g]][k];
                bits |= 1u << (24 + s);
            }
        }
        if (mask && bits) mask[k] |= bits;
    }
}

struct FluxArgs {
    const double *in[12];  // rho_L,rho_R,vx_L,vx_R,vy_L,vy_R,P_L,P_R,Bx_L,Bx_R,By_L,By_R
    double *out[5];        // Mass, Momx, Momy, Energy, By
};

// compute_fluxes, src/mhd_flux.f90:139-191
template <int SOLVER>
__global__ void k_fluxes(FluxArgs a, size_t n, double gamma) {
    GRID_LOOP(k, n) {
        FaceState L{a.in[0][k], a.in[2][k], a.in[4][k], a.in[6][k], a.in[8][k], a.in[10][k]};
        FaceState R{a.in[1][k], a.in[3][k], a.in[5][k], a.in[7][k], a.in[9][k], a.in[11][k]};
        Flux5 f = riemann<SOLVER>(L, R, gamma);
        a.out[0][k] = f.mass;
        a.out[1][k] = f.mom_n;
        a.out[2][k] = f.mom_t;
        a.out[3][k] = f.energy;
        a.out[4][k] = f.bt;
    }
}

// update_conserved, src/mhd_flux.f90:20-76
__global__ void k_update(double *F, const double *fX, const double *fY, double dx, double dt,
                         int nx, int ny, int px, int py) {
    const double dtdx = dt * dx;
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        double v = F[k];
        if (i > 0) v = v - dtdx * fX[k] + dtdx * fX[k - 1];
        else if (px) v = v - dtdx * fX[k] + dtdx * fX[k + nx - 1];
        else v = v - dtdx * fX[k];
        if (j > 0) v = v - dtdx * fY[k] + dtdx * fY[k - nx];
        else if (py) v = v - dtdx * fY[k] + dtdx * fY[k + (size_t)nx * (ny - 1)];
        else v = v - dtdx * fY[k];
        F[k] = v;
    }
}

// -Ez at corner (i, j), src/mhd_flux.f90:108-129
__device__ __forceinline__ double neg_ez(const double *fByX, const double *fBxY, int i, int j,
                                         int nx, int ny, int pxh, int pyh) {
    size_t k = (size_t)i + (size_t)nx * j;
    double up = (j < ny - 1) ? fByX[k + nx] : (pyh ? fByX[i] : fByX[k]);
    double rt = (i < nx - 1) ? fBxY[k + 1] : (pxh ? fBxY[(size_t)nx * j] : fBxY[k]);
    double Ez = 0.25 * (-fByX[k] - up + fBxY[k] + rt);
    return -Ez;
}

// constrained_transport, src/mhd_flux.f90:79-136 (curl of -Ez as compute_curl_2d)
__global__ void k_ct(double *bx, double *by, const double *fByX, const double *fBxY, double dx,
                     double dt, int nx, int ny, int pxl, int pxh, int pyl, int pyh) {
    const double inv_dx = 1.0 / dx;
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        double dbx, dby;
        if (j > 0)
            dbx = (neg_ez(fByX, fBxY, i, j, nx, ny, pxh, pyh) -
                   neg_ez(fByX, fBxY, i, j - 1, nx, ny, pxh, pyh)) * inv_dx;
        else if (pyl)
            dbx = (neg_ez(fByX, fBxY, i, 0, nx, ny, pxh, pyh) -
                   neg_ez(fByX, fBxY, i, ny - 1, nx, ny, pxh, pyh)) * inv_dx;
        else
            dbx = (neg_ez(fByX, fBxY, i, 1, nx, ny, pxh, pyh) -
                   neg_ez(fByX, fBxY, i, 0, nx, ny, pxh, pyh)) * inv_dx;
        if (i > 0)
            dby = -((neg_ez(fByX, fBxY, i, j, nx, ny, pxh, pyh) -
                     neg_ez(fByX, fBxY, i - 1, j, nx, ny, pxh, pyh)) * inv_dx);
        else if (pxl)
            dby = -((neg_ez(fByX, fBxY, 0, j, nx, ny, pxh, pyh) -
                     neg_ez(fByX, fBxY, nx - 1, j, nx, ny, pxh, pyh)) * inv_dx);
        else
            dby = -((neg_ez(fByX, fBxY, 1, j, nx, ny, pxh, pyh) -
                     neg_ez(fByX, fBxY, 0, j, nx, ny, pxh, pyh)) * inv_dx);
        bx[k] = bx[k] + dt * dbx;
        by[k] = by[k] + dt * dby;
    }
}

// compute_curl_2d, src/mhd_field_ops.f90:15-57
__global__ void k_curl(const double *Az, double dx, int nx, int ny, double *bx, double *by,
                       int pxl, int pyl) {
    const double inv_dx = 1.0 / dx;
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        if (j > 0) bx[k] = (Az[k] - Az[k - nx]) * inv_dx;
        else if (pyl) bx[k] = (Az[k] - Az[k + (size_t)nx * (ny - 1)]) * inv_dx;
        else bx[k] = (Az[k + nx] - Az[k]) * inv_dx;
        if (i > 0) by[k] = -((Az[k] - Az[k - 1]) * inv_dx);
        else if (pxl) by[k] = -((Az[k] - Az[k + nx - 1]) * inv_dx);
        else by[k] = -((Az[k + 1] - Az[k]) * inv_dx);
    }
}

// compute_divB, src/mhd_field_ops.f90:60-117
__global__ void k_divb(const double *bx, const double *by, double dx, int nx, int ny,
                       double *divB, int pxl, int pyl) {
    const double inv_dx = 1.0 / dx;
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        bool xin = i > 0 || pxl, yin = j > 0 || pyl;
        size_t km = i > 0 ? k - 1 : k + nx - 1;
        size_t kb = j > 0 ? k - nx : k + (size_t)nx * (ny - 1);
        double d;
        if (xin && yin) d = (bx[k] - bx[km] + by[k] - by[kb]) * inv_dx;
        else if (xin) d = (bx[k] - bx[km]) * inv_dx;
        else if (yin) d = (by[k] - by[kb]) * inv_dx;
        else d = 0.0;
        divB[k] = d;
    }
}

// average_face_to_cell_B, src/mhd_field_ops.f90:120-156
__global__ void k_avg_b(const double *b_x, const double *b_y, int nx, int ny, double *Bx,
                        double *By, int pxl, int pyl) {
    GRID_LOOP(k, (size_t)nx * ny) {
        int i = (int)(k % nx), j = (int)(k / nx);
        if (i > 0) Bx[k] = 0.5 * (b_x[k] + b_x[k - 1]);
        else Bx[k] = pxl ? 0.5 * (b_x[k] + b_x[k + nx - 1]) : b_x[k];
        if (j > 0) By[k] = 0.5 * (b_y[k] + b_y[k - nx]);
        else By[k] = pyl ? 0.5 * (b_y[k] + b_y[k + (size_t)nx * (ny - 1)]) : b_y[k];
    }
}

__global__ void k_get_cons(const double *rho, const double *vx, const double *vy,
                           const double *P, const double *Bx, const double *By, double g,
                           double vol, size_t n, double *M, double *Mx, double *My,
                           double *E) {
    GRID_LOOP(k, n) {
        prim_to_cons(rho[k], vx[k], vy[k], P[k], Bx[k], By[k], g, vol, M[k], Mx[k], My[k], E[k]);
    }
}

__global__ void k_get_prim(const double *M, const double *Mx, const double *My, const double *E,
                           const double *Bx, const double *By, double g, double vol, size_t n,
                           double *rho, double *vx, double *vy, double *P) {
    GRID_LOOP(k, n) {
        cons_to_prim(M[k], Mx[k], My[k], E[k], Bx[k], By[k], g, vol, rho[k], vx[k], vy[k], P[k]);
    }
}

__global__ void k_floors(double *M, double *Mx, double *My, double *E, const double *Bx,
                         const double *By, double vol, double g, size_t n) {
    GRID_LOOP(k, n) {
        double m = M[k], mx = Mx[k], my = My[k], e = E[k];
        conserved_floors(m, mx, my, e, Bx[k], By[k], vol, g);
        M[k] = m;
        Mx[k] = mx;
        My[k] = my;
        E[k] = e;
    }
}

// ---- host helpers -------------------------------------------------------------

struct DevBuf {
    double *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// One device arena for a drop-in call: `count` arrays of `n` doubles.
struct Arena {
    double *base = nullptr;
    size_t n = 0;
    ~Arena() {
        if (base) (void)hipFree(base);
    }
    int alloc(size_t count, size_t n_) {
        n = n_;
        hipError_t e = hipMalloc(&base, count * n_ * sizeof(double) + 16);
        if (e != hipSuccess) {
            set_error("hipMalloc of %zu bytes failed", count * n_ * sizeof(double));
            return CHIMERA_ENOMEM;
        }
        return CHIMERA_OK;
    }
    double *operator[](size_t i) const { return base + i * n; }
};

int h2d(double *d, const double *h, size_t n) {
    CH_REQUIRE(h, "NULL input array");
    CH_HIP(hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice));
    return CHIMERA_OK;
}
int d2h(double *h, const double *d, size_t n) {
    CH_REQUIRE(h, "NULL output array");
    CH_HIP(hipMemcpy(h, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return CHIMERA_OK;
}
int finish(const char *what) {
    CH_TRY(launch_check(what));
    CH_HIP(hipDeviceSynchronize());
    return CHIMERA_OK;
}

int check_grid(int nx, int ny) {
    CH_REQUIRE(nx >= 2 && ny >= 2, "grid must be at least 2 x 2 (got %d x %d)", nx, ny);
    return CHIMERA_OK;
}

int check_flags(const chimera_bc_t *bc) {
    CH_REQUIRE(bc, "bc is NULL");
    CH_REQUIRE(bc_flag_ok(bc->xlo) && bc_flag_ok(bc->xhi) && bc_flag_ok(bc->ylo) &&
                   bc_flag_ok(bc->yhi),
               "boundary flags must be 1..4");
    return CHIMERA_OK;
}

int launch_fluxes(int solver, const FluxArgs &a, size_t n, double gamma, hipStream_t s) {
    switch (solver) {
    case RUSANOV: hipLaunchKernelGGL(k_fluxes<RUSANOV>, dim3(blocks_for(n)), dim3(BS), 0, s, a, n, gamma); break;
    case HLLE: hipLaunchKernelGGL(k_fluxes<HLLE>, dim3(blocks_for(n)), dim3(BS), 0, s, a, n, gamma); break;
    case HLLD: hipLaunchKernelGGL(k_fluxes<HLLD>, dim3(blocks_for(n)), dim3(BS), 0, s, a, n, gamma); break;
    default: set_error("riemann_solver must be 1..3"); return CHIMERA_EINVAL;
    }
    return launch_check("k_fluxes");
}

// compute_stage_fluxes (src/main.f90:346-403) as the reference's routine sequence on
// device buffers: ghost fill, limiter (or gradients), reconstruction x6, thermal
// check, x and rotated y fluxes.
int dev_stage_fluxes(const double *prims, double *flux, uint32_t *mask, double dx, double gamma,
                     int nx, int ny, const chimera_bc_t *bc, const chimera_scheme_t *sch,
                     hipStream_t s) {
    size_t n = (size_t)nx * ny, np = (size_t)(nx + 2) * (ny + 2);
    Arena pad, tmp;
    CH_TRY(pad.alloc(1, np));
    CH_TRY(tmp.alloc(26, n));  // 2 slopes + 24 faces
    double *gx = tmp[0], *gy = tmp[1];
    auto face = [&](int f, int side) { return tmp[2 + 4 * f + side]; };
    if (mask) CH_HIP(hipMemsetAsync(mask, 0, n * sizeof(uint32_t), s));
    for (int f = 0; f < 6; ++f) {
        const double *pf = prims + (size_t)f * n;
        hipLaunchKernelGGL(k_fill_ghost, dim3(blocks_for(np)), dim3(BS), 0, s, pf, pad[0],
                           (const double *)nullptr, nx, ny, bc->xlo, bc->xhi, bc->ylo, bc->yhi);
        if (!sch->use_slope_limiting)
            hipLaunchKernelGGL(k_gradients, dim3(blocks_for(n)), dim3(BS), 0, s, pad[0], dx, nx, ny,
                               gx, gy);
        else if (sch->slope_limiter == LIM_MC)
            hipLaunchKernelGGL(k_limiter<LIM_MC>, dim3(blocks_for(n)), dim3(BS), 0, s, pad[0], dx,
                               nx, ny, gx, gy);
        else
            hipLaunchKernelGGL(k_limiter<LIM_VAN_LEER>, dim3(blocks_for(n)), dim3(BS), 0, s, pad[0],
                               dx, nx, ny, gx, gy);
        hipLaunchKernelGGL(k_reconstruction, dim3(blocks_for(n)), dim3(BS), 0, s, pf, gx, gy, dx,
                           nx, ny, face(f, 0), face(f, 1), face(f, 2), face(f, 3),
                           (f == 0 || f == 3) ? 1 : 0, bc->xlo, bc->xhi, bc->ylo, bc->yhi,
                           sch->fallback_2_mood, mask, f);
        CH_TRY(launch_check("stage reconstruction"));
    }
    ThermalArgs ta{};
    const int grp[6] = {3, 4, 5, 0, 1, 2};
    for (int g = 0; g < 6; ++g)
        for (int side = 0; side < 4; ++side) ta.face[4 * g + side] = face(grp[g], side);
    for (int f = 0; f < 6; ++f) ta.cell[f] = prims + (size_t)f * n;
    hipLaunchKernelGGL(k_thermal, dim3(blocks_for(n)), dim3(BS), 0, s, ta, n, mask);
    FluxArgs ax{{face(0, 0), face(0, 1), face(1, 0), face(1, 1), face(2, 0), face(2, 1),
                 face(3, 0), face(3, 1), face(4, 0), face(4, 1), face(5, 0), face(5, 1)},
                {flux + 0 * n, flux + 2 * n, flux + 4 * n, flux + 6 * n, flux + 8 * n}};
    CH_TRY(launch_fluxes(sch->riemann_solver, ax, n, gamma, s));
    FluxArgs ay{{face(0, 2), face(0, 3), face(2, 2), face(2, 3), face(1, 2), face(1, 3),
                 face(3, 2), face(3, 3), face(5, 2), face(5, 3), face(4, 2), face(4, 3)},
                {flux + 1 * n, flux + 5 * n, flux + 3 * n, flux + 7 * n, flux + 9 * n}};
    CH_TRY(launch_fluxes(sch->riemann_solver, ay, n, gamma, s));
    CH_HIP(hipStreamSynchronize(s));  // the arenas are freed on return
    return CHIMERA_OK;
}

}  // namespace
}  // namespace chimera

using namespace chimera;

extern "C" int chimera_fill_ghost_cells(const double *f, double *f_pad, const double *f_ambient,
                                        int nx, int ny, const chimera_bc_t *bc) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    size_t n = (size_t)nx * ny, np = (size_t)(nx + 2) * (ny + 2);
    Arena a;
    CH_TRY(a.alloc(3, np));
    CH_TRY(h2d(a[0], f, n));
    if (f_ambient) CH_TRY(h2d(a[2], f_ambient, n));
    hipLaunchKernelGGL(k_fill_ghost, dim3(blocks_for(np)), dim3(BS), 0, 0, a[0], a[1],
                       f_ambient ? (const double *)a[2] : nullptr, nx, ny, bc->xlo, bc->xhi,
                       bc->ylo, bc->yhi);
    CH_TRY(finish("k_fill_ghost"));
    return d2h(f_pad, a[1], np);
}

extern "C" int chimera_compute_gradients(const double *f_pad, double dx, int nx, int ny,
                                         double *f_dx, double *f_dy) {
    CH_TRY(check_grid(nx, ny));
    size_t n = (size_t)nx * ny, np = (size_t)(nx + 2) * (ny + 2);
    Arena a;
    CH_TRY(a.alloc(3, np));
    CH_TRY(h2d(a[0], f_pad, np));
    hipLaunchKernelGGL(k_gradients, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], dx, nx, ny, a[1],
                       a[2]);
    CH_TRY(finish("k_gradients"));
    CH_TRY(d2h(f_dx, a[1], n));
    return d2h(f_dy, a[2], n);
}

extern "C" int chimera_apply_slope_limiter(const double *f_pad, double dx, int nx, int ny,
                                           double *f_dx, double *f_dy, int slope_limiter) {
    CH_TRY(check_grid(nx, ny));
    CH_REQUIRE(slope_limiter == 1 || slope_limiter == 2, "slope_limiter must be 1 or 2");
    size_t n = (size_t)nx * ny, np = (size_t)(nx + 2) * (ny + 2);
    Arena a;
    CH_TRY(a.alloc(3, np));
    CH_TRY(h2d(a[0], f_pad, np));
    if (slope_limiter == LIM_MC)
        hipLaunchKernelGGL(k_limiter<LIM_MC>, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], dx, nx, ny,
                           a[1], a[2]);
    else
        hipLaunchKernelGGL(k_limiter<LIM_VAN_LEER>, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], dx,
                           nx, ny, a[1], a[2]);
    CH_TRY(finish("k_limiter"));
    CH_TRY(d2h(f_dx, a[1], n));
    return d2h(f_dy, a[2], n);
}

extern "C" int chimera_reconstruction(const double *f, const double *f_dx, const double *f_dy,
                                      double dx, int nx, int ny, double *f_XL, double *f_XR,
                                      double *f_YL, double *f_YR, int rho_or_p,
                                      const chimera_bc_t *bc, int fallback_2_mood) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(7, n));
    CH_TRY(h2d(a[0], f, n));
    CH_TRY(h2d(a[1], f_dx, n));
    CH_TRY(h2d(a[2], f_dy, n));
    hipLaunchKernelGGL(k_reconstruction, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], a[2], dx,
                       nx, ny, a[3], a[4], a[5], a[6], rho_or_p ? 1 : 0, bc->xlo, bc->xhi, bc->ylo,
                       bc->yhi, fallback_2_mood ? 1 : 0, (uint32_t *)nullptr, 0);
    CH_TRY(finish("k_reconstruction"));
    CH_TRY(d2h(f_XL, a[3], n));
    CH_TRY(d2h(f_XR, a[4], n));
    CH_TRY(d2h(f_YL, a[5], n));
    return d2h(f_YR, a[6], n);
}

extern "C" int chimera_thermal_pressure_check(
    int nx, int ny, double *P_XL, double *P_XR, double *P_YL, double *P_YR, double *Bx_XL,
    double *Bx_XR, double *Bx_YL, double *Bx_YR, double *By_XL, double *By_XR, double *By_YL,
    double *By_YR, double *rho_XL, double *rho_XR, double *rho_YL, double *rho_YR, double *vx_XL,
    double *vx_XR, double *vx_YL, double *vx_YR, double *vy_XL, double *vy_XR, double *vy_YL,
    double *vy_YR, const double *rho_0, const double *vx_0, const double *vy_0, const double *P_0,
    const double *Bx_0, const double *By_0) {
    CH_TRY(check_grid(nx, ny));
    size_t n = (size_t)nx * ny;
    double *hf[24] = {P_XL,   P_XR,   P_YL,   P_YR,   Bx_XL, Bx_XR, Bx_YL, Bx_YR,
                      By_XL,  By_XR,  By_YL,  By_YR,  rho_XL, rho_XR, rho_YL, rho_YR,
                      vx_XL,  vx_XR,  vx_YL,  vx_YR,  vy_XL, vy_XR, vy_YL, vy_YR};
    const double *hc[6] = {rho_0, vx_0, vy_0, P_0, Bx_0, By_0};
    Arena a;
    CH_TRY(a.alloc(30, n));
    ThermalArgs ta{};
    for (int q = 0; q < 24; ++q) {
        CH_TRY(h2d(a[q], hf[q], n));
        ta.face[q] = a[q];
    }
    for (int q = 0; q < 6; ++q) {
        CH_TRY(h2d(a[24 + q], hc[q], n));
        ta.cell[q] = a[24 + q];
    }
    hipLaunchKernelGGL(k_thermal, dim3(blocks_for(n)), dim3(BS), 0, 0, ta, n, (uint32_t *)nullptr);
    CH_TRY(finish("k_thermal"));
    for (int q = 0; q < 24; ++q) CH_TRY(d2h(hf[q], a[q], n));
    return CHIMERA_OK;
}

extern "C" int chimera_compute_fluxes(const double *rho_L, const double *rho_R,
                                      const double *vx_L, const double *vx_R, const double *vy_L,
                                      const double *vy_R, const double *P_L, const double *P_R,
                                      const double *Bx_L, const double *Bx_R, const double *By_L,
                                      const double *By_R, double gamma, int nx, int ny,
                                      double *flux_Mass, double *flux_Momx, double *flux_Momy,
                                      double *flux_Energy, double *flux_By, int riemann_solver) {
    CH_REQUIRE(nx >= 1 && ny >= 1, "nx, ny must be >= 1");
    CH_REQUIRE(riemann_solver >= 1 && riemann_solver <= 3, "riemann_solver must be 1..3");
    size_t n = (size_t)nx * ny;
    const double *hin[12] = {rho_L, rho_R, vx_L, vx_R, vy_L, vy_R, P_L, P_R, Bx_L, Bx_R, By_L, By_R};
    double *hout[5] = {flux_Mass, flux_Momx, flux_Momy, flux_Energy, flux_By};
    Arena a;
    CH_TRY(a.alloc(17, n));
    FluxArgs fa{};
    for (int q = 0; q < 12; ++q) {
        CH_TRY(h2d(a[q], hin[q], n));
        fa.in[q] = a[q];
    }
    for (int q = 0; q < 5; ++q) fa.out[q] = a[12 + q];
    CH_TRY(launch_fluxes(riemann_solver, fa, n, gamma, 0));
    CH_TRY(finish("k_fluxes"));
    for (int q = 0; q < 5; ++q) CH_TRY(d2h(hout[q], a[12 + q], n));
    return CHIMERA_OK;
}

extern "C" int chimera_dev_compute_fluxes(const double *in12, double *out5, double gamma, int n,
                                          int riemann_solver, void *stream) {
    CH_REQUIRE(in12 && out5 && n >= 0, "bad arguments");
    FluxArgs fa{};
    for (int q = 0; q < 12; ++q) fa.in[q] = in12 + (size_t)q * n;
    for (int q = 0; q < 5; ++q) fa.out[q] = out5 + (size_t)q * n;
    if (n == 0) return CHIMERA_OK;
    return launch_fluxes(riemann_solver, fa, (size_t)n, gamma, (hipStream_t)stream);
}

extern "C" int chimera_update_conserved(double *F, const double *flux_F_X,
                                        const double *flux_F_Y, double dx, double dt, int nx,
                                        int ny, const chimera_bc_t *bc) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(3, n));
    CH_TRY(h2d(a[0], F, n));
    CH_TRY(h2d(a[1], flux_F_X, n));
    CH_TRY(h2d(a[2], flux_F_Y, n));
    hipLaunchKernelGGL(k_update, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], a[2], dx, dt, nx,
                       ny, bc->xlo == BC_PERIODIC, bc->ylo == BC_PERIODIC);
    CH_TRY(finish("k_update"));
    return d2h(F, a[0], n);
}

extern "C" int chimera_constrained_transport(double *bx, double *by, const double *flux_By_X,
                                             const double *flux_Bx_Y, double dx, double dt,
                                             int nx, int ny, const chimera_bc_t *bc) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(4, n));
    CH_TRY(h2d(a[0], bx, n));
    CH_TRY(h2d(a[1], by, n));
    CH_TRY(h2d(a[2], flux_By_X, n));
    CH_TRY(h2d(a[3], flux_Bx_Y, n));
    hipLaunchKernelGGL(k_ct, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], a[2], a[3], dx, dt, nx,
                       ny, bc->xlo == BC_PERIODIC, bc->xhi == BC_PERIODIC, bc->ylo == BC_PERIODIC,
                       bc->yhi == BC_PERIODIC);
    CH_TRY(finish("k_ct"));
    CH_TRY(d2h(bx, a[0], n));
    return d2h(by, a[1], n);
}

extern "C" int chimera_compute_curl_2d(const double *Az, double dx, int nx, int ny, double *bx,
                                       double *by, const chimera_bc_t *bc) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(3, n));
    CH_TRY(h2d(a[0], Az, n));
    hipLaunchKernelGGL(k_curl, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], dx, nx, ny, a[1], a[2],
                       bc->xlo == BC_PERIODIC, bc->ylo == BC_PERIODIC);
    CH_TRY(finish("k_curl"));
    CH_TRY(d2h(bx, a[1], n));
    return d2h(by, a[2], n);
}

extern "C" int chimera_compute_divB(const double *bx, const double *by, double dx, int nx,
                                    int ny, double *divB, const chimera_bc_t *bc) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(3, n));
    CH_TRY(h2d(a[0], bx, n));
    CH_TRY(h2d(a[1], by, n));
    hipLaunchKernelGGL(k_divb, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], dx, nx, ny, a[2],
                       bc->xlo == BC_PERIODIC, bc->ylo == BC_PERIODIC);
    CH_TRY(finish("k_divb"));
    return d2h(divB, a[2], n);
}

extern "C" int chimera_average_face_to_cell_B(const double *b_x, const double *b_y, int nx,
                                              int ny, double *Bx, double *By,
                                              const chimera_bc_t *bc) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(4, n));
    CH_TRY(h2d(a[0], b_x, n));
    CH_TRY(h2d(a[1], b_y, n));
    hipLaunchKernelGGL(k_avg_b, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], nx, ny, a[2], a[3],
                       bc->xlo == BC_PERIODIC, bc->ylo == BC_PERIODIC);
    CH_TRY(finish("k_avg_b"));
    CH_TRY(d2h(Bx, a[2], n));
    return d2h(By, a[3], n);
}

extern "C" int chimera_get_conserved(const double *rho, const double *vx, const double *vy,
                                     const double *P, const double *Bx, const double *By,
                                     double gamma, double vol, int nx, int ny, double *Mass,
                                     double *Momx, double *Momy, double *Energy) {
    CH_REQUIRE(nx >= 1 && ny >= 1, "nx, ny must be >= 1");
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(10, n));
    const double *in[6] = {rho, vx, vy, P, Bx, By};
    for (int q = 0; q < 6; ++q) CH_TRY(h2d(a[q], in[q], n));
    hipLaunchKernelGGL(k_get_cons, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], a[2], a[3], a[4],
                       a[5], gamma, vol, n, a[6], a[7], a[8], a[9]);
    CH_TRY(finish("k_get_cons"));
    double *out[4] = {Mass, Momx, Momy, Energy};
    for (int q = 0; q < 4; ++q) CH_TRY(d2h(out[q], a[6 + q], n));
    return CHIMERA_OK;
}

extern "C" int chimera_get_primitive(const double *Mass, const double *Momx, const double *Momy,
                                     const double *Energy, const double *Bx, const double *By,
                                     double gamma, double vol, int nx, int ny, double *rho,
                                     double *vx, double *vy, double *P) {
    CH_REQUIRE(nx >= 1 && ny >= 1, "nx, ny must be >= 1");
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(10, n));
    const double *in[6] = {Mass, Momx, Momy, Energy, Bx, By};
    for (int q = 0; q < 6; ++q) CH_TRY(h2d(a[q], in[q], n));
    hipLaunchKernelGGL(k_get_prim, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], a[2], a[3], a[4],
                       a[5], gamma, vol, n, a[6], a[7], a[8], a[9]);
    CH_TRY(finish("k_get_prim"));
    double *out[4] = {rho, vx, vy, P};
    for (int q = 0; q < 4; ++q) CH_TRY(d2h(out[q], a[6 + q], n));
    return CHIMERA_OK;
}

extern "C" int chimera_apply_conserved_floors(double *Mass, double *Momx, double *Momy,
                                              double *Energy, const double *Bx,
                                              const double *By, double vol, double gamma, int nx,
                                              int ny) {
    CH_REQUIRE(nx >= 1 && ny >= 1, "nx, ny must be >= 1");
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(6, n));
    double *io[4] = {Mass, Momx, Momy, Energy};
    for (int q = 0; q < 4; ++q) CH_TRY(h2d(a[q], io[q], n));
    CH_TRY(h2d(a[4], Bx, n));
    CH_TRY(h2d(a[5], By, n));
    hipLaunchKernelGGL(k_floors, dim3(blocks_for(n)), dim3(BS), 0, 0, a[0], a[1], a[2], a[3], a[4],
                       a[5], vol, gamma, n);
    CH_TRY(finish("k_floors"));
    for (int q = 0; q < 4; ++q) CH_TRY(d2h(io[q], a[q], n));
    return CHIMERA_OK;
}

extern "C" int chimera_dev_stage_fluxes(const double *prims, double *flux, uint32_t *mask,
                                        double dx, double gamma, int nx, int ny,
                                        const chimera_bc_t *bc, const chimera_scheme_t *scheme,
                                        void *stream) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    CH_TRY(check_scheme(scheme));
    CH_REQUIRE(prims && flux, "NULL argument");
    return dev_stage_fluxes(prims, flux, mask, dx, gamma, nx, ny, bc, scheme, (hipStream_t)stream);
}

extern "C" int chimera_stage_fluxes(const double *prims, double *flux, uint32_t *mask, double dx,
                                    double gamma, int nx, int ny, const chimera_bc_t *bc,
                                    const chimera_scheme_t *scheme) {
    CH_TRY(check_grid(nx, ny));
    CH_TRY(check_flags(bc));
    CH_TRY(check_scheme(scheme));
    size_t n = (size_t)nx * ny;
    Arena a;
    CH_TRY(a.alloc(16, n));
    CH_TRY(h2d(a[0], prims, 6 * n));
    uint32_t *dmask = nullptr;
    if (mask) CH_HIP(hipMalloc(&dmask, n * sizeof(uint32_t)));
    int rc = dev_stage_fluxes(a[0], a[6], dmask, dx, gamma, nx, ny, bc, scheme, 0);
    if (rc == CHIMERA_OK) rc = d2h(flux, a[6], 10 * n);
    if (rc == CHIMERA_OK && mask &&
        hipMemcpy(mask, dmask, n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("mask copy failed");
        rc = CHIMERA_EHIP;
    }
    if (dmask) (void)hipFree(dmask);
    return rc;
}
