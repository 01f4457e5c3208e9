// Initial conditions of the reference's problem types 1-4 on the device.
//
// initialize_problem (src/mhd_init.f90:259-288) picks gamma, the boundary flags and a
// set-up -- setup_OT_vortex (:296-311), setup_KH_instab (:319-352), setup_FL_advec
// (:360-391), setup_rotor (:399-450) -- on the grid of src/main.f90:33-39 (dx = 1/N,
// cell centres (i - 1/2) dx), and main converts the primitives with get_conserved
// (src/main.f90:49, src/mhd_change_states.f90:16-58).  One thread per cell writes
// the conserved state (Mass, Momx, Momy, Energy) and the face field (b_x, b_y).
//
// Every expression is the reference's, in its order, with IEEE division and square
// root, so the state has the bits of the host set-up.  The face field of the
// Orszag-Tang and field-loop problems is the backward-difference curl of a potential
// sampled at cell centres (compute_curl_2d, src/mhd_field_ops.f90:15-57, periodic) and
// the cell field its face average (average_face_to_cell_B, :120-156); a thread
// evaluates the potential at its cell and the three neighbours below and to the left
// instead of reading a potential array.  The Orszag-Tang sines and cosines depend on x
// or y alone: the host evaluates them once per column with the C library, as the
// reference's compiled code calls it, and the kernel reads them from a table.  The
// Kelvin-Helmholtz exp is per cell and runs on the device with the C library's bits
// (libm_exp.hpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "libm_exp.hpp"
#include "resident.hpp"

namespace chimera {

namespace {

// src/mhd_config.f90:54-56, 59-60
constexpr double PI = 3.1415926535897932;
constexpr double TWO_PI = 6.2831853071795865;
constexpr double FOUR_PI = 12.566370614359173;
constexpr double FLOOR = 1.0e-12;  // rho_floor, P_floor

enum { OT = 1, KH = 2, FL = 3, ROTOR = 4 };

struct SetupArgs {
    int n;
    double dx, inv_dx, vol, gamma;
    double c0, c1;      // OT: rho, P before the magnetic term; rotor: cell and face B_x
    const double *tab;  // OT: sin(2 pi x_i) [0, n), A_z's x term [n, 2n), its y term [2n, 3n)
    double *u[6];
};

__device__ __forceinline__ double centre(int i, double dx) { return ((double)i + 0.5) * dx; }

// A_z at the centre of cell (i, j): setup_OT_vortex :303-304 (from the table),
// setup_FL_advec :375-383.
template <int PROB>
__device__ __forceinline__ double potential(const SetupArgs &a, int i, int j) {
    if constexpr (PROB == OT) {
        return a.tab[a.n + i] + a.tab[2 * a.n + j];
    } else {
        const double X = centre(i, a.dx) - 0.5, Y = centre(j, a.dx) - 0.5;
        const double r = sqrt(X * X + Y * Y);
        return r < 0.3 ? 1.0e-3 * (0.3 - r) : 0.0;
    }
}

template <int PROB>
__global__ void __launch_bounds__(256) setup_kernel(SetupArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= a.n || j >= a.n) return;
    double rho, vx, vy, P, Bx, By, bx, by;
    if constexpr (PROB == OT || PROB == FL) {
        const int im = i == 0 ? a.n - 1 : i - 1, jm = j == 0 ? a.n - 1 : j - 1;
        const double A = potential<PROB>(a, i, j), Aw = potential<PROB>(a, im, j);
        const double As = potential<PROB>(a, i, jm), Asw = potential<PROB>(a, im, jm);
        // compute_curl_2d: b_x(i,j) = (A(i,j) - A(i,j-1))/dx, b_y(i,j) = -(A(i,j) - A(i-1,j))/dx
        bx = (A - As) * a.inv_dx;
        by = -((A - Aw) * a.inv_dx);
        const double bx_w = (Aw - Asw) * a.inv_dx;
        const double by_s = -((As - Asw) * a.inv_dx);
        Bx = 0.5 * (bx + bx_w);
        By = 0.5 * (by + by_s);
        if constexpr (PROB == OT) {
            rho = a.c0;
            vx = -a.tab[j];
            vy = a.tab[i];
            P = a.c1 + 0.5 * (Bx * Bx + By * By);
        } else {
            rho = 1.0;
            vx = -2.0;
            vy = -1.0;
            P = 1.0 + 0.5 * (Bx * Bx + By * By);
        }
    } else if constexpr (PROB == KH) {
        const double X = centre(i, a.dx), Y = centre(j, a.dx);
        const bool inner = fabs(Y - 0.5) <= 0.25;
        rho = inner ? 2.0 : 1.0;
        vx = inner ? 0.5 : -0.5;
        const double ex = X - 0.5, ey = Y - 0.25;
        const double t = -(ex * ex + ey * ey) / (2.0 * (0.05 * 0.05));  // |t| <= 162.5
        vy = 0.5 * exp_libm(t);
        P = 2.5;
        Bx = By = bx = by = 0.0;
    } else {  // rotor
        const double rx = centre(i, a.dx) - 0.5, ry = centre(j, a.dx) - 0.5;
        const double r = sqrt(rx * rx + ry * ry);
        constexpr double r0 = 0.1, r1 = 0.115, u0 = 2.0;
        rho = 1.0;
        vx = vy = 0.0;
        if (r <= r0) {
            rho = 10.0;
            vx = -u0 * (ry / r0);
            vy = u0 * (rx / r0);
        } else if (r < r1) {
            const double f = (r1 - r) / (r1 - r0);
            rho = 1.0 + 9.0 * f;
            vx = -f * u0 * (ry / r);
            vy = f * u0 * (rx / r);
        }
        Bx = a.c0;
        By = 0.0;
        P = 1.0 + 0.5 * (Bx * Bx + By * By);
        bx = a.c0;
        by = 0.0;
    }
    // get_conserved
    const double r = fmax(rho, FLOOR), p = fmax(P, FLOOR);
    const double M = r * a.vol;
    const double half_b2 = 0.5 * (Bx * Bx + By * By);
    const double E =
        ((p - half_b2) / (a.gamma - 1.0) + 0.5 * r * (vx * vx + vy * vy) + half_b2) * a.vol;
    const size_t k = (size_t)a.n * j + i;
    a.u[0][k] = M;
    a.u[1][k] = M * vx;
    a.u[2][k] = M * vy;
    a.u[3][k] = E;
    a.u[4][k] = bx;
    a.u[5][k] = by;
}

}  // namespace

int problem_info(int problem, int n, double *dx, double *gamma, chimera_bc_t *bc) {
    CH_REQUIRE(problem >= 1 && problem <= 5, "problem_type %d: the reference defines 1-5",
               problem);
    CH_REQUIRE(problem != 5, "problem_type 5 (random fields) is set up by the GRF generator, "
                             "not by the closed-form set-ups");
    CH_REQUIRE(n >= 4 && (long long)n * n <= MAX_GRID_CELLS,
               "N = %d: need 4 <= N and N^2 <= %lld", n, MAX_GRID_CELLS);
    if (dx) *dx = 1.0 / n;
    if (gamma) *gamma = (problem == OT || problem == FL) ? 5.0 / 3.0 : 1.4;
    if (bc) {
        const int f = problem == ROTOR ? CHIMERA_BC_OUTFLOW : CHIMERA_BC_PERIODIC;
        *bc = chimera_bc_t{f, f, f, f};
    }
    return CHIMERA_OK;
}

int launch_setup_problem(int problem, int n, double *const *u, hipStream_t stream) {
    double dx = 0.0, gamma = 0.0;
    CH_TRY(problem_info(problem, n, &dx, &gamma, nullptr));
    for (int q = 0; q < 6; ++q) CH_REQUIRE(u[q], "state[%d] is NULL", q);
    SetupArgs a{};
    a.n = n;
    a.dx = dx;
    a.inv_dx = 1.0 / dx;
    a.vol = dx * dx;
    a.gamma = gamma;
    for (int q = 0; q < 6; ++q) a.u[q] = u[q];
    double *tab = nullptr;
    if (problem == OT) {
        a.c0 = 25.0 / (36.0 * PI);
        a.c1 = 5.0 / (12.0 * PI);
        std::vector<double> h(3 * (size_t)n);
        const double sx = FOUR_PI * std::sqrt(FOUR_PI), sy = TWO_PI * std::sqrt(FOUR_PI);
        for (int i = 0; i < n; ++i) {
            const double x = ((double)i + 0.5) * dx;
            h[i] = std::sin(TWO_PI * x);
            h[n + i] = std::cos(FOUR_PI * x) / sx;
            h[2 * n + i] = std::cos(TWO_PI * x) / sy;
        }
        CH_HIP(hipMalloc(&tab, h.size() * sizeof(double)));
        hipError_t e = hipMemcpyAsync(tab, h.data(), h.size() * sizeof(double),
                                      hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);  // h is freed on return
        if (e != hipSuccess) {
            (void)hipFree(tab);
            set_error("OT table upload: %s", hipGetErrorString(e));
            return CHIMERA_EHIP;
        }
        a.tab = tab;
    } else if (problem == ROTOR) {
        a.c0 = 5.0 / std::sqrt(4.0 * PI);
    }
    const dim3 block(64, 4), grid((n + 63) / 64, (n + 3) / 4);
    switch (problem) {
        case OT: setup_kernel<OT><<<grid, block, 0, stream>>>(a); break;
        case KH: setup_kernel<KH><<<grid, block, 0, stream>>>(a); break;
        case FL: setup_kernel<FL><<<grid, block, 0, stream>>>(a); break;
        default: setup_kernel<ROTOR><<<grid, block, 0, stream>>>(a); break;
    }
    int rc = launch_check("setup_kernel");
    hipError_t e = hipStreamSynchronize(stream);
    if (tab) (void)hipFree(tab);
    CH_TRY(rc);
    CH_HIP(e);
    return CHIMERA_OK;
}

}  // namespace chimera

extern "C" int chimera_problem_info(int problem_type, int n, double *dx, double *gamma,
                                    chimera_bc_t *bc) {
    return chimera::problem_info(problem_type, n, dx, gamma, bc);
}

extern "C" int chimera_setup_problem(int problem_type, int n, double *const *state,
                                     void *stream) {
    CH_REQUIRE(state, "NULL state");
    return chimera::launch_setup_problem(problem_type, n, state, (hipStream_t)stream);
}
