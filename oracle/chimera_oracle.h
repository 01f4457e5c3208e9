/*
 * CPU oracle for the CHIMERA SSP-RK3 stage update -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference Fortran hot path (tjnb0/CHIMERA,
 * src/main.f90, src/mhd_*.f90).  Every function cites the reference lines it
 * follows and keeps the reference's floating-point operation order so that,
 * compiled with -ffp-contract=off, it reproduces the reference bit for bit.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / reported CPU baseline.  The
 * product path (chimera_amd) never calls it.
 *
 * Layout: every field is FP64, Fortran (nx, ny) column-major == C [ny][nx],
 * element (i, j) (0-based) at index i + nx*j.  Padded fields are
 * (nx+2) x (ny+2) with physical cell (i, j) at (i+1) + (nx+2)*(j+1).
 *
 * Boundary flags follow src/mhd_config.f90:40-43:
 *   1 = periodic, 2 = fixed, 3 = outflow, 4 = inflow.
 * bc[] is ordered {xlo, xhi, ylo, yhi}.
 */
#ifndef CHIMERA_ORACLE_H
#define CHIMERA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_PERIODIC = 1, ORC_BC_FIXED = 2, ORC_BC_OUTFLOW = 3, ORC_BC_INFLOW = 4 };
enum { ORC_LIM_VAN_LEER = 1, ORC_LIM_MC = 2 };
enum { ORC_RUSANOV = 1, ORC_HLLE = 2, ORC_HLLD = 3 };
enum { ORC_DT_REFERENCE = 0, ORC_DT_CFL = 1, ORC_DT_PRESCRIBED = 2 };

/* scheme[] = {riemann_solver, slope_limiter, use_slope_limiting, fallback_2_MOOD} */

void orc_fill_ghost_cells(int nx, int ny, const int *bc, const double *f,
                          double *f_pad, const double *f_ambient);
void orc_compute_gradients(int nx, int ny, const double *f_pad, double dx,
                           double *f_dx, double *f_dy);
void orc_apply_slope_limiter(int nx, int ny, int limiter, const double *f_pad,
                             double dx, double *f_dx, double *f_dy);
void orc_reconstruction(int nx, int ny, const int *bc, int mood, const double *f,
                        const double *f_dx, const double *f_dy, double dx,
                        double *XL, double *XR, double *YL, double *YR,
                        int rho_or_p, unsigned *mood_mask, int field_index);
/* faces[24] = {P_XL,P_XR,P_YL,P_YR, Bx_*, By_*, rho_*, vx_*, vy_*} (reference
 * argument order); cells[6] = {rho, vx, vy, P, Bx, By}. */
void orc_thermal_pressure_check(int n, double *const *faces, const double *const *cells,
                                unsigned *mask);
/* in[12] = {rho_L,rho_R,vx_L,vx_R,vy_L,vy_R,P_L,P_R,Bx_L,Bx_R,By_L,By_R};
 * out[5] = {Mass, Momx, Momy, Energy, By}. */
void orc_compute_fluxes(int n, int solver, double gamma, const double *const *in,
                        double *const *out);
void orc_update_conserved(int nx, int ny, const int *bc, double *F, const double *fX,
                          const double *fY, double dx, double dt);
void orc_constrained_transport(int nx, int ny, const int *bc, double *bx, double *by,
                               const double *fByX, const double *fBxY, double dx,
                               double dt);
void orc_compute_curl_2d(int nx, int ny, const int *bc, const double *Az, double dx,
                         double *bx, double *by);
void orc_compute_divB(int nx, int ny, const int *bc, const double *bx, const double *by,
                      double dx, double *divB);
void orc_average_face_to_cell_B(int nx, int ny, const int *bc, const double *b_x,
                                const double *b_y, double *Bx, double *By);
void orc_get_conserved(int n, const double *rho, const double *vx, const double *vy,
                       const double *P, const double *Bx, const double *By, double gamma,
                       double vol, double *Mass, double *Momx, double *Momy,
                       double *Energy);
void orc_get_primitive(int n, const double *Mass, const double *Momx, const double *Momy,
                       const double *Energy, const double *Bx, const double *By,
                       double gamma, double vol, double *rho, double *vx, double *vy,
                       double *P);
void orc_apply_conserved_floors(int n, double *Mass, double *Momx, double *Momy,
                                double *Energy, const double *Bx, const double *By,
                                double vol, double gamma);
/* max over cells of cf + |v| (main.f90:75-79); dt = 0.3 * dx / max. */
double orc_cfl_max_speed(int n, double gamma, const double *rho, const double *vx,
                         const double *vy, const double *P, const double *Bx,
                         const double *By);

/* L(u) of one stage (main.f90:346-403) from the stage primitives.
 * prims[6] = {rho, vx, vy, P, Bx, By}; flux[10] = {Mass_X, Mass_Y, Momx_X, Momx_Y,
 * Momy_X, Momy_Y, Energy_X, Energy_Y, By_X, Bx_Y}.  mask (optional, nx*ny) gets
 * bit 4*f+k when MOOD reset face k (XL,XR,YL,YR) of field f (rho,vx,vy,P,Bx,By),
 * and bit 24+k when the thermal-pressure check reset face k. */
int orc_stage_fluxes(int nx, int ny, const int *bc, const int *scheme, double gamma,
                     double dx, const double *const *prims, double *const *flux,
                     unsigned *mask);

/* Run nsteps SSP-RK3 steps (main.f90:65-303) in place on
 * state[6] = {Mass, Momx, Momy, Energy, b_x, b_y}.
 *   dt_mode 0: reference policy (CFL, clamp to next output time, 1e-12f floor)
 *           1: CFL with the 1e-12f floor, no output clamp
 *           2: dt_in[s] prescribed for step s
 *   t_io / out_count_io: simulation time and output counter (mode 0 clamp);
 *   t_end > 0 stops early once t >= t_end (the reference's `do while (t < tEnd)`).
 *   dt_out (optional) receives each step's dt.  divb_out (optional, 2 per step)
 *   receives {max|divB|, sum|divB|} after each step.
 * Returns the number of steps taken, or <0 on bad arguments. */
int orc_run(int nx, int ny, const int *bc, const int *scheme, double gamma, double dx,
            double *const *state, int nsteps, int dt_mode, const double *dt_in,
            double *dt_out, double *t_io, int *out_count_io, double t_end,
            double *divb_out, unsigned *mask_stage1);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
