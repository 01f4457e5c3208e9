/*
 * chimera_hip.h -- C ABI of libchimera_hip.so, the MI355X (gfx950) implementation of
 * CHIMERA's SSP-RK3 stage update (tjnb0/CHIMERA).
 *
 * Arrays are FP64 in the reference's Fortran (nx, ny) column-major layout, which is
 * C [ny][nx]: element (i, j) at i + nx*j.  Boundary flags use the reference's
 * constants (src/mhd_config.f90:40-43): 1 periodic, 2 fixed, 3 outflow, 4 inflow;
 * the main loop never supplies an ambient state, so 2/3/4 all act as zero-gradient
 * (src/mhd_bc.f90:62-67, 76-81, 92-97, 106-111, src/main.f90:349-354).  Every
 * function returns 0 on success and a negative chimera_status on failure;
 * chimera_last_error() describes the last failure on the calling thread.
 *
 * Layers:
 *   1. Drop-ins: one function per reference routine, same argument meaning and order
 *      as the Fortran, plus the BC flags / scheme choices the Fortran reads from
 *      module mhd_config.  Host pointers; data is copied to the GPU and back per
 *      call.  (The *_dev variants take device pointers and a HIP stream.)
 *   2. Resident single-grid API (chimera_ctx_*): state lives in HBM, whole SSP-RK3
 *      steps run as fused LDS-tiled stage kernels, dt is reduced on the device.
 *   3. Batched ensemble (chimera_batch_*) and adjoint (chimera_adj_*).
 *   4. An allocation-free step (chimera_step_noalloc).
 */
#ifndef CHIMERA_HIP_H
#define CHIMERA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum chimera_status {
    CHIMERA_OK = 0,
    CHIMERA_EINVAL = -1,      /* bad argument (sizes, flags, null pointer) */
    CHIMERA_EHIP = -2,        /* HIP runtime error */
    CHIMERA_ENOMEM = -3,      /* device allocation failed */
    CHIMERA_EUNSUPPORTED = -4 /* valid in the reference, not supported here */
} chimera_status;

enum { CHIMERA_BC_PERIODIC = 1, CHIMERA_BC_FIXED = 2, CHIMERA_BC_OUTFLOW = 3, CHIMERA_BC_INFLOW = 4 };
enum { CHIMERA_LIMITER_VAN_LEER = 1, CHIMERA_LIMITER_MC = 2 };
enum { CHIMERA_RUSANOV = 1, CHIMERA_HLLE = 2, CHIMERA_HLLD = 3 };
/* dt policies of the resident loop (src/main.f90:74-83) */
enum {
    CHIMERA_DT_REFERENCE = 0,  /* 0.3*min(dx/(cf+|v|)), clamp to next output time, 1e-12 floor */
    CHIMERA_DT_CFL = 1,        /* as 0 without the output-time clamp */
    CHIMERA_DT_PRESCRIBED = 2  /* dt sequence supplied by the caller */
};

/* Per-side boundary flags: BC_xlo, BC_xhi, BC_ylo, BC_yhi of src/mhd_config.f90:48-49. */
typedef struct chimera_bc {
    int xlo, xhi, ylo, yhi;
} chimera_bc_t;

/* Scheme choices that are compile-time parameters in src/mhd_config.f90:27-37. */
typedef struct chimera_scheme {
    int riemann_solver;     /* CHIMERA_RUSANOV / _HLLE / _HLLD */
    int slope_limiter;      /* CHIMERA_LIMITER_VAN_LEER / _MC */
    int use_slope_limiting; /* 0: central gradients (compute_gradients) */
    int fallback_2_mood;    /* 0: no MOOD bounds check */
} chimera_scheme_t;

const char *chimera_last_error(void);
int chimera_device_count(int *count);

/* ------------------------------------------------------------------------------
 * 1. Drop-ins (host pointers).  Each replaces the Fortran routine named.
 * ---------------------------------------------------------------------------- */

/* mhd_bc::fill_ghost_cells, src/mhd_bc.f90:36-141.  f_pad is (nx+2) x (ny+2);
 * f_ambient may be NULL (zero-gradient for fixed/inflow, as main.f90 uses it).
 * Unlike the reference this also accepts nx != ny. */
int chimera_fill_ghost_cells(const double *f, double *f_pad, const double *f_ambient, int nx,
                             int ny, const chimera_bc_t *bc);

/* mhd_derivatives::compute_gradients, src/mhd_derivatives.f90:15-47 */
int chimera_compute_gradients(const double *f_pad, double dx, int nx, int ny, double *f_dx,
                              double *f_dy);

/* mhd_derivatives::apply_slope_limiter, src/mhd_derivatives.f90:50-127 */
int chimera_apply_slope_limiter(const double *f_pad, double dx, int nx, int ny, double *f_dx,
                                double *f_dy, int slope_limiter);

/* mhd_derivatives::reconstruction, src/mhd_derivatives.f90:130-229 */
int chimera_reconstruction(const double *f, const double *f_dx, const double *f_dy, double dx,
                           int nx, int ny, double *f_XL, double *f_XR, double *f_YL,
                           double *f_YR, int rho_or_p, const chimera_bc_t *bc,
                           int fallback_2_mood);

/* mhd_derivatives::thermal_pressure_check, src/mhd_derivatives.f90:232-293 */
int chimera_thermal_pressure_check(int nx, int ny, double *P_XL, double *P_XR, double *P_YL,
                                   double *P_YR, double *Bx_XL, double *Bx_XR, double *Bx_YL,
                                   double *Bx_YR, double *By_XL, double *By_XR, double *By_YL,
                                   double *By_YR, double *rho_XL, double *rho_XR,
                                   double *rho_YL, double *rho_YR, double *vx_XL,
                                   double *vx_XR, double *vx_YL, double *vx_YR, double *vy_XL,
                                   double *vy_XR, double *vy_YL, double *vy_YR,
                                   const double *rho_0, const double *vx_0, const double *vy_0,
                                   const double *P_0, const double *Bx_0, const double *By_0);

/* mhd_flux::compute_fluxes, src/mhd_flux.f90:139-191 (rusanov 197-288, hlle 291-405,
 * hlld 408-779). */
int chimera_compute_fluxes(const double *rho_L, const double *rho_R, const double *vx_L,
                           const double *vx_R, const double *vy_L, const double *vy_R,
                           const double *P_L, const double *P_R, const double *Bx_L,
                           const double *Bx_R, const double *By_L, const double *By_R,
                           double gamma, int nx, int ny, double *flux_Mass, double *flux_Momx,
                           double *flux_Momy, double *flux_Energy, double *flux_By,
                           int riemann_solver);

/* mhd_flux::update_conserved, src/mhd_flux.f90:20-76 */
int chimera_update_conserved(double *F, const double *flux_F_X, const double *flux_F_Y,
                             double dx, double dt, int nx, int ny, const chimera_bc_t *bc);

/* mhd_flux::constrained_transport, src/mhd_flux.f90:79-136 */
int chimera_constrained_transport(double *bx, double *by, const double *flux_By_X,
                                  const double *flux_Bx_Y, double dx, double dt, int nx, int ny,
                                  const chimera_bc_t *bc);

/* mhd_field_ops::compute_curl_2d, src/mhd_field_ops.f90:15-57 */
int chimera_compute_curl_2d(const double *Az, double dx, int nx, int ny, double *bx,
                            double *by, const chimera_bc_t *bc);

/* mhd_field_ops::compute_divB, src/mhd_field_ops.f90:60-117 */
int chimera_compute_divB(const double *bx, const double *by, double dx, int nx, int ny,
                         double *divB, const chimera_bc_t *bc);

/* mhd_field_ops::average_face_to_cell_B, src/mhd_field_ops.f90:120-156 */
int chimera_average_face_to_cell_B(const double *b_x, const double *b_y, int nx, int ny,
                                   double *Bx, double *By, const chimera_bc_t *bc);

/* mhd_change_states::get_conserved, src/mhd_change_states.f90:16-58 */
int chimera_get_conserved(const double *rho, const double *vx, const double *vy,
                          const double *P, const double *Bx, const double *By, double gamma,
                          double vol, int nx, int ny, double *Mass, double *Momx,
                          double *Momy, double *Energy);

/* mhd_change_states::get_primitive, src/mhd_change_states.f90:61-113 */
int chimera_get_primitive(const double *Mass, const double *Momx, const double *Momy,
                          const double *Energy, const double *Bx, const double *By,
                          double gamma, double vol, int nx, int ny, double *rho, double *vx,
                          double *vy, double *P);

/* mhd_change_states::apply_conserved_floors, src/mhd_change_states.f90:116-166 */
int chimera_apply_conserved_floors(double *Mass, double *Momx, double *Momy, double *Energy,
                                   const double *Bx, const double *By, double vol,
                                   double gamma, int nx, int ny);

/* main.f90:346-403 compute_stage_fluxes from the stage primitives
 * prims = {rho, vx, vy, P, Bx, By} (6 arrays of nx*ny, consecutive) into
 * flux = {Mass_X, Mass_Y, Momx_X, Momx_Y, Momy_X, Momy_Y, Energy_X, Energy_Y, By_X,
 * Bx_Y} (10 consecutive arrays).  mask (optional) gets the MOOD / thermal-check bits
 * (bit 4*f+s: MOOD reset face s in {XL,XR,YL,YR} of field f in {rho,vx,vy,P,Bx,By};
 * bit 24+s: thermal check reset face s). */
int chimera_stage_fluxes(const double *prims, double *flux, uint32_t *mask, double dx,
                         double gamma, int nx, int ny, const chimera_bc_t *bc,
                         const chimera_scheme_t *scheme);

/* Device-pointer forms of the drop-ins used by the fused path's tests and the
 * Python binding (same semantics, no copies, asynchronous on `stream`). */
int chimera_dev_stage_fluxes(const double *prims, double *flux, uint32_t *mask, double dx,
                             double gamma, int nx, int ny, const chimera_bc_t *bc,
                             const chimera_scheme_t *scheme, void *stream);
int chimera_dev_compute_fluxes(const double *in12, double *out5, double gamma, int n,
                               int riemann_solver, void *stream);

/* ------------------------------------------------------------------------------
 * 2. Resident single-grid API: the main.f90:65-303 loop with state in HBM.
 * ---------------------------------------------------------------------------- */
typedef struct chimera_ctx chimera_ctx_t;

/* Allocates three state copies (6 FP64 fields each) on the current device.
 * stream may be NULL (the context creates its own). */
int chimera_ctx_create(chimera_ctx_t **ctx, int nx, int ny, double dx, double gamma,
                       const chimera_bc_t *bc, const chimera_scheme_t *scheme, void *stream);
int chimera_ctx_destroy(chimera_ctx_t *ctx);
/* state = {Mass, Momx, Momy, Energy, b_x, b_y}; host or device pointers. */
int chimera_ctx_upload(chimera_ctx_t *ctx, const double *const *state);
int chimera_ctx_download(chimera_ctx_t *ctx, double *const *state);

/* Initial conditions of the reference's problem types 1-4 generated on the device:
 * initialize_problem (src/mhd_init.f90:259-288) with setup_OT_vortex (:296-311),
 * setup_KH_instab (:319-352), setup_FL_advec (:360-391) or setup_rotor (:399-450) on
 * the grid of src/main.f90:33-39, then get_conserved (src/main.f90:49); the state has
 * the bits of the reference program's.  Problem type 5 (random fields) is rejected
 * (CHIMERA_EINVAL): the GRF generator sets it up.
 * chimera_problem_info: the problem's dx (1/N), gamma and boundary flags (any output
 * may be NULL); host only.
 * chimera_setup_problem: into six N x N device fields state[0..5] = {Mass, Momx, Momy,
 * Energy, b_x, b_y}, on stream (NULL: the null stream); returns once they are written.
 * chimera_ctx_setup_problem: into the context's state, in place of an upload; the
 * context must have the problem's grid (N x N, dx, gamma, periodicity). */
int chimera_problem_info(int problem_type, int n, double *dx, double *gamma, chimera_bc_t *bc);
int chimera_setup_problem(int problem_type, int n, double *const *state, void *stream);
int chimera_ctx_setup_problem(chimera_ctx_t *ctx, int problem_type);
/* Run nsteps SSP-RK3 steps.  dt_in: nsteps values (host memory) for
 * CHIMERA_DT_PRESCRIBED, else NULL.  use_graph: replay a captured hipGraph per step. */
int chimera_ctx_steps(chimera_ctx_t *ctx, int nsteps, int dt_mode, const double *dt_in,
                      int use_graph);
/* Simulation time and the reference's outputCount (main.f90:61-62,279-292). */
int chimera_ctx_get_time(chimera_ctx_t *ctx, double *t, int *out_count);
int chimera_ctx_set_time(chimera_ctx_t *ctx, double t, int out_count);
/* dt of the last step taken (device value copied back). */
int chimera_ctx_last_dt(chimera_ctx_t *ctx, double *dt);
/* CFL dt of the current state under dt_mode (no step taken). */
int chimera_ctx_cfl_dt(chimera_ctx_t *ctx, int dt_mode, double *dt);
/* max|divB| and sum|divB| of the current state (compute_divB, field_ops:60-117). */
int chimera_ctx_divb(chimera_ctx_t *ctx, double *max_abs, double *sum_abs);
/* MOOD / thermal masks of stage 1 of the next step (layout as chimera_stage_fluxes). */
int chimera_ctx_stage1_mask(chimera_ctx_t *ctx, uint32_t *mask_host);
/* Synchronise the context's stream. */
int chimera_ctx_sync(chimera_ctx_t *ctx);
/* The stream the context launches on (hipStream_t). */
void *chimera_ctx_stream(chimera_ctx_t *ctx);
/* Timing of the fused stage kernel: after chimera_ctx_set_timing(ctx, 1) every stage
 * launch of chimera_ctx_steps with use_graph = 0 is bracketed by HIP events on the
 * context's stream; chimera_ctx_stage_kernel_ms returns the average duration (ms) and
 * the number of launches since set_timing (graph replays are not timed: launches stays
 * 0 and ms 0 when only graph steps ran). */
int chimera_ctx_set_timing(chimera_ctx_t *ctx, int enable);
int chimera_ctx_stage_kernel_ms(chimera_ctx_t *ctx, double *ms, int *launches);
/* Tile shape of the fused stage kernel: 0 = 32x8, 1 = 16x16, 2 = 64x4 cells (256
 * threads), 3 = 32x32 cells (1024 threads; Rusanov and HLLE only), 4 = 32x16 cells
 * (512 threads).  A new context uses 4 for Rusanov and HLLE on grids of at least one
 * 32x16 tile per CU, 0 otherwise.  Every shape gives the same bits. */
#define CHIMERA_CTX_NUM_SHAPES 5
int chimera_ctx_set_tile_shape(chimera_ctx_t *ctx, int shape);
int chimera_ctx_get_tile_shape(chimera_ctx_t *ctx, int *shape, int *tx, int *ty);
/* Time every tile shape on one stage of the current state (`reps` launches each),
 * keep the fastest; ms_out (optional) receives CHIMERA_CTX_NUM_SHAPES averages (-1 for a
 * shape the solver does not have). */
int chimera_ctx_autotune(chimera_ctx_t *ctx, int reps, int *best_shape, double *ms_out);

/* ------------------------------------------------------------------------------
 * 3a. Batched ensemble: many independent grids of uneven sizes advanced together.
 *     Each RK stage is ONE persistent launch over a tile table spanning all members;
 *     dt is per member (CFL-reduced on the device); steps can be replayed as hipGraphs.
 *     Replaces the sequential loop of single runs in scripts/run_mhd_MC.py:57-73.
 * ---------------------------------------------------------------------------- */
typedef struct chimera_batch chimera_batch_t;

/* nx, ny, dx, gamma: one entry per member; bc: one entry per member; scheme shared. */
int chimera_batch_create(chimera_batch_t **batch, int members, const int *nx, const int *ny,
                         const double *dx, const double *gamma, const chimera_bc_t *bc,
                         const chimera_scheme_t *scheme, void *stream);
int chimera_batch_destroy(chimera_batch_t *batch);
/* state = {Mass, Momx, Momy, Energy, b_x, b_y} of one member; host or device pointers. */
int chimera_batch_upload(chimera_batch_t *batch, int member, const double *const *state);
int chimera_batch_download(chimera_batch_t *batch, int member, double *const *state);
/* nsteps steps of every member (dt_mode CHIMERA_DT_REFERENCE or CHIMERA_DT_CFL). */
int chimera_batch_steps(chimera_batch_t *batch, int nsteps, int dt_mode, int use_graph);
int chimera_batch_member_time(chimera_batch_t *batch, int member, double *t, int *out_count,
                              double *last_dt);
int chimera_batch_sync(chimera_batch_t *batch);
void *chimera_batch_stream(chimera_batch_t *batch);
/* Total cells over all members. */
int chimera_batch_cells(chimera_batch_t *batch, long long *cells);
/* Tile shape (as chimera_ctx_set_tile_shape) and blocks per CU of the stage launch
 * (0..256; 0 = 16 x the blocks a CU holds at once).  Every block loops over the tile
 * table; counts above what a CU holds oversubscribe the launch, which keeps the blocks
 * sharing a CU out of phase.  Every setting gives the same bits. */
int chimera_batch_set_tiling(chimera_batch_t *batch, int shape, int blocks_per_cu);
int chimera_batch_get_tiling(chimera_batch_t *batch, int *shape, int *blocks_per_cu);
/* Time every (tile shape, blocks per CU) candidate on one stage (1 .. the occupancy
 * limit, then 2 .. 32 times it), keep the fastest. */
int chimera_batch_autotune(chimera_batch_t *batch, int reps, int *best_shape,
                           int *best_blocks_per_cu);
/* Average duration (ms) and count of the fused-stage launches of chimera_batch_steps
 * with use_graph = 0 since chimera_batch_set_timing(batch, 1) (HIP events on the
 * batch's stream; graph replays are not timed). */
int chimera_batch_set_timing(chimera_batch_t *batch, int enable);
int chimera_batch_stage_kernel_ms(chimera_batch_t *batch, double *ms, int *launches);

/* ------------------------------------------------------------------------------
 * 3b. Adjoint (discrete, frozen dt): reverse of one SSP-RK3 step.  Hand-written
 *     reverse kernels for reconstruction (limiter, MOOD, face floors, thermal check),
 *     Riemann fluxes, update, CT, outflow correction, floors and blends, taking the
 *     branches the forward pass took.  All pointers are device pointers (6 arrays of
 *     nx*ny: Mass, Momx, Momy, Energy, b_x, b_y).  No counterpart in the reference,
 *     which has no derivatives.
 * ---------------------------------------------------------------------------- */
typedef struct chimera_adj chimera_adj_t;

int chimera_adj_create(chimera_adj_t **adj, int nx, int ny, double dx, double gamma,
                       const chimera_bc_t *bc, const chimera_scheme_t *scheme, void *stream);
int chimera_adj_destroy(chimera_adj_t *adj);
/* state_out = one step of state_in with the given dt (the adjoint's forward kernels,
 * bit-identical to the fused path). */
int chimera_adj_forward_step(chimera_adj_t *adj, const double *const *state_in, double dt,
                             double *const *state_out);
/* lam_in = (dU^{n+1}/dU^n)^T lam_out at U^n = state_n with dt frozen. */
int chimera_adj_step_vjp(chimera_adj_t *adj, const double *const *state_n, double dt,
                         const double *const *lam_out, double *const *lam_in);
/* Reverse sweep over one checkpoint segment of nsteps steps: recompute the segment from
 * state0 = U^0 with the fused forward stage kernel and the dts dt_in[0..nsteps-1] (host
 * memory), keeping every stage input and every stage's flux-updated conserved state
 * (3 * nsteps + 1 states and 3 * nsteps four-field sets of device memory, allocated on
 * first use and kept), then lam <- (dU^{nsteps}/dU^0)^T lam in place (device arrays). */
int chimera_adj_segment_vjp(chimera_adj_t *adj, const double *const *state0, const double *dt_in,
                            int nsteps, double *const *lam);
/* The whole reverse sweep of an nsteps run checkpointed every seg_steps steps: states
 * holds 6 device pointers per checkpoint (checkpoint k = U^{k * seg_steps}, k = 0 ..
 * ceil(nsteps / seg_steps) - 1), then lam <- (dU^{nsteps}/dU^0)^T lam in place.  The
 * same bits as chimera_adj_segment_vjp over the segments from last to first; the
 * recompute of segment k-1 runs on a second stream while segment k is reversed (twice
 * the trajectory memory of one segment).  Replaces the per-segment loop of the
 * autograd backward; no counterpart in the reference (it has no derivatives). */
int chimera_adj_sweep(chimera_adj_t *adj, const double *const *states, const double *dt_in,
                      int nsteps, int seg_steps, double *const *lam);
/* Forward of the last checkpoint segment of a run the next chimera_adj_sweep reverses:
 * nsteps (<= seg_steps) steps from state0 (device) with the dts dt_in (host), keeping
 * the segment's trajectory where that sweep (a run of nseg segments checkpointed every
 * seg_steps steps) looks for it; the final state goes to state_out (device).  A sweep
 * whose last segment has the same state0 pointers, dts, seg_steps and nseg skips that
 * segment's recompute; any other trajectory use of the workspace drops the recording.
 * No counterpart in the reference. */
int chimera_adj_record(chimera_adj_t *adj, const double *const *state0, const double *dt_in,
                       int nsteps, int seg_steps, int nseg, double *const *state_out);
void *chimera_adj_stream(chimera_adj_t *adj);

/* Which reverse kernels the adjoint runs: 1 (default) the fused LDS-tiled pair that
 * recompute the forward stage in place of reading stored intermediates, 0 the
 * stored-intermediate kernels.  Both give the same bits; 0 is the cross-check.  No
 * counterpart in the reference (it has no derivatives). */
int chimera_adj_set_fused(chimera_adj_t *adj, int fused);
/* Tile shapes of the two fused reverse kernels: 0 = 16x16, 1 = 16x8, 2 = 8x8, 3 = 32x8
 * (-1 keeps the current one).  Every shape gives the same bits. */
int chimera_adj_set_tiles(chimera_adj_t *adj, int shape_a, int shape_b);

/* Riemann-solver VJPs of n x-direction faces (host arrays): in[12][n] holds the face
 * states in the order of chimera_compute_fluxes (rho_L, rho_R, vx_L, vx_R, vy_L, vy_R,
 * P_L, P_R, Bx_L, Bx_R, By_L, By_R), lam_flux[5][n] the adjoint of (mass, Momx, Momy,
 * energy, By) flux, out[12][n] receives the adjoint of each input.  method 0: the
 * hand-written reverse the adjoint uses, 1: forward-mode dual numbers (one pass per
 * input), the cross-check of the first.  No counterpart in the reference. */
int chimera_face_vjp(int solver, int method, int n, double gamma, const double *in,
                     const double *lam_flux, double *out);

/* ---- batched adjoint (ensembles of uneven grids) -------------------------------------
 * The adjoint of chimera_adj_* for many independent grids at once: every forward and
 * reverse stage is one persistent launch over a tile table spanning all members.  Each
 * member has its own size, dx, gamma, BCs and prescribed dt sequence; the scheme is
 * shared.  Per-member results are bit-identical to the single-grid adjoint.  State and
 * lam arguments are members x 6 device pointers (member m's field q at [6*m + q]); dt
 * arrays are members x nsteps host doubles (member m's steps at [m*nsteps ..]).  No
 * counterpart in the reference (it has no derivatives). */
typedef struct chimera_badj chimera_badj_t;
int chimera_badj_create(chimera_badj_t **badj, int members, const int *nx, const int *ny,
                        const double *dx, const double *gamma, const chimera_bc_t *bc,
                        const chimera_scheme_t *scheme, void *stream);
int chimera_badj_destroy(chimera_badj_t *badj);
/* nsteps steps of every member from state_in to state_out (may be the same arrays);
 * keeps no trajectory (two intermediate stage states per member). */
int chimera_badj_forward(chimera_badj_t *badj, const double *const *state_in, const double *dt_in,
                         int nsteps, double *const *state_out);
/* One checkpoint segment of every member: recompute from state0 keeping every stage
 * input and every stage's flux-updated conserved state (per member 3 * nsteps + 1
 * states and 3 * nsteps four-field sets, allocated on first use and kept), then
 * lam <- (dU^{nsteps}/dU^0)^T lam in place. */
int chimera_badj_segment_vjp(chimera_badj_t *badj, const double *const *state0,
                             const double *dt_in, int nsteps, double *const *lam);
/* The whole reverse sweep of an nsteps run checkpointed every seg_steps steps, for every
 * member: states holds nseg x members x 6 device pointers (checkpoint k of member m at
 * 6 * (k * members + m), k = 0 .. ceil(nsteps / seg_steps) - 1), dt_in members x nsteps
 * (host), lam members x 6 device pointers, updated in place.  The bits of
 * chimera_badj_segment_vjp over the segments from last to first, in one call.  No
 * counterpart in the reference (it has no derivatives). */
int chimera_badj_sweep(chimera_badj_t *badj, const double *const *states, const double *dt_in,
                       int nsteps, int seg_steps, double *const *lam);
/* Forward of the last checkpoint segment of a run for every member (nsteps <= seg_steps
 * steps from state0, members x 6 device pointers; dt_in members x nsteps, host), keeping
 * the trajectories the next chimera_badj_sweep of a run of nseg segments would
 * recompute for it; the final states go to state_out (members x 6).  That sweep skips
 * the recompute when its last segment has the same state0 pointers and dts; any other
 * trajectory use drops the recording.  No counterpart in the reference. */
int chimera_badj_record(chimera_badj_t *badj, const double *const *state0, const double *dt_in,
                        int nsteps, int seg_steps, int nseg, double *const *state_out);
/* Tile shapes of the two reverse passes (default 1) and of the forward stages
 * (default 0): 0 = 16x16, 1 = 32x8 (-1 keeps).  Every shape gives the same bits. */
int chimera_badj_set_tiles(chimera_badj_t *badj, int shape_a, int shape_b, int shape_f);
/* Tile autotuner: times one forward step of every member from state0 (members x 6
 * device pointers; dt_in: members x 1, host) per forward shape and one reverse step per
 * (A, B) shape pair, reps times each after one untimed run, and keeps the fastest.
 * Uses scratch arrays of its own.  best (may be NULL): the chosen {forward, A, B};
 * ms (may be NULL): mean ms of forward shapes 0, 1, then of the pairs (0,0), (0,1),
 * (1,0), (1,1).  No counterpart in the reference. */
int chimera_badj_autotune(chimera_badj_t *badj, const double *const *state0, const double *dt_in,
                          int reps, int *best, double *ms);
void *chimera_badj_stream(chimera_badj_t *badj);

/* ------------------------------------------------------------------------------
 * 4. Allocation-free step.  No counterpart in the reference: for callers that cannot
 *    allocate device memory (code under HIP graph capture, tight serving loops).
 * ---------------------------------------------------------------------------- */
/* Largest grid (nx * ny) of chimera_step_noalloc. */
#define CHIMERA_NOALLOC_MAX_CELLS (512 * 512)
/* state_out = one SSP-RK3 step of state_in with the given dt: the bits of
 * chimera_adj_forward_step (and of chimera_ctx_steps with that dt).  state_in /
 * state_out: 6 device arrays of nx * ny (Mass, Momx, Momy, Energy, b_x, b_y);
 * state_out may be state_in.  Asynchronous on `stream` (NULL: the null stream).  Calls
 * no hipMalloc / hipFree: the two intermediate stage states and the device dt live in
 * a static device array of the library (one per GPU, about 25 MB, held until the
 * process ends), so grids above CHIMERA_NOALLOC_MAX_CELLS cells are rejected with
 * CHIMERA_EINVAL, and calls that overlap on one GPU (other streams or threads) race
 * on that array. */
int chimera_step_noalloc(const double *const *state_in, double dt, double *const *state_out,
                         int nx, int ny, double dx, double gamma, const chimera_bc_t *bc,
                         const chimera_scheme_t *scheme, void *stream);

#ifdef __cplusplus
}
#endif
#endif
