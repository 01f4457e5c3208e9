// Launch helpers shared by the resident single-grid driver, the batched ensemble
// and the adjoint.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "stage_kernel.hpp"

namespace chimera {

// Tile shapes of the stage kernels (single grid and batched ensemble).  One thread per
// cell in the update phase.  The 256-cell shapes 32 x 8, 16 x 16 and 64 x 4 keep
// 41-49 KiB of LDS (32 x 8: primitives of the 36 x 12 halo tile, the four conserved
// fluxes of 33 x 8 x faces and 32 x 9 y faces, the transverse B flux of 33 x 10 x faces
// and 34 x 9 y faces), three blocks per CU.  32 x 32 is run by 1024 threads (Rusanov
// and HLLE; HLLD needs more than the 128 VGPRs a 1024-thread block leaves each lane):
// per cell 1.27 primitive conversions and 2.19 Riemann faces instead of 1.69 and 2.48
// for 32 x 8, in 144 KiB of LDS, one block (16 waves) per CU.  32 x 16 is run by 512
// threads (1.41 conversions, 2.29 faces); its 76 KiB of LDS let two blocks share a CU,
// so that one block's barrier waits overlap the other's work.
struct TileDims {
    int tx, ty;
};
constexpr int RES_NUM_SHAPES = 5;
constexpr TileDims RES_TILE_SHAPES[RES_NUM_SHAPES] = {
    {32, 8}, {16, 16}, {64, 4}, {32, 32}, {32, 16}};
// Largest grid (cells per field) of the fused kernels: they address a field with 32-bit
// byte offsets.
constexpr long long MAX_GRID_CELLS = 1ll << 29;
// block size of the adjoint's forward-stage kernels (256-cell tiles)
constexpr int STAGE_THREADS = 256;

// Per-grid step bookkeeping kept on the device so whole steps can be captured.
struct StepState {
    double t;
    double dt;
    unsigned long long smax_bits;  // max signal speed, order-preserving bits
    int out_count;
    int step;
};

struct StageLaunch {
    GridView g;
    int stage, use_lim, mood, solver, limiter, shape;
    int tiles_x, ntiles, grid;
};

struct CflArgs {
    const double *u[6];
    int nx, ny, px, py;
    double vol, gamma;
    unsigned long long *smax_bits;
};

int launch_stage(const StageLaunch &L, hipStream_t stream);
int launch_cfl(const CflArgs &a, hipStream_t stream);
int launch_dt_finalize(StepState *st, int mode, double dx, const double *dt_seq,
                       hipStream_t stream);
int launch_set_state(StepState *st, double t, int oc, int step, hipStream_t stream);
int launch_zero_u32(uint32_t *p, size_t n, hipStream_t stream);
int launch_divb(const double *bx, const double *by, int nx, int ny, int px, int py, double dx,
                unsigned long long *maxbits, double *sum, hipStream_t stream);

// Initial conditions of problem types 1-4 (setup.hip): dx, gamma and boundary flags of
// problem `problem` on N x N cells, and the set-up written into the six N x N device
// fields u[0..5] (returns once the fields are written).
int problem_info(int problem, int n, double *dx, double *gamma, chimera_bc_t *bc);
int launch_setup_problem(int problem, int n, double *const *u, hipStream_t stream);

}  // namespace chimera
