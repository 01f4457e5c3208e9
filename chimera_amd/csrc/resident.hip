// Resident single-grid SSP-RK3 driver (the src/main.f90:65-303 loop with the state
// in HBM) and the small device kernels around the fused stage kernel: CFL
// max-reduction, dt policy, step bookkeeping and the div B monitor.
#include <cstring>
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "resident.hpp"

namespace chimera {

thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_bc(const chimera_bc_t *bc, int *px, int *py) {
    CH_REQUIRE(bc != nullptr, "bc is NULL");
    CH_REQUIRE(bc_flag_ok(bc->xlo) && bc_flag_ok(bc->xhi) && bc_flag_ok(bc->ylo) &&
                   bc_flag_ok(bc->yhi),
               "boundary flags must be 1..4 (got %d %d %d %d)", bc->xlo, bc->xhi, bc->ylo,
               bc->yhi);
    bool pxl = bc->xlo == BC_PERIODIC, pxh = bc->xhi == BC_PERIODIC;
    bool pyl = bc->ylo == BC_PERIODIC, pyh = bc->yhi == BC_PERIODIC;
    if (pxl != pxh || pyl != pyh) {
        set_error("periodic boundary on one side of an axis only is not supported");
        return CHIMERA_EUNSUPPORTED;
    }
    if (px) *px = pxl;
    if (py) *py = pyl;
    return CHIMERA_OK;
}

int check_scheme(const chimera_scheme_t *s) {
    CH_REQUIRE(s != nullptr, "scheme is NULL");
    CH_REQUIRE(s->riemann_solver >= 1 && s->riemann_solver <= 3, "riemann_solver must be 1..3");
    CH_REQUIRE(s->slope_limiter == 1 || s->slope_limiter == 2, "slope_limiter must be 1 or 2");
    return CHIMERA_OK;
}

// ---------------------------------------------------------------------------------
// kernels

// 512-thread Rusanov / HLLE blocks: at most 128 VGPRs, so that two blocks (four waves
// per SIMD) share a CU.  256-thread HLLD blocks: at most 168 (three waves per SIMD, as
// their LDS allows); the compiler's own 170 left two (32 x 8 and 64 x 4 8-9% slower).
template <int SOLVER, int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX * TY)
__attribute__((amdgpu_waves_per_eu((TX * TY == 512 && SOLVER != HLLD) ? 4
                                   : ((TX * TY == 256 && SOLVER == HLLD) ? 3 : 1))))
void stage_kernel(GridView g, int stage, int use_lim, int mood, int tiles_x, int ntiles) {
    __shared__ double lds[TileShape<TX, TY>::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;  // nb is a multiple of 8
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    stage_tile<SOLVER, LIMITER, TX, TY>(g, stage, use_lim != 0, mood != 0, tile % tiles_x,
                                        tile / tiles_x, lds);
}

// Stage 1 with the MOOD / thermal-check bits of every owned face ORed into g.mask
// (chimera_ctx_stage1_mask); 32 x 8 tiles.
template <int SOLVER, int LIMITER>
__global__ __launch_bounds__(256) void stage_mask_kernel(GridView g, int use_lim, int mood,
                                                          int tiles_x, int ntiles) {
    __shared__ double lds[TileShape<32, 8>::LDS_DOUBLES];
    const int b = blockIdx.x, nb = gridDim.x;  // nb is a multiple of 8
    const int tile = (b % 8) * (nb / 8) + b / 8;
    if (tile >= ntiles) return;
    stage_tile<SOLVER, LIMITER, 32, 8, STORE_NONE, true>(g, 1, use_lim != 0, mood != 0,
                                                          tile % tiles_x, tile / tiles_x, lds);
}

using StageFn = void (*)(GridView, int, int, int, int, int);
using MaskFn = void (*)(GridView, int, int, int, int);

MaskFn mask_fn(int solver, int limiter) {
    static const MaskFn table[3][2] = {
        {stage_mask_kernel<RUSANOV, LIM_VAN_LEER>, stage_mask_kernel<RUSANOV, LIM_MC>},
        {stage_mask_kernel<HLLE, LIM_VAN_LEER>, stage_mask_kernel<HLLE, LIM_MC>},
        {stage_mask_kernel<HLLD, LIM_VAN_LEER>, stage_mask_kernel<HLLD, LIM_MC>}};
    return table[solver - 1][limiter - 1];
}

#define CH_SHAPES(S, L)                                                                 \
    {stage_kernel<S, L, 32, 8>, stage_kernel<S, L, 16, 16>, stage_kernel<S, L, 64, 4>,     \
     stage_kernel<S, L, 32, 32>, stage_kernel<S, L, 32, 16>}
#define CH_SHAPES_HLLD(S, L)                                                            \
    {stage_kernel<S, L, 32, 8>, stage_kernel<S, L, 16, 16>, stage_kernel<S, L, 64, 4>,     \
     nullptr, stage_kernel<S, L, 32, 16>}

// nullptr: shape not available for that solver
StageFn stage_fn(int solver, int limiter, int shape) {
    static const StageFn table[3][2][RES_NUM_SHAPES] = {
        {CH_SHAPES(RUSANOV, LIM_VAN_LEER), CH_SHAPES(RUSANOV, LIM_MC)},
        {CH_SHAPES(HLLE, LIM_VAN_LEER), CH_SHAPES(HLLE, LIM_MC)},
        {CH_SHAPES_HLLD(HLLD, LIM_VAN_LEER), CH_SHAPES_HLLD(HLLD, LIM_MC)}};
    return table[solver - 1][limiter - 1][shape];
}
#undef CH_SHAPES
#undef CH_SHAPES_HLLD

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// max over cells of cf + |v| of the primitives of state u (src/main.f90:69-79).  Blocks
// stride over rows and threads over a row's cells (coalesced, no index division); the
// max is exact, so the order of the reduction does not change dt.
__global__ __launch_bounds__(256) void cfl_reduce_kernel(CflArgs a) {
    __shared__ double part[4];
    const int nx = a.nx, ny = a.ny;
    double s = 0.0;
    for (int j = blockIdx.x; j < ny; j += gridDim.x) {
        const int jm = j > 0 ? j - 1 : (a.py ? ny - 1 : 0);
        const size_t row = (size_t)nx * j, rowm = (size_t)nx * jm;
        for (int i = threadIdx.x; i < nx; i += 256) {
            const int im = i > 0 ? i - 1 : (a.px ? nx - 1 : 0);
            const size_t idx = row + i;
            double Bx = 0.5 * (a.u[4][idx] + a.u[4][row + im]);
            double By = 0.5 * (a.u[5][idx] + a.u[5][rowm + i]);
            double rho, vx, vy, P;
            cons_to_prim(a.u[0][idx], a.u[1][idx], a.u[2][idx], a.u[3][idx], Bx, By, a.gamma,
                         a.vol, rho, vx, vy, P);
            s = fmax(s, cfl_speed(rho, vx, vy, P, Bx, By, a.gamma));
        }
    }
    s = wave_max(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(part[0], part[1]), fmax(part[2], part[3]));
        atomicMax(a.smax_bits, encode_nonneg(m));
    }
}

// dt of the step (src/main.f90:79-83), or the prescribed value, then the step's advance
// of t and of the output counter (src/main.f90:273-292) in the same thread: the stages
// read only dt, so t can move before they run.
__global__ void dt_finalize_kernel(StepState *st, int mode, double dx, const double *dt_seq) {
    double dt;
    if (mode == CHIMERA_DT_PRESCRIBED) {
        dt = dt_seq[st->step];
    } else {
        double smax = __longlong_as_double((long long)st->smax_bits);
        st->smax_bits = 0ull;
        dt = COURANT * (dx / smax);
        if (mode == CHIMERA_DT_REFERENCE) {
            double t_next = st->out_count * T_OUT;
            if (st->t + dt > t_next) dt = t_next - st->t;
        }
        dt = fmax(dt, (double)1e-12f);  // single-precision literal in the reference
    }
    st->dt = dt;
    const double t = st->t + dt;
    st->t = t;
    if (t >= st->out_count * T_OUT) st->out_count = st->out_count + 1;
    st->step = st->step + 1;
}

__global__ void set_state_kernel(StepState *st, double t, int out_count, int step) {
    st->t = t;
    st->out_count = out_count;
    st->step = step;
    st->smax_bits = 0ull;
}

__global__ void zero_u32_kernel(uint32_t *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// compute_divB (src/mhd_field_ops.f90:60-117) reduced to max|divB| and sum|divB|.
__global__ __launch_bounds__(256) void divb_reduce_kernel(const double *bx, const double *by,
                                                          int nx, int ny, int px, int py,
                                                          double dx, unsigned long long *maxbits,
                                                          double *sum) {
    __shared__ double pm[4], ps[4];
    const size_t n = (size_t)nx * ny;
    const double inv_dx = 1.0 / dx;
    double m = 0.0, s = 0.0;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < n;
         idx += (size_t)gridDim.x * 256) {
        int i = (int)(idx % nx), j = (int)(idx / nx);
        double d;
        bool xin = i > 0 || px, yin = j > 0 || py;
        int im = i > 0 ? i - 1 : nx - 1, jm = j > 0 ? j - 1 : ny - 1;
        double dxb = bx[idx] - bx[(size_t)im + (size_t)nx * j];
        double dyb = by[idx] - by[(size_t)i + (size_t)nx * jm];
        if (xin && yin) d = (dxb + by[idx] - by[(size_t)i + (size_t)nx * jm]) * inv_dx;
        else if (xin) d = dxb * inv_dx;
        else if (yin) d = dyb * inv_dx;
        else d = 0.0;
        double a = fabs(d);
        m = fmax(m, a);
        s += a;
    }
    m = wave_max(m);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) {
        pm[threadIdx.x >> 6] = m;
        ps[threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMax(maxbits, encode_nonneg(fmax(fmax(pm[0], pm[1]), fmax(pm[2], pm[3]))));
        atomicAdd(sum, (ps[0] + ps[1]) + (ps[2] + ps[3]));
    }
}

int launch_stage(const StageLaunch &L, hipStream_t stream) {
    if (L.g.mask) {   // the diagnostics variant (stage 1, 32 x 8 tiles)
        const int tiles_x = (L.g.nx + 31) / 32, ntiles = tiles_x * ((L.g.ny + 7) / 8);
        hipLaunchKernelGGL(mask_fn(L.solver, L.limiter), dim3((ntiles + 7) / 8 * 8), dim3(256), 0,
                           stream, L.g, L.use_lim, L.mood, tiles_x, ntiles);
        return launch_check("stage_mask_kernel");
    }
    StageFn fn = stage_fn(L.solver, L.limiter, L.shape);
    CH_REQUIRE(fn, "tile shape %d is not available for riemann_solver %d", L.shape, L.solver);
    const TileDims d = RES_TILE_SHAPES[L.shape];
    hipLaunchKernelGGL(fn, dim3(L.grid), dim3(d.tx * d.ty), 0, stream, L.g, L.stage, L.use_lim,
                       L.mood, L.tiles_x, L.ntiles);
    return launch_check("stage_kernel");
}

int launch_cfl(const CflArgs &a, hipStream_t stream) {
    const int blocks = std::max(1, std::min(a.ny, 4096));   // one row per block and round
    hipLaunchKernelGGL(cfl_reduce_kernel, dim3(blocks), dim3(256), 0, stream, a);
    return launch_check("cfl_reduce_kernel");
}

int launch_dt_finalize(StepState *st, int mode, double dx, const double *dt_seq,
                       hipStream_t stream) {
    hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(1), 0, stream, st, mode, dx, dt_seq);
    return launch_check("dt_finalize_kernel");
}

int launch_set_state(StepState *st, double t, int oc, int step, hipStream_t stream) {
    hipLaunchKernelGGL(set_state_kernel, dim3(1), dim3(1), 0, stream, st, t, oc, step);
    return launch_check("set_state_kernel");
}

int launch_zero_u32(uint32_t *p, size_t n, hipStream_t stream) {
    int blocks = (int)std::min<size_t>((n + 255) / 256, 1024);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(blocks), dim3(256), 0, stream, p, n);
    return launch_check("zero_u32_kernel");
}

int launch_divb(const double *bx, const double *by, int nx, int ny, int px, int py, double dx,
                unsigned long long *maxbits, double *sum, hipStream_t stream) {
    size_t n = (size_t)nx * ny;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(divb_reduce_kernel, dim3(blocks), dim3(256), 0, stream, bx, by, nx, ny, px,
                       py, dx, maxbits, sum);
    return launch_check("divb_reduce_kernel");
}

}  // namespace chimera

using namespace chimera;

// ---------------------------------------------------------------------------------
// resident context

struct chimera_ctx {
    int nx = 0, ny = 0, px = 0, py = 0;
    double dx = 0, gamma = 0;
    chimera_scheme_t sch{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *base = nullptr;
    double *S[3][6] = {};
    int cur = 0;
    StepState *st = nullptr;
    double *dt_seq = nullptr;
    int dt_cap = 0;
    unsigned long long *scratch_u64 = nullptr;  // [0] cfl / divb max, [1] divb sum (as double)
    uint32_t *mask = nullptr;
    int shape = 0;
    int tiles_x = 0, tiles_y = 0, ntiles = 0, grid = 0;
    hipGraphExec_t gexec[2][3] = {};
    bool timing = false;
    std::vector<hipEvent_t> ev;  // event pool: pairs around fused stage launches
    size_t ev_used = 0;          // events recorded and not yet collected
    double stage_ms = 0.0;
    int stage_launches = 0;
};

// Next pair of timing events from the pool.
static int take_events(chimera_ctx *c, hipEvent_t *a, hipEvent_t *b) {
    while (c->ev.size() < c->ev_used + 2) {
        hipEvent_t e;
        CH_HIP(hipEventCreate(&e));
        c->ev.push_back(e);
    }
    *a = c->ev[c->ev_used];
    *b = c->ev[c->ev_used + 1];
    c->ev_used += 2;
    return CHIMERA_OK;
}

// Fold the recorded event pairs into the running average (synchronises).
static int collect_events(chimera_ctx *c) {
    if (c->ev_used == 0) return CHIMERA_OK;
    CH_HIP(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
        float ms = 0.f;
        CH_HIP(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
        c->stage_ms += ms;
        c->stage_launches += 1;
    }
    c->ev_used = 0;
    return CHIMERA_OK;
}

extern "C" const char *chimera_last_error(void) { return chimera::g_err; }

static void drop_graphs(chimera_ctx *c) {
    for (auto &row : c->gexec)
        for (auto &g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

static void set_shape(chimera_ctx *c, int shape) {
    c->shape = shape;
    const TileDims d = RES_TILE_SHAPES[shape];
    c->tiles_x = (c->nx + d.tx - 1) / d.tx;
    c->tiles_y = (c->ny + d.ty - 1) / d.ty;
    c->ntiles = c->tiles_x * c->tiles_y;
    c->grid = (c->ntiles + 7) / 8 * 8;
    drop_graphs(c);
}

extern "C" int chimera_device_count(int *count) {
    CH_REQUIRE(count, "count is NULL");
    CH_HIP(hipGetDeviceCount(count));
    return CHIMERA_OK;
}

static StageLaunch stage_launch(chimera_ctx *c, int stage, double *const *uin,
                                double *const *u0, double *const *uout, uint32_t *mask) {
    StageLaunch L{};
    for (int q = 0; q < 6; ++q) {
        L.g.uin[q] = uin[q];
        L.g.u0[q] = u0[q];
        L.g.uout[q] = uout[q];
    }
    L.g.dt = &c->st->dt;
    L.g.mask = mask;
    L.g.nx = c->nx;
    L.g.ny = c->ny;
    L.g.px = c->px;
    L.g.py = c->py;
    L.g.dx = c->dx;
    L.g.vol = c->dx * c->dx;
    L.g.gamma = c->gamma;
    L.stage = stage;
    L.use_lim = c->sch.use_slope_limiting;
    L.mood = c->sch.fallback_2_mood;
    L.solver = c->sch.riemann_solver;
    L.limiter = c->sch.slope_limiter;
    L.shape = c->shape;
    L.tiles_x = c->tiles_x;
    L.ntiles = c->ntiles;
    L.grid = c->grid;
    return L;
}

static CflArgs cfl_args(chimera_ctx *c, double *const *u, unsigned long long *bits) {
    CflArgs a{};
    for (int q = 0; q < 6; ++q) a.u[q] = u[q];
    a.nx = c->nx;
    a.ny = c->ny;
    a.px = c->px;
    a.py = c->py;
    a.vol = c->dx * c->dx;
    a.gamma = c->gamma;
    a.smax_bits = bits;
    return a;
}

static int enqueue_step(chimera_ctx *c, int mode, bool timed) {
    double *const *X = c->S[c->cur];
    double *const *Y = c->S[1 - c->cur];
    double *const *Z = c->S[2];
    if (mode != CHIMERA_DT_PRESCRIBED)
        CH_TRY(launch_cfl(cfl_args(c, X, &c->st->smax_bits), c->stream));
    CH_TRY(launch_dt_finalize(c->st, mode, c->dx, c->dt_seq, c->stream));
    double *const *io[3][3] = {{X, X, Y}, {Y, X, Z}, {Z, X, Y}};
    for (int k = 0; k < 3; ++k) {
        StageLaunch L = stage_launch(c, k + 1, io[k][0], io[k][1], io[k][2], nullptr);
        if (timed) {
            hipEvent_t a, b;
            CH_TRY(take_events(c, &a, &b));
            CH_HIP(hipEventRecord(a, c->stream));
            CH_TRY(launch_stage(L, c->stream));
            CH_HIP(hipEventRecord(b, c->stream));
        } else {
            CH_TRY(launch_stage(L, c->stream));
        }
    }
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_create(chimera_ctx_t **out, int nx, int ny, double dx, double gamma,
                                  const chimera_bc_t *bc, const chimera_scheme_t *scheme,
                                  void *stream) {
    CH_REQUIRE(out, "ctx out-pointer is NULL");
    CH_REQUIRE(nx >= 4 && ny >= 4, "grid must be at least 4 x 4 (got %d x %d)", nx, ny);
    CH_REQUIRE((long long)nx * ny <= MAX_GRID_CELLS, "grid of %d x %d cells exceeds %lld cells", nx, ny,
               MAX_GRID_CELLS);
    CH_REQUIRE(dx > 0.0 && gamma > 1.0, "dx must be > 0 and gamma > 1");
    int px, py;
    CH_TRY(check_bc(bc, &px, &py));
    CH_TRY(check_scheme(scheme));
    chimera_ctx *c = new chimera_ctx();
    c->nx = nx;
    c->ny = ny;
    c->px = px;
    c->py = py;
    c->dx = dx;
    c->gamma = gamma;
    c->sch = *scheme;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            set_error("hipStreamCreate failed");
            return CHIMERA_EHIP;
        }
        c->own_stream = true;
    }
    size_t n = (size_t)nx * ny;
    hipError_t e = hipMalloc(&c->base, 18 * n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->st, sizeof(StepState));
    if (e == hipSuccess) e = hipMalloc(&c->scratch_u64, 4 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->st, 0, sizeof(StepState));
    if (e != hipSuccess) {
        set_error("device allocation for a %d x %d grid failed: %s", nx, ny, hipGetErrorString(e));
        chimera_ctx_destroy(c);
        return CHIMERA_ENOMEM;
    }
    for (int s = 0; s < 3; ++s)
        for (int q = 0; q < 6; ++q) c->S[s][q] = c->base + (size_t)(6 * s + q) * n;
    // default tile: 32 x 16 (two 512-thread blocks per CU) once the grid has a tile per
    // CU and the solver leaves room for two such blocks (not HLLD); HLLD 16 x 16 (2.5-9%
    // faster than 32 x 8 on 512^2-2048^2 grids at three waves per SIMD); 32 x 8 otherwise
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 0;
    const long long tiles16 = (long long)((nx + 31) / 32) * ((ny + 15) / 16);
    const bool hlld = scheme->riemann_solver == HLLD;
    set_shape(c, (!hlld && cus > 0 && tiles16 >= cus) ? 4 : (hlld ? 1 : 0));
    int rc = launch_set_state(c->st, 0.0, 1, 0, c->stream);
    if (rc == CHIMERA_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = CHIMERA_EHIP;
    if (rc != CHIMERA_OK) {
        chimera_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_destroy(chimera_ctx_t *c) {
    if (!c) return CHIMERA_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (auto e : c->ev) (void)hipEventDestroy(e);
    (void)hipFree(c->base);
    (void)hipFree(c->st);
    (void)hipFree(c->dt_seq);
    (void)hipFree(c->scratch_u64);
    (void)hipFree(c->mask);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_upload(chimera_ctx_t *c, const double *const *state) {
    CH_REQUIRE(c && state, "NULL argument");
    size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
    for (int q = 0; q < 6; ++q) {
        CH_REQUIRE(state[q], "state[%d] is NULL", q);
        CH_HIP(hipMemcpyAsync(c->S[c->cur][q], state[q], bytes, hipMemcpyDefault, c->stream));
    }
    CH_HIP(hipStreamSynchronize(c->stream));
    return CHIMERA_OK;
}

// The initial state of problem type 1-4 written on the device (setup.hip), in place of an
// upload; the grid must be the problem's: N x N cells, dx = 1/N, its gamma and its
// periodicity (chimera_problem_info).
extern "C" int chimera_ctx_setup_problem(chimera_ctx_t *c, int problem_type) {
    CH_REQUIRE(c, "NULL argument");
    double dx = 0.0, gamma = 0.0;
    chimera_bc_t bc{};
    CH_TRY(problem_info(problem_type, c->nx, &dx, &gamma, &bc));
    int px = 0, py = 0;
    CH_TRY(check_bc(&bc, &px, &py));
    CH_REQUIRE(c->ny == c->nx && c->dx == dx && c->gamma == gamma && c->px == px && c->py == py,
               "problem_type %d needs an N x N grid with dx = 1/N, gamma = %.17g and %s "
               "boundaries (grid %d x %d, dx = %.17g, gamma = %.17g)",
               problem_type, gamma, px ? "periodic" : "outflow", c->nx, c->ny, c->dx, c->gamma);
    double *u[6];
    for (int q = 0; q < 6; ++q) u[q] = c->S[c->cur][q];
    return launch_setup_problem(problem_type, c->nx, u, c->stream);
}

extern "C" int chimera_ctx_download(chimera_ctx_t *c, double *const *state) {
    CH_REQUIRE(c && state, "NULL argument");
    size_t bytes = (size_t)c->nx * c->ny * sizeof(double);
    for (int q = 0; q < 6; ++q) {
        CH_REQUIRE(state[q], "state[%d] is NULL", q);
        CH_HIP(hipMemcpyAsync(state[q], c->S[c->cur][q], bytes, hipMemcpyDefault, c->stream));
    }
    CH_HIP(hipStreamSynchronize(c->stream));
    return CHIMERA_OK;
}

// Upload a prescribed dt sequence and restart the device step counter at 0.
static int prepare_prescribed(chimera_ctx *c, int nsteps, const double *dt_in) {
    CH_REQUIRE(dt_in, "CHIMERA_DT_PRESCRIBED needs dt_in");
    if (nsteps > c->dt_cap) {
        (void)hipFree(c->dt_seq);
        c->dt_seq = nullptr;
        c->dt_cap = 0;
        // graphs captured with the old pointer are stale
        for (auto &row : c->gexec)
            if (row[CHIMERA_DT_PRESCRIBED]) {
                (void)hipGraphExecDestroy(row[CHIMERA_DT_PRESCRIBED]);
                row[CHIMERA_DT_PRESCRIBED] = nullptr;
            }
        CH_HIP(hipMalloc(&c->dt_seq, (size_t)nsteps * sizeof(double)));
        c->dt_cap = nsteps;
    }
    CH_HIP(hipMemcpyAsync(c->dt_seq, dt_in, (size_t)nsteps * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    StepState h;
    CH_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(hipStreamSynchronize(c->stream));
    return launch_set_state(c->st, h.t, h.out_count, 0, c->stream);
}

// One full step (graph launch or plain launches); flips the current state buffer.
static int launch_one_step(chimera_ctx *c, int dt_mode, bool use_graph, bool timed) {
    if (use_graph) {
        hipGraphExec_t &ge = c->gexec[c->cur][dt_mode];
        if (!ge) {
            hipGraph_t graph;
            CH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_step(c, dt_mode, false);
            hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (rc != CHIMERA_OK) return rc;
            CH_HIP(e);
            CH_HIP(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
        }
        CH_HIP(hipGraphLaunch(ge, c->stream));
    } else {
        CH_TRY(enqueue_step(c, dt_mode, timed));
    }
    c->cur = 1 - c->cur;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_steps(chimera_ctx_t *c, int nsteps, int dt_mode, const double *dt_in,
                                 int use_graph) {
    CH_REQUIRE(c, "ctx is NULL");
    CH_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    CH_REQUIRE(dt_mode >= 0 && dt_mode <= 2, "dt_mode must be 0, 1 or 2");
    if (nsteps == 0) return CHIMERA_OK;
    if (dt_mode == CHIMERA_DT_PRESCRIBED) CH_TRY(prepare_prescribed(c, nsteps, dt_in));
    const bool timed = c->timing && !use_graph;
    if (timed && c->ev_used + 6 * (size_t)nsteps > c->ev.size()) {
        // grow the pool before the launches so event creation stays out of the loop
        CH_TRY(collect_events(c));
        while (c->ev.size() < 6 * (size_t)nsteps) {
            hipEvent_t e;
            CH_HIP(hipEventCreate(&e));
            c->ev.push_back(e);
        }
    }
    for (int s = 0; s < nsteps; ++s) CH_TRY(launch_one_step(c, dt_mode, use_graph != 0, timed));
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_sync(chimera_ctx_t *c) {
    CH_REQUIRE(c, "ctx is NULL");
    CH_HIP(hipStreamSynchronize(c->stream));
    return CHIMERA_OK;
}

extern "C" void *chimera_ctx_stream(chimera_ctx_t *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int chimera_ctx_get_time(chimera_ctx_t *c, double *t, int *out_count) {
    CH_REQUIRE(c, "ctx is NULL");
    StepState h;
    CH_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(hipStreamSynchronize(c->stream));
    if (t) *t = h.t;
    if (out_count) *out_count = h.out_count;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_set_time(chimera_ctx_t *c, double t, int out_count) {
    CH_REQUIRE(c, "ctx is NULL");
    CH_TRY(launch_set_state(c->st, t, out_count, 0, c->stream));
    CH_HIP(hipStreamSynchronize(c->stream));
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_last_dt(chimera_ctx_t *c, double *dt) {
    CH_REQUIRE(c && dt, "NULL argument");
    StepState h;
    CH_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(hipStreamSynchronize(c->stream));
    *dt = h.dt;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_cfl_dt(chimera_ctx_t *c, int dt_mode, double *dt) {
    CH_REQUIRE(c && dt, "NULL argument");
    CH_REQUIRE(dt_mode == CHIMERA_DT_REFERENCE || dt_mode == CHIMERA_DT_CFL,
               "dt_mode must be 0 or 1");
    CH_HIP(hipMemsetAsync(c->scratch_u64, 0, sizeof(unsigned long long), c->stream));
    CH_TRY(launch_cfl(cfl_args(c, c->S[c->cur], c->scratch_u64), c->stream));
    unsigned long long bits = 0;
    StepState h;
    CH_HIP(hipMemcpyAsync(&bits, c->scratch_u64, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(hipMemcpyAsync(&h, c->st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(hipStreamSynchronize(c->stream));
    double smax;
    std::memcpy(&smax, &bits, sizeof(smax));
    double d = COURANT * (c->dx / smax);
    if (dt_mode == CHIMERA_DT_REFERENCE) {
        double t_next = h.out_count * T_OUT;
        if (h.t + d > t_next) d = t_next - h.t;
    }
    *dt = d > (double)1e-12f ? d : (double)1e-12f;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_divb(chimera_ctx_t *c, double *max_abs, double *sum_abs) {
    CH_REQUIRE(c, "ctx is NULL");
    CH_HIP(hipMemsetAsync(c->scratch_u64, 0, 2 * sizeof(unsigned long long), c->stream));
    CH_TRY(launch_divb(c->S[c->cur][4], c->S[c->cur][5], c->nx, c->ny, c->px, c->py, c->dx,
                       c->scratch_u64, reinterpret_cast<double *>(c->scratch_u64 + 1), c->stream));
    unsigned long long h[2];
    CH_HIP(hipMemcpyAsync(h, c->scratch_u64, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CH_HIP(hipStreamSynchronize(c->stream));
    double m, s;
    std::memcpy(&m, &h[0], sizeof(m));
    std::memcpy(&s, &h[1], sizeof(s));
    if (max_abs) *max_abs = m;
    if (sum_abs) *sum_abs = s;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_stage1_mask(chimera_ctx_t *c, uint32_t *mask_host) {
    CH_REQUIRE(c && mask_host, "NULL argument");
    size_t n = (size_t)c->nx * c->ny;
    if (!c->mask) CH_HIP(hipMalloc(&c->mask, n * sizeof(uint32_t)));
    CH_TRY(launch_zero_u32(c->mask, n, c->stream));
    // the stage output goes to the scratch copy; the state is not advanced
    StageLaunch L = stage_launch(c, 1, c->S[c->cur], c->S[c->cur], c->S[2], c->mask);
    CH_TRY(launch_stage(L, c->stream));
    CH_HIP(hipMemcpyAsync(mask_host, c->mask, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          c->stream));
    CH_HIP(hipStreamSynchronize(c->stream));
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_set_timing(chimera_ctx_t *c, int enable) {
    CH_REQUIRE(c, "ctx is NULL");
    CH_TRY(collect_events(c));
    c->timing = enable != 0;
    c->stage_ms = 0.0;
    c->stage_launches = 0;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_stage_kernel_ms(chimera_ctx_t *c, double *ms, int *launches) {
    CH_REQUIRE(c, "ctx is NULL");
    CH_TRY(collect_events(c));
    if (ms) *ms = c->stage_launches ? c->stage_ms / c->stage_launches : 0.0;
    if (launches) *launches = c->stage_launches;
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_set_tile_shape(chimera_ctx_t *c, int shape) {
    CH_REQUIRE(c, "ctx is NULL");
    CH_REQUIRE(shape >= 0 && shape < RES_NUM_SHAPES, "tile shape must be 0..%d",
               RES_NUM_SHAPES - 1);
    CH_REQUIRE(stage_fn(c->sch.riemann_solver, c->sch.slope_limiter, shape),
               "tile shape %d is not available for riemann_solver %d", shape,
               c->sch.riemann_solver);
    CH_HIP(hipStreamSynchronize(c->stream));
    set_shape(c, shape);
    return CHIMERA_OK;
}

extern "C" int chimera_ctx_get_tile_shape(chimera_ctx_t *c, int *shape, int *tx, int *ty) {
    CH_REQUIRE(c, "ctx is NULL");
    if (shape) *shape = c->shape;
    if (tx) *tx = RES_TILE_SHAPES[c->shape].tx;
    if (ty) *ty = RES_TILE_SHAPES[c->shape].ty;
    return CHIMERA_OK;
}

// Time each tile shape on one stage of the current state (output to the scratch
// copy, the state is not advanced) and keep the fastest.  ms_out (optional) gets
// CHIMERA_CTX_NUM_SHAPES timings (-1 for a shape the solver does not have).
extern "C" int chimera_ctx_autotune(chimera_ctx_t *c, int reps, int *best, double *ms_out) {
    CH_REQUIRE(c, "ctx is NULL");
    if (reps < 1) reps = 3;
    hipEvent_t a, b;
    CH_HIP(hipEventCreate(&a));
    CH_HIP(hipEventCreate(&b));
    double best_ms = 1e300;
    int best_shape = c->shape;
    for (int sh = 0; sh < RES_NUM_SHAPES; ++sh) {
        if (!stage_fn(c->sch.riemann_solver, c->sch.slope_limiter, sh)) {
            if (ms_out) ms_out[sh] = -1.0;
            continue;
        }
        set_shape(c, sh);
        StageLaunch L = stage_launch(c, 2, c->S[c->cur], c->S[c->cur], c->S[2], nullptr);
        CH_TRY(launch_stage(L, c->stream));  // warm-up
        CH_HIP(hipEventRecord(a, c->stream));
        for (int r = 0; r < reps; ++r) CH_TRY(launch_stage(L, c->stream));
        CH_HIP(hipEventRecord(b, c->stream));
        CH_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        CH_HIP(hipEventElapsedTime(&ms, a, b));
        double per = ms / reps;
        if (ms_out) ms_out[sh] = per;
        if (per < best_ms) {
            best_ms = per;
            best_shape = sh;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    set_shape(c, best_shape);
    if (best) *best = best_shape;
    return CHIMERA_OK;
}
