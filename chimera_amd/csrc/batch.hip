// Batched ensemble of independent grids of uneven sizes (the GRF Monte-Carlo runs
// that scripts/run_mhd_MC.py:57-73 launches one after another).
//
// All members live in one HBM arena.  Every RK stage is ONE persistent launch over
// a tile table that spans all members: the table is sorted by member size
// (largest first) and split into 8 equal contiguous segments, one per XCD; block b
// (on XCD b % 8) walks its XCD's segment with stride gridDim/8, so the tiles of a
// member stay on one XCD's L2.  dt is per member: a batched CFL kernel max-reduces
// each member's signal speed with one atomic per tile, and a per-member finalize
// applies the reference dt policy (src/main.f90:74-83).  A whole step is capturable
// as a hipGraph.
#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "resident.hpp"

namespace chimera {

struct MemberDesc {
    double *S[3][6];  // three state copies x (Mass, Momx, Momy, Energy, b_x, b_y)
    StepState *st;
    int nx, ny, px, py;
    double dx, vol, gamma;
};

struct TileRef {
    int member, tx, ty;
};

__device__ __forceinline__ void xcd_segment(int ntiles, int &begin, int &end, int &slot,
                                            int &slots) {
    const int xcd = blockIdx.x & 7;
    slot = blockIdx.x >> 3;
    slots = gridDim.x >> 3;
    begin = (int)((long long)ntiles * xcd / 8);
    end = (int)((long long)ntiles * (xcd + 1) / 8);
}

// A 512-thread block (32 x 16) is two waves per SIMD, so two blocks per CU need <= 128
// VGPRs (Rusanov and HLLE need about that; HLLD needs more and keeps one block).  HLLD's
// 256-thread blocks: at most 168 VGPRs, three waves per SIMD as their LDS allows (the
// compiler's 177-179 left two; 12-15% faster).
template <int SOLVER, int LIMITER, int TX, int TY>
__global__ __launch_bounds__(TX * TY)
__attribute__((amdgpu_waves_per_eu((TX * TY == 512 && SOLVER != HLLD) ? 4
                                   : ((TX * TY == 256 && SOLVER == HLLD) ? 3 : 1))))
void batch_stage_kernel(
    const MemberDesc *__restrict__ mem, const TileRef *__restrict__ tiles, int ntiles, int stage,
    int cur, int use_lim, int mood) {
    __shared__ double lds[TileShape<TX, TY>::LDS_DOUBLES];
    int begin, end, slot, slots;
    xcd_segment(ntiles, begin, end, slot, slots);
    const int X = cur, Y = 1 - cur, Z = 2;
    const int in_i = stage == 1 ? X : (stage == 2 ? Y : Z);
    const int out_i = stage == 2 ? Z : Y;
    for (int t = begin + slot; t < end; t += slots) {
        const TileRef tr = tiles[t];
        const MemberDesc &m = mem[tr.member];
        GridView g;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            g.uin[q] = m.S[in_i][q];
            g.u0[q] = m.S[X][q];
            g.uout[q] = m.S[out_i][q];
        }
        g.dt = &m.st->dt;
        g.mask = nullptr;
        g.nx = m.nx;
        g.ny = m.ny;
        g.px = m.px;
        g.py = m.py;
        g.dx = m.dx;
        g.vol = m.vol;
        g.gamma = m.gamma;
        stage_tile<SOLVER, LIMITER, TX, TY>(g, stage, use_lim != 0, mood != 0, tr.tx, tr.ty, lds);
        __syncthreads();  // the next tile reuses LDS
    }
}

// Per-member max of cf + |v| over the tiles of the u^n state (src/main.f90:69-79).
template <int TX, int TY>
__global__ __launch_bounds__(TX * TY) void batch_cfl_kernel(const MemberDesc *__restrict__ mem,
                                                            const TileRef *__restrict__ tiles,
                                                            int ntiles, int cur) {
    constexpr int NW = TX * TY / 64;   // waves per block
    __shared__ double part[NW];
    int begin, end, slot, slots;
    xcd_segment(ntiles, begin, end, slot, slots);
    for (int t = begin + slot; t < end; t += slots) {
        const TileRef tr = tiles[t];
        const MemberDesc &m = mem[tr.member];
        const int i = tr.tx * TX + (int)(threadIdx.x % TX);
        const int j = tr.ty * TY + (int)(threadIdx.x / TX);
        double s = 0.0;
        if (i < m.nx && j < m.ny) {
            double *const *u = m.S[cur];
            const size_t k = (size_t)i + (size_t)m.nx * j;
            const int im = i > 0 ? i - 1 : (m.px ? m.nx - 1 : 0);
            const int jm = j > 0 ? j - 1 : (m.py ? m.ny - 1 : 0);
            double Bx = 0.5 * (u[4][k] + u[4][(size_t)im + (size_t)m.nx * j]);
            double By = 0.5 * (u[5][k] + u[5][(size_t)i + (size_t)m.nx * jm]);
            double rho, vx, vy, P;
            cons_to_prim(u[0][k], u[1][k], u[2][k], u[3][k], Bx, By, m.gamma, m.vol, rho, vx, vy, P);
            s = cfl_speed(rho, vx, vy, P, Bx, By, m.gamma);
        }
        for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off, 64));
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mx = part[0];
            for (int w = 1; w < NW; ++w) mx = fmax(mx, part[w]);
            atomicMax(&m.st->smax_bits, encode_nonneg(mx));
        }
        __syncthreads();
    }
}

__global__ void batch_dt_kernel(const MemberDesc *__restrict__ mem, int members, int mode) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= members) return;
    StepState *st = mem[m].st;
    double smax = __longlong_as_double((long long)st->smax_bits);
    st->smax_bits = 0ull;
    double dt = COURANT * (mem[m].dx / smax);
    if (mode == CHIMERA_DT_REFERENCE) {
        double t_next = st->out_count * T_OUT;
        if (st->t + dt > t_next) dt = t_next - st->t;
    }
    st->dt = fmax(dt, (double)1e-12f);
}

__global__ void batch_step_end_kernel(const MemberDesc *__restrict__ mem, int members) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= members) return;
    StepState *st = mem[m].st;
    double t = st->t + st->dt;
    st->t = t;
    if (t >= st->out_count * T_OUT) st->out_count = st->out_count + 1;
    st->step = st->step + 1;
}

__global__ void batch_init_kernel(StepState *st, int members) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= members) return;
    st[m].t = 0.0;
    st[m].dt = 0.0;
    st[m].smax_bits = 0ull;
    st[m].out_count = 1;
    st[m].step = 0;
}

using BStageFn = void (*)(const MemberDesc *, const TileRef *, int, int, int, int, int);
using BCflFn = void (*)(const MemberDesc *, const TileRef *, int, int);

// The single-grid driver's tile shapes (RES_TILE_SHAPES, same indices and the same
// availability: no 32 x 32 for HLLD); nullptr: shape not available for that solver.
#define CH_BSHAPES(S, L)                                                                   \
    {batch_stage_kernel<S, L, 32, 8>, batch_stage_kernel<S, L, 16, 16>,                    \
     batch_stage_kernel<S, L, 64, 4>, batch_stage_kernel<S, L, 32, 32>,                    \
     batch_stage_kernel<S, L, 32, 16>}
#define CH_BSHAPES_HLLD(S, L)                                                              \
    {batch_stage_kernel<S, L, 32, 8>, batch_stage_kernel<S, L, 16, 16>,                    \
     batch_stage_kernel<S, L, 64, 4>, nullptr, batch_stage_kernel<S, L, 32, 16>}

static BStageFn bstage_fn(int solver, int limiter, int shape) {
    static const BStageFn table[3][2][RES_NUM_SHAPES] = {
        {CH_BSHAPES(RUSANOV, LIM_VAN_LEER), CH_BSHAPES(RUSANOV, LIM_MC)},
        {CH_BSHAPES(HLLE, LIM_VAN_LEER), CH_BSHAPES(HLLE, LIM_MC)},
        {CH_BSHAPES_HLLD(HLLD, LIM_VAN_LEER), CH_BSHAPES_HLLD(HLLD, LIM_MC)}};
    return table[solver - 1][limiter - 1][shape];
}
#undef CH_BSHAPES
#undef CH_BSHAPES_HLLD

static BCflFn bcfl_fn(int shape) {
    static const BCflFn table[RES_NUM_SHAPES] = {
        batch_cfl_kernel<32, 8>, batch_cfl_kernel<16, 16>, batch_cfl_kernel<64, 4>,
        batch_cfl_kernel<32, 32>, batch_cfl_kernel<32, 16>};
    return table[shape];
}

static int shape_threads(int shape) {
    return RES_TILE_SHAPES[shape].tx * RES_TILE_SHAPES[shape].ty;
}

}  // namespace chimera

using namespace chimera;

struct chimera_batch {
    int members = 0;
    chimera_scheme_t sch{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<MemberDesc> hmem;
    MemberDesc *dmem = nullptr;
    StepState *st = nullptr;
    double *arena = nullptr;
    long long cells = 0;
    TileRef *dtiles[RES_NUM_SHAPES] = {};
    int ntiles[RES_NUM_SHAPES] = {};
    int shape = 0;
    int blocks_per_cu = 0;  // 0: occupancy-derived
    int grid = 0;
    int cu_count = 0;
    int cur = 0;
    hipGraphExec_t gexec[2][2] = {};
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t ev_used = 0;
    double stage_ms = 0.0;
    int stage_launches = 0;
};

static void batch_drop_graphs(chimera_batch *b) {
    for (auto &row : b->gexec)
        for (auto &g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

// Blocks of the persistent stage launch.  Default: 16 x the blocks a CU holds at once.
// The launch then has more blocks than fit: blocks start whenever a slot frees up, so
// the blocks sharing a CU run out of phase (one block's halo loads and barriers overlap
// the other's arithmetic) instead of marching through the phases of every tile in step,
// and each block walks fewer tiles, which shortens the launch's tail (measured on the
// 512-member GRF ensemble, 32 x 16 tiles: 4.65 ms per stage with 2 blocks per CU, 4.34
// with 16, 4.26 with 32 or 64).
constexpr int BATCH_OVERSUBSCRIBE = 16;

static int batch_set_grid(chimera_batch *b) {
    int bpc = b->blocks_per_cu;
    if (bpc <= 0) {
        int occ = 0;
        BStageFn fn = bstage_fn(b->sch.riemann_solver, b->sch.slope_limiter, b->shape);
        CH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fn,
                                                            shape_threads(b->shape), 0));
        bpc = (occ > 0 ? occ : 1) * BATCH_OVERSUBSCRIBE;
    }
    long long g = (long long)b->cu_count * bpc;
    const long long tiles8 = ((long long)b->ntiles[b->shape] + 7) / 8 * 8;   // no idle blocks
    g = std::min(g, std::max(8ll, tiles8));
    g = std::max(8ll, g / 8 * 8);
    b->grid = g;
    batch_drop_graphs(b);
    return CHIMERA_OK;
}

static int batch_tables(chimera_batch *b) {
    // members ordered largest first, tiles row-major within a member
    std::vector<int> order(b->members);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return (long long)b->hmem[a].nx * b->hmem[a].ny > (long long)b->hmem[c].nx * b->hmem[c].ny;
    });
    for (int sh = 0; sh < RES_NUM_SHAPES; ++sh) {
        const TileDims d = RES_TILE_SHAPES[sh];
        std::vector<TileRef> t;
        for (int m : order) {
            int tx = (b->hmem[m].nx + d.tx - 1) / d.tx, ty = (b->hmem[m].ny + d.ty - 1) / d.ty;
            for (int y = 0; y < ty; ++y)
                for (int x = 0; x < tx; ++x) t.push_back({m, x, y});
        }
        b->ntiles[sh] = (int)t.size();
        CH_HIP(hipMalloc(&b->dtiles[sh], t.size() * sizeof(TileRef)));
        CH_HIP(hipMemcpy(b->dtiles[sh], t.data(), t.size() * sizeof(TileRef), hipMemcpyHostToDevice));
    }
    return CHIMERA_OK;
}

static int batch_launch_stage(chimera_batch *b, int stage) {
    BStageFn fn = bstage_fn(b->sch.riemann_solver, b->sch.slope_limiter, b->shape);
    hipLaunchKernelGGL(fn, dim3(b->grid), dim3(shape_threads(b->shape)), 0, b->stream, b->dmem,
                       b->dtiles[b->shape], b->ntiles[b->shape], stage, b->cur,
                       b->sch.use_slope_limiting, b->sch.fallback_2_mood);
    return launch_check("batch_stage_kernel");
}

static int batch_enqueue_step(chimera_batch *b, int mode, bool timed) {
    const int mb = (b->members + 127) / 128;
    hipLaunchKernelGGL(bcfl_fn(b->shape), dim3(b->grid), dim3(shape_threads(b->shape)), 0,
                       b->stream,
                       b->dmem, b->dtiles[b->shape], b->ntiles[b->shape], b->cur);
    CH_TRY(launch_check("batch_cfl_kernel"));
    hipLaunchKernelGGL(batch_dt_kernel, dim3(mb), dim3(128), 0, b->stream, b->dmem, b->members, mode);
    CH_TRY(launch_check("batch_dt_kernel"));
    for (int k = 1; k <= 3; ++k) {
        if (timed) {
            while (b->ev.size() < b->ev_used + 2) {
                hipEvent_t e;
                CH_HIP(hipEventCreate(&e));
                b->ev.push_back(e);
            }
            CH_HIP(hipEventRecord(b->ev[b->ev_used], b->stream));
            CH_TRY(batch_launch_stage(b, k));
            CH_HIP(hipEventRecord(b->ev[b->ev_used + 1], b->stream));
            b->ev_used += 2;
        } else {
            CH_TRY(batch_launch_stage(b, k));
        }
    }
    hipLaunchKernelGGL(batch_step_end_kernel, dim3(mb), dim3(128), 0, b->stream, b->dmem, b->members);
    return launch_check("batch_step_end_kernel");
}

static int batch_collect(chimera_batch *b) {
    if (b->ev_used == 0) return CHIMERA_OK;
    CH_HIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i + 1 < b->ev_used; i += 2) {
        float ms = 0.f;
        CH_HIP(hipEventElapsedTime(&ms, b->ev[i], b->ev[i + 1]));
        b->stage_ms += ms;
        b->stage_launches += 1;
    }
    b->ev_used = 0;
    return CHIMERA_OK;
}

extern "C" int chimera_batch_destroy(chimera_batch_t *b) {
    if (!b) return CHIMERA_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (auto e : b->ev) (void)hipEventDestroy(e);
    for (auto &t : b->dtiles) (void)hipFree(t);
    (void)hipFree(b->dmem);
    (void)hipFree(b->st);
    (void)hipFree(b->arena);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CHIMERA_OK;
}

extern "C" int chimera_batch_create(chimera_batch_t **out, int members, const int *nx,
                                    const int *ny, const double *dx, const double *gamma,
                                    const chimera_bc_t *bc, const chimera_scheme_t *scheme,
                                    void *stream) {
    CH_REQUIRE(out && nx && ny && dx && gamma && bc, "NULL argument");
    CH_REQUIRE(members >= 1, "members must be >= 1");
    CH_TRY(check_scheme(scheme));
    std::vector<MemberDesc> hm(members);
    long long cells = 0;
    for (int m = 0; m < members; ++m) {
        CH_REQUIRE(nx[m] >= 4 && ny[m] >= 4, "member %d: grid must be at least 4 x 4", m);
        CH_REQUIRE((long long)nx[m] * ny[m] <= MAX_GRID_CELLS, "member %d: grid exceeds %lld cells", m,
                   MAX_GRID_CELLS);
        CH_REQUIRE(dx[m] > 0.0 && gamma[m] > 1.0, "member %d: dx must be > 0, gamma > 1", m);
        int px, py;
        CH_TRY(check_bc(&bc[m], &px, &py));
        hm[m].nx = nx[m];
        hm[m].ny = ny[m];
        hm[m].px = px;
        hm[m].py = py;
        hm[m].dx = dx[m];
        hm[m].vol = dx[m] * dx[m];
        hm[m].gamma = gamma[m];
        cells += (long long)nx[m] * ny[m];
    }
    chimera_batch *b = new chimera_batch();
    b->members = members;
    b->sch = *scheme;
    b->hmem = hm;
    b->cells = cells;
    auto fail = [&](int rc) {
        chimera_batch_destroy(b);
        return rc;
    };
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            set_error("hipStreamCreate failed");
            return fail(CHIMERA_EHIP);
        }
        b->own_stream = true;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        set_error("cannot query the device");
        return fail(CHIMERA_EHIP);
    }
    hipError_t e = hipMalloc(&b->arena, (size_t)cells * 18 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&b->st, members * sizeof(StepState));
    if (e == hipSuccess) e = hipMalloc(&b->dmem, members * sizeof(MemberDesc));
    if (e != hipSuccess) {
        set_error("device allocation for %lld cells failed: %s", cells, hipGetErrorString(e));
        return fail(CHIMERA_ENOMEM);
    }
    size_t off = 0;
    for (int m = 0; m < members; ++m) {
        size_t n = (size_t)b->hmem[m].nx * b->hmem[m].ny;
        for (int s = 0; s < 3; ++s)
            for (int q = 0; q < 6; ++q) b->hmem[m].S[s][q] = b->arena + off + (size_t)(6 * s + q) * n;
        b->hmem[m].st = b->st + m;
        off += 18 * n;
    }
    if (hipMemcpy(b->dmem, b->hmem.data(), members * sizeof(MemberDesc), hipMemcpyHostToDevice) !=
        hipSuccess) {
        set_error("descriptor upload failed");
        return fail(CHIMERA_EHIP);
    }
    int rc = batch_tables(b);
    // default tile (as a single grid's): 32 x 16 once the ensemble has a tile per CU and
    // the solver leaves room for two such blocks per CU (not HLLD), HLLD 16 x 16, else
    // 32 x 8; chimera_batch_set_tiling / chimera_batch_autotune change it
    const bool hlld = b->sch.riemann_solver == HLLD;
    b->shape = (!hlld && b->ntiles[4] >= b->cu_count) ? 4 : (hlld ? 1 : 0);
    if (rc == CHIMERA_OK) rc = batch_set_grid(b);
    if (rc != CHIMERA_OK) return fail(rc);
    hipLaunchKernelGGL(batch_init_kernel, dim3((members + 127) / 128), dim3(128), 0, b->stream, b->st,
                       members);
    if (launch_check("batch_init_kernel") != CHIMERA_OK ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return fail(CHIMERA_EHIP);
    *out = b;
    return CHIMERA_OK;
}

extern "C" int chimera_batch_upload(chimera_batch_t *b, int member, const double *const *state) {
    CH_REQUIRE(b && state, "NULL argument");
    CH_REQUIRE(member >= 0 && member < b->members, "member %d out of range", member);
    const MemberDesc &m = b->hmem[member];
    size_t bytes = (size_t)m.nx * m.ny * sizeof(double);
    for (int q = 0; q < 6; ++q) {
        CH_REQUIRE(state[q], "state[%d] is NULL", q);
        CH_HIP(hipMemcpyAsync(m.S[b->cur][q], state[q], bytes, hipMemcpyDefault, b->stream));
    }
    CH_HIP(hipStreamSynchronize(b->stream));
    return CHIMERA_OK;
}

extern "C" int chimera_batch_download(chimera_batch_t *b, int member, double *const *state) {
    CH_REQUIRE(b && state, "NULL argument");
    CH_REQUIRE(member >= 0 && member < b->members, "member %d out of range", member);
    const MemberDesc &m = b->hmem[member];
    size_t bytes = (size_t)m.nx * m.ny * sizeof(double);
    for (int q = 0; q < 6; ++q) {
        CH_REQUIRE(state[q], "state[%d] is NULL", q);
        CH_HIP(hipMemcpyAsync(state[q], m.S[b->cur][q], bytes, hipMemcpyDefault, b->stream));
    }
    CH_HIP(hipStreamSynchronize(b->stream));
    return CHIMERA_OK;
}

extern "C" int chimera_batch_steps(chimera_batch_t *b, int nsteps, int dt_mode, int use_graph) {
    CH_REQUIRE(b, "batch is NULL");
    CH_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    CH_REQUIRE(dt_mode == CHIMERA_DT_REFERENCE || dt_mode == CHIMERA_DT_CFL,
               "batched dt_mode must be 0 or 1");
    const bool timed = b->timing && !use_graph;
    for (int s = 0; s < nsteps; ++s) {
        if (use_graph) {
            hipGraphExec_t &ge = b->gexec[b->cur][dt_mode];
            if (!ge) {
                hipGraph_t graph;
                CH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                int rc = batch_enqueue_step(b, dt_mode, false);
                hipError_t e = hipStreamEndCapture(b->stream, &graph);
                if (rc != CHIMERA_OK) return rc;
                CH_HIP(e);
                CH_HIP(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
                (void)hipGraphDestroy(graph);
            }
            CH_HIP(hipGraphLaunch(ge, b->stream));
        } else {
            CH_TRY(batch_enqueue_step(b, dt_mode, timed));
        }
        b->cur = 1 - b->cur;
    }
    return CHIMERA_OK;
}

extern "C" int chimera_batch_member_time(chimera_batch_t *b, int member, double *t, int *out_count,
                                         double *last_dt) {
    CH_REQUIRE(b, "batch is NULL");
    CH_REQUIRE(member >= 0 && member < b->members, "member %d out of range", member);
    StepState h;
    CH_HIP(hipMemcpyAsync(&h, b->st + member, sizeof(h), hipMemcpyDeviceToHost, b->stream));
    CH_HIP(hipStreamSynchronize(b->stream));
    if (t) *t = h.t;
    if (out_count) *out_count = h.out_count;
    if (last_dt) *last_dt = h.dt;
    return CHIMERA_OK;
}

extern "C" int chimera_batch_sync(chimera_batch_t *b) {
    CH_REQUIRE(b, "batch is NULL");
    CH_HIP(hipStreamSynchronize(b->stream));
    return CHIMERA_OK;
}

extern "C" void *chimera_batch_stream(chimera_batch_t *b) { return b ? (void *)b->stream : nullptr; }

extern "C" int chimera_batch_cells(chimera_batch_t *b, long long *cells) {
    CH_REQUIRE(b && cells, "NULL argument");
    *cells = b->cells;
    return CHIMERA_OK;
}

extern "C" int chimera_batch_set_tiling(chimera_batch_t *b, int shape, int blocks_per_cu) {
    CH_REQUIRE(b, "batch is NULL");
    CH_REQUIRE(shape >= 0 && shape < RES_NUM_SHAPES, "tile shape must be 0..%d",
               RES_NUM_SHAPES - 1);
    CH_REQUIRE(bstage_fn(b->sch.riemann_solver, b->sch.slope_limiter, shape),
               "tile shape %d is not available for riemann_solver %d", shape,
               b->sch.riemann_solver);
    CH_REQUIRE(blocks_per_cu >= 0 && blocks_per_cu <= 256, "blocks_per_cu must be 0..256");
    CH_HIP(hipStreamSynchronize(b->stream));
    b->shape = shape;
    b->blocks_per_cu = blocks_per_cu;
    return batch_set_grid(b);
}

extern "C" int chimera_batch_get_tiling(chimera_batch_t *b, int *shape, int *blocks_per_cu) {
    CH_REQUIRE(b, "batch is NULL");
    if (shape) *shape = b->shape;
    if (blocks_per_cu) *blocks_per_cu = b->grid / std::max(1, b->cu_count);
    return CHIMERA_OK;
}

// Time one stage (the stage-1 form: reads u^n, writes the copy that stage 1 of the
// next step overwrites anyway; the members' states are not advanced) for every tile
// shape and blocks-per-CU count (1 to the occupancy limit, then 2 to 32 times it); keep
// the fastest.
extern "C" int chimera_batch_autotune(chimera_batch_t *b, int reps, int *best_shape,
                                      int *best_bpc) {
    CH_REQUIRE(b, "batch is NULL");
    if (reps < 1) reps = 3;
    hipEvent_t e0, e1;
    CH_HIP(hipEventCreate(&e0));
    CH_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int bs = b->shape, bb = b->blocks_per_cu;
    const int saved_cur = b->cur;
    for (int sh = 0; sh < RES_NUM_SHAPES; ++sh) {
        int occ = 0;
        BStageFn fn = bstage_fn(b->sch.riemann_solver, b->sch.slope_limiter, sh);
        if (!fn) continue;
        CH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)fn,
                                                            shape_threads(sh), 0));
        // 1 .. occ resident blocks per CU, then oversubscribed grids of 2, 4, .. 32 x occ
        std::vector<int> cands;
        for (int bpc = 1; bpc <= std::max(1, occ); ++bpc) cands.push_back(bpc);
        for (int f = 2; f <= 32; f *= 2) cands.push_back(std::max(1, occ) * f);
        for (int bpc : cands) {
            b->shape = sh;
            b->blocks_per_cu = bpc;
            CH_TRY(batch_set_grid(b));
            CH_TRY(batch_launch_stage(b, 1));
            CH_HIP(hipEventRecord(e0, b->stream));
            for (int r = 0; r < reps; ++r) CH_TRY(batch_launch_stage(b, 1));
            CH_HIP(hipEventRecord(e1, b->stream));
            CH_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CH_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms / reps < best) {
                best = ms / reps;
                bs = sh;
                bb = bpc;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->cur = saved_cur;
    b->shape = bs;
    b->blocks_per_cu = bb;
    CH_TRY(batch_set_grid(b));
    if (best_shape) *best_shape = bs;
    if (best_bpc) *best_bpc = bb;
    return CHIMERA_OK;
}

extern "C" int chimera_batch_set_timing(chimera_batch_t *b, int enable) {
    CH_REQUIRE(b, "batch is NULL");
    CH_TRY(batch_collect(b));
    b->timing = enable != 0;
    b->stage_ms = 0.0;
    b->stage_launches = 0;
    return CHIMERA_OK;
}

extern "C" int chimera_batch_stage_kernel_ms(chimera_batch_t *b, double *ms, int *launches) {
    CH_REQUIRE(b, "batch is NULL");
    CH_TRY(batch_collect(b));
    if (ms) *ms = b->stage_launches ? b->stage_ms / b->stage_launches : 0.0;
    if (launches) *launches = b->stage_launches;
    return CHIMERA_OK;
}
